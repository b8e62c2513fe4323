// Host side of the exodus commit, free of HIP so that a CPU build (the
// thread-sanitizer harness under tests/emu) runs the same code:
//
// - Voids: the catalogue as accepted, with the reference's 32^3 void cell
//   list (cell_linked_list.c:41-59, 181-232) and its exact overlap scan
//   (sif__overlap_exact, src/finder/utils.c:850-960), the conflict search
//   and compaction of SIF_FINDER_DISPLACE (exodus_finder.c:684-1032);
// - host_paint: sif__mark_sphere / sif__clear_sphere (utils.c:965-1039) on
//   a host mask;
// - WorkPool: the fork-join pool the per-candidate tests run on.
//
// Threading contract: link() and every mutating call run on one thread;
// overlap_linked() and the other const reads may run on the pool's threads
// at once, after link() has run.
#pragma once
#include <immintrin.h>

#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "overlap.h"
#include "paint.h"

namespace sif_amd {

// The accepted voids with their 32^3 cell list (cell_linked_list.c:41-59,
// :181-210; head push, periodic cell index).
struct Voids {
  std::vector<float> cx, cy, cz, r;
  std::vector<uint64_t> n_in;
  std::vector<int32_t> head, next;
  size_t linked = 0;  // voids [0, linked) are in the cell list
  float inv_l = 0.0f, box = 0.0f, max_r = 0.0f;
  bool dirty = true;
  // reads of the cell list made while voids were left unlinked (none when
  // the callers keep to link() before the pool's scans)
  mutable std::atomic<uint64_t> unlinked_reads{0};
  void init(float box_) {
    box = box_;
    inv_l = (float)VOID_CELLS / box;
    head.assign(VOID_CELLS * VOID_CELLS * VOID_CELLS, -1);
  }
  uint32_t cell_of(float x, float y, float z) const {
    return (void_axis_cell(x, inv_l) * VOID_CELLS + void_axis_cell(y, inv_l)) * VOID_CELLS +
           void_axis_cell(z, inv_l);
  }
  // The cell list is only read by the reference's scan, which the commit
  // needs rarely: voids are linked into it when it is next read.
  void add(float x, float y, float z, float rr, uint64_t nin) {
    cx.push_back(x);
    cy.push_back(y);
    cz.push_back(z);
    r.push_back(rr);
    n_in.push_back(nin);
    if (rr > max_r)
      max_r = rr;
    dirty = true;
  }
  void link() {
    next.resize(cx.size());
    for (; linked < cx.size(); linked++) {
      const uint32_t c = cell_of(cx[linked], cy[linked], cz[linked]);
      next[linked] = head[c];
      head[c] = (int32_t)linked;
    }
  }
  // utils.c:850-960
  bool overlap_exact(float x, float y, float z, float rr, float f, float scale) {
    link();
    return overlap_linked(x, y, z, rr, f, scale);
  }
  // the same, with every void already linked: reads only, so callable from
  // several threads at once once link() has run on one. A read with voids
  // left unlinked would miss them; it is counted (unlinked_reads), and the
  // run that made it fails rather than return a catalogue built on it.
  bool overlap_linked(float x, float y, float z, float rr, float f, float scale) const {
    if (linked != cx.size())
      unlinked_reads.fetch_add(1, std::memory_order_relaxed);
    const float half = box * 0.5f;
    const float sr = scale * (rr + max_r - f * (rr < max_r ? rr : max_r));
    const int32_t lo[3] = {(int32_t)floorf((x - sr) * inv_l), (int32_t)floorf((y - sr) * inv_l),
      (int32_t)floorf((z - sr) * inv_l)};
    const int32_t hi[3] = {(int32_t)floorf((x + sr) * inv_l), (int32_t)floorf((y + sr) * inv_l),
      (int32_t)floorf((z + sr) * inv_l)};
    const bool interior = (x - sr >= 0.0f) && (x + sr < box) && (y - sr >= 0.0f) &&
                          (y + sr < box) && (z - sr >= 0.0f) && (z + sr < box);
    for (int32_t a = lo[0]; a <= hi[0]; a++) {
      const uint32_t wa = void_wrap(a);
      for (int32_t b = lo[1]; b <= hi[1]; b++) {
        const uint32_t wb = void_wrap(b);
        for (int32_t c = lo[2]; c <= hi[2]; c++) {
          const uint32_t cell = (wa * VOID_CELLS + wb) * VOID_CELLS + void_wrap(c);
          for (int32_t i = head[cell]; i != -1; i = next[i]) {
            const float ri = r[i];
            const float mr = rr < ri ? rr : ri;
            const float tot = scale * (rr + ri - f * mr);
            const float t2 = tot * tot;
            if (interior) {
              const float dx = x - cx[i], dy = y - cy[i], dz = z - cz[i];
              if (dx * dx + dy * dy + dz * dz < t2)
                return true;
            } else {
              float dx = fabsf(x - cx[i]);
              dx = dx > half ? box - dx : dx;
              const float dx2 = dx * dx;
              if (dx2 >= t2)
                continue;
              float dy = fabsf(y - cy[i]);
              dy = dy > half ? box - dy : dy;
              const float dy2 = dy * dy;
              if (dx2 + dy2 >= t2)
                continue;
              float dz = fabsf(z - cz[i]);
              dz = dz > half ? box - dz : dz;
              if (dx2 + dy2 + dz * dz < t2)
                return true;
            }
          }
        }
      }
    }
    return false;
  }

  // --- SIF_FINDER_DISPLACE (exodus_finder.c:684-1032) ---

  // The cell window within `reach` of a point, whole along an axis it spans
  // (exodus_finder.c:706-717).
  void window(float x, float y, float z, float reach, int32_t lo[3], int32_t hi[3]) const {
    const float p[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
      lo[a] = (int32_t)floorf((p[a] - reach) * inv_l);
      hi[a] = (int32_t)floorf((p[a] + reach) * inv_l);
      if (hi[a] - lo[a] >= (int32_t)VOID_CELLS) {
        lo[a] = 0;
        hi[a] = VOID_CELLS - 1;
      }
    }
  }
  // Visits the listed voids of the window in the reference's order: cells in
  // window order, each chain from its head.
  template <class F>
  void visit(float x, float y, float z, float reach, F&& fn) {
    link();
    int32_t lo[3], hi[3];
    window(x, y, z, reach, lo, hi);
    for (int32_t a = lo[0]; a <= hi[0]; a++)
      for (int32_t b = lo[1]; b <= hi[1]; b++)
        for (int32_t c = lo[2]; c <= hi[2]; c++) {
          const uint32_t cell = (void_wrap(a) * VOID_CELLS + void_wrap(b)) * VOID_CELLS +
                                void_wrap(c);
          for (int32_t i = head[cell]; i != -1; i = next[i])
            fn((uint32_t)i);
        }
  }
  // folded axis distances to void i
  void dist(float x, float y, float z, uint32_t i, float* dx, float* dy, float* dz) const {
    const float half = 0.5f * box;
    float ax = fabsf(x - cx[i]), ay = fabsf(y - cy[i]), az = fabsf(z - cz[i]);
    *dx = ax > half ? box - ax : ax;
    *dy = ay > half ? box - ay : ay;
    *dz = az > half ? box - az : az;
  }
  // find_conflicts (:696-750): the voids a sphere collides with, in visiting
  // order, and the largest radius among them (-1 for none)
  float conflicts(float x, float y, float z, float rr, float f, float scale,
    std::vector<uint32_t>* out) {
    out->clear();
    float largest = -1.0f;
    const float reach = scale * (rr + max_r - f * (rr < max_r ? rr : max_r));
    visit(x, y, z, reach, [&](uint32_t i) {
      float dx, dy, dz;
      dist(x, y, z, i, &dx, &dy, &dz);
      const float ri = r[i];
      const float tot = scale * (rr + ri - f * (rr < ri ? rr : ri));
      if (dx * dx + dy * dy + dz * dz >= tot * tot)
        return;
      out->push_back(i);
      if (ri > largest)
        largest = ri;
    });
    return largest;
  }
  // cell_linked_list.c:212-232
  void unlink(uint32_t i) {
    link();
    int32_t* at = &head[cell_of(cx[i], cy[i], cz[i])];
    while (*at != -1 && *at != (int32_t)i)
      at = &next[*at];
    if (*at != -1)
      *at = next[i];
  }
  // compact_catalogue (:1004-1032): the displaced voids (radius negated)
  // dropped, the rest re-listed in catalogue order. The largest radius is
  // kept: the reference never lowers it.
  void compact() {
    size_t j = 0;
    for (size_t i = 0; i < cx.size(); i++) {
      if (!(r[i] > 0.0f))
        continue;
      cx[j] = cx[i];
      cy[j] = cy[i];
      cz[j] = cz[i];
      r[j] = r[i];
      n_in[j] = n_in[i];
      j++;
    }
    cx.resize(j);
    cy.resize(j);
    cz.resize(j);
    r.resize(j);
    n_in.resize(j);
    head.assign(VOID_CELLS * VOID_CELLS * VOID_CELLS, -1);
    linked = 0;
    link();
  }
};

// sif__mark_sphere / sif__clear_sphere (utils.c:965-1039) on a host mask: the
// reference's walk of the cube of nodes around the centre, with paint.h's
// float steps.
void host_paint(std::vector<uint64_t>& mask, float x, float y, float z, float r, float l,
  uint32_t n, bool set) {
  PaintSphere P;
  paint_sphere(P, x, y, z, r, l, n);
  for (int32_t gx = P.ox - P.c; gx <= P.ox + P.c; gx++) {
    const float ddx = paint_dist(P, gx, P.cx);
    const float dx2 = ddx * ddx;
    if (dx2 > P.r2)
      continue;
    const uint64_t wx = paint_wrap(gx, P.n);
    for (int32_t gy = P.oy - P.c; gy <= P.oy + P.c; gy++) {
      const float ddy = paint_dist(P, gy, P.cy);
      const float dxy2 = dx2 + ddy * ddy;
      if (dxy2 > P.r2)
        continue;
      const uint64_t wy = paint_wrap(gy, P.n);
      for (int32_t gz = P.oz - P.c; gz <= P.oz + P.c; gz++) {
        const float ddz = paint_dist(P, gz, P.cz);
        if (dxy2 + ddz * ddz <= P.r2) {
          const uint64_t cell = (wx * n + wy) * n + paint_wrap(gz, P.n);
          if (set)
            mask[cell >> 6] |= 1ull << (cell & 63);
          else
            mask[cell >> 6] &= ~(1ull << (cell & 63));
        }
      }
    }
  }
}

// A fork-join pool for the host's independent per-candidate tests: run(n,
// fn) calls fn(begin, end) over [0, n) in chunks from every worker and the
// caller, and returns when all are done.
class WorkPool {
 public:
  explicit WorkPool(unsigned workers) {
    for (unsigned i = 0; i < workers; i++)
      threads_.emplace_back([this]() { loop(); });
  }
  ~WorkPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_.store(true);
    }
    cv_.notify_all();
    for (std::thread& t : threads_)
      t.join();
  }
  // A run ends when its chunks are done, not when every worker has come by:
  // a worker still waking from sleep when the caller has taken the last
  // chunk never joins, and the caller waits only for the workers inside.
  void run(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn) {
    if (threads_.empty() || n <= GRAIN) {
      fn(0, n);
      return;
    }
    job_ = &fn;
    n_ = n;
    next_.store(0, std::memory_order_relaxed);
    const uint64_t g = gen_.load(std::memory_order_relaxed) + 1;
    open_.store(g, std::memory_order_seq_cst);
    {
      // under the lock: a worker about to sleep sees the new job
      std::lock_guard<std::mutex> lk(mu_);
      gen_.store(g, std::memory_order_release);
    }
    if (sleeping_.load() > 0)
      cv_.notify_all();
    work(fn, n);
    // closed: a worker that arrives now sees it and leaves without the job;
    // one that arrived before is inside and is waited for (each side writes
    // its flag before reading the other's, both sequentially consistent)
    open_.store(0, std::memory_order_seq_cst);
    while (inside_.load(std::memory_order_seq_cst) != 0)
      _mm_pause();
    job_ = nullptr;
  }

 private:
  static constexpr uint64_t GRAIN = 64;
  // Pauses a waiting worker spins before it sleeps (some tens of
  // microseconds): the commit's pool runs come a few tens of microseconds
  // apart, and a wake-up from sleep costs about as much.
  static constexpr uint32_t SPIN = 2048;
  void work(const std::function<void(uint64_t, uint64_t)>& fn, uint64_t n) {
    for (;;) {
      const uint64_t b = next_.fetch_add(GRAIN);
      if (b >= n)
        return;
      fn(b, b + GRAIN < n ? b + GRAIN : n);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      uint64_t g;
      for (uint32_t spins = 0;
           (g = gen_.load(std::memory_order_acquire)) == seen && !stop_.load(); spins++) {
        if (spins < SPIN) {
          _mm_pause();
          continue;
        }
        std::unique_lock<std::mutex> lk(mu_);
        sleeping_++;
        cv_.wait(lk, [&]() { return stop_.load() || gen_.load() != seen; });
        sleeping_--;
        spins = 0;
      }
      if (stop_.load())
        return;
      seen = g;
      inside_.fetch_add(1, std::memory_order_seq_cst);
      if (open_.load(std::memory_order_seq_cst) == g)
        work(*job_, n_);
      inside_.fetch_sub(1, std::memory_order_seq_cst);
    }
  }
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_;
  const std::function<void(uint64_t, uint64_t)>* job_ = nullptr;
  uint64_t n_ = 0;
  std::atomic<uint64_t> next_{0};
  std::atomic<uint64_t> open_{0};  // the generation taking workers, 0 when closed
  std::atomic<unsigned> inside_{0};
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> sleeping_{0};
  std::atomic<bool> stop_{false};
};

}  // namespace sif_amd
