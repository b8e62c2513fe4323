// The exodus driver on device-resident data.
//
// Replaces sif__exodus_run (src/finder/exodus_finder.c:1585-2492) for a
// periodic box with SIF_FINDER_SEEDS_CELLS on their nodes, and the helpers it
// calls on the commit path: accept_void (656-683), sif__overlap_quick and
// sif__overlap_exact (src/finder/utils.c:819-960), sif__mark_sphere
// (utils.c:965-1033) and the void cell list (cell_linked_list.c:181-210).
//
// Per rung, on the GPU: the filtered c2r (top-hat at the rung, CIC window
// divided out), the candidate scan and sort, and a prefilter of every
// candidate against the catalogue as the rung begins (exodus_finder.c:
// 871-944). The survivors come to the host in order. They are then evaluated
// in speculative batches -- a batch's sphere growths run on the GPU against the
// catalogue as the batch begins -- and committed one by one on the host, where
// each candidate is tested again against the catalogue as it now stands, at
// the rung radius and at its grown radius, exactly the tests of
// exodus_finder.c:2181-2329. The catalogue is therefore that of a sequential
// pass, whatever the batch sizes.
//
// The crossing witness (exodus_finder.c:2021-2035) is not run: it only refuses
// candidates that the commit would refuse, so leaving it out costs GPU work
// on those candidates and changes nothing in the catalogue.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <immintrin.h>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "common.h"
#include "internal.h"
#include "overlap.h"
#include "paint.h"
#include "rung_voids.h"
#include "batch_pairs.h"
#include "host_commit.h"
#include "sif/structures/catalogue.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {

// Growth batches: the reference's adaptive rule (exodus_finder.c:2331-2349)
// with a floor that keeps the GPU full while the batch's late, large spheres
// finish (about 2048 wavefronts resident; a floor of 8192 grows a 512^3
// step's candidates in fewer, fuller launches than 4096 did). The window
// only groups candidates, so any floor and ceiling give the same
// catalogue; sif_amd_set_batch_window() lowers them (BATCH_ROOM is what the
// slots hold).
constexpr uint32_t BATCH_FLOOR_DEFAULT = 8192;
constexpr uint32_t BATCH_ROOM = 32768;
uint32_t g_batch_floor = BATCH_FLOOR_DEFAULT;
uint32_t g_batch_max = BATCH_ROOM;
// Smoothings that only feed the scan list their candidate cells
// (sif_amd_set_cell_lists).
bool g_cell_lists = true;

// A scan of the void cell list made while voids were left unlinked would have
// missed them: the run fails rather than return a catalogue built on it.
int all_reads_linked(const Voids& V) {
  if (V.unlinked_reads.load() == 0)
    return SIF_OK;
  error_set(SIF_ERR_INVALID, "exodus: internal: the void cell list was read before linking");
  return SIF_ERR_INVALID;
}

struct DeviceState {
  DeviceBuffer spec, work, mask, keys, flat, keep, surv, skeys, luts, axis;
  DeviceBuffer pf_counts;  // per rung: masked, proxy, overlap
  DeviceBuffer why, since;  // the prefilter's verdicts; the candidates masked since (4)
  uint64_t* h_since = nullptr;  // pinned copy of `since`
  DeviceBuffer field2;     // the second smoothed field (rungs alternate)
  // or, when only scanned, the two rungs' cell lists in their place
  DeviceBuffer lkeys[2], llocal[2], lcount[2];
  DeviceBuffer rspec, rfield;  // a survey's randoms: spectrum, smoothed grid
};
thread_local DeviceState g_dev;

// A displacement rung's growth window on the device.
struct DisplaceBuffers {
  DeviceBuffer flat, cx, cy, cz, rad, reason, nin;
  int reserve(uint32_t n) {
    int st = flat.reserve((size_t)n * 4);
    if (st == SIF_OK) st = cx.reserve((size_t)n * 4);
    if (st == SIF_OK) st = cy.reserve((size_t)n * 4);
    if (st == SIF_OK) st = cz.reserve((size_t)n * 4);
    if (st == SIF_OK) st = rad.reserve((size_t)n * 4);
    if (st == SIF_OK) st = reason.reserve((size_t)n * 4);
    if (st == SIF_OK) st = nin.reserve((size_t)n * 8);
    return st;
  }
};
thread_local DisplaceBuffers disp_bufs;

// The pool's workers beside the calling thread: SIF_AMD_HOST_THREADS - 1 when
// set (1 to 64 threads in all), else one fewer than the hardware threads, at
// most 15.
WorkPool& host_pool() {
  // a thread on streams of its own is one of several runs side by side:
  // its per-candidate tests run inline
  if (on_own_streams()) {
    static thread_local WorkPool inline_pool(0);
    return inline_pool;
  }
  static WorkPool pool([]() {
    const char* env = getenv("SIF_AMD_HOST_THREADS");
    const long want = env ? strtol(env, nullptr, 10) : 0;
    if (want >= 1 && want <= 64)
      return (unsigned)(want - 1);
    return std::min(std::max(std::thread::hardware_concurrency(), 2u) - 1u, 15u);
  }());
  return pool;
}

// One growth batch in flight: device buffers, pinned host copies of its
// inputs and results, and the event that marks the results in host memory.
struct BatchSlot {
  DeviceBuffer d_flat, d_cx, d_cy, d_cz;
  // the batch's pair effects' scratch (batch_pairs.hip)
  DeviceBuffer d_cnt32, d_off, d_setmax, d_cells;
  // The launch's results in one block, copied to host memory at once:
  // radius, reason, state, effect count, enclosed count, algorithmic bytes,
  // then the effects list. The typed pointers into it are set per launch
  // (regions sized by its count), on the device and in the pinned copy.
  DeviceBuffer d_out;
  uint8_t* h_out = nullptr;
  float* p_rad = nullptr;
  uint32_t* p_reason = nullptr;
  uint32_t* p_state = nullptr;
  uint8_t* p_cnt = nullptr;
  uint64_t* p_nin = nullptr;
  uint64_t* p_algo = nullptr;
  PairEffect* p_eff = nullptr;
  uint64_t out_bytes = 0;
  float* h_rad = nullptr;
  uint32_t* h_reason = nullptr;
  uint32_t* h_state = nullptr;
  uint8_t* h_cnt = nullptr;
  uint64_t* h_nin = nullptr;
  uint64_t* h_algo = nullptr;
  PairEffect* h_eff = nullptr;
  uint64_t eff_capacity = 0;  // effects downloaded with the batch
  uint32_t* h_flat = nullptr;
  std::vector<uint8_t> edge;
  hipEvent_t done = nullptr;
  uint32_t count = 0;
  bool algo = false;
  // The regions for a launch of bc candidates.
  void place(uint32_t bc) {
    auto a8 = [](uint64_t v) { return (v + 7) & ~(uint64_t)7; };
    uint64_t o = 0;
    const uint64_t o_rad = o;
    o = a8(o + 4ull * bc);
    const uint64_t o_reason = o;
    o = a8(o + 4ull * bc);
    const uint64_t o_state = o;
    o = a8(o + 4ull * bc);
    const uint64_t o_cnt = o;
    o = a8(o + bc);
    const uint64_t o_nin = o;
    o += 8ull * bc;
    const uint64_t o_algo = o;
    o += 8ull * bc;
    const uint64_t o_eff = o;
    eff_capacity = (uint64_t)bc * BATCH_EFF_PER;
    o += eff_capacity * sizeof(PairEffect);
    out_bytes = o;
    uint8_t* d = d_out.as<uint8_t>();
    p_rad = reinterpret_cast<float*>(d + o_rad);
    p_reason = reinterpret_cast<uint32_t*>(d + o_reason);
    p_state = reinterpret_cast<uint32_t*>(d + o_state);
    p_cnt = d + o_cnt;
    p_nin = reinterpret_cast<uint64_t*>(d + o_nin);
    p_algo = reinterpret_cast<uint64_t*>(d + o_algo);
    p_eff = reinterpret_cast<PairEffect*>(d + o_eff);
    h_rad = reinterpret_cast<float*>(h_out + o_rad);
    h_reason = reinterpret_cast<uint32_t*>(h_out + o_reason);
    h_state = reinterpret_cast<uint32_t*>(h_out + o_state);
    h_cnt = h_out + o_cnt;
    h_nin = reinterpret_cast<uint64_t*>(h_out + o_nin);
    h_algo = reinterpret_cast<uint64_t*>(h_out + o_algo);
    h_eff = reinterpret_cast<PairEffect*>(h_out + o_eff);
  }
};
// a launch's result block at its largest
constexpr uint64_t BATCH_OUT_BYTES =
  (uint64_t)BATCH_ROOM * (4 + 4 + 4 + 1 + 8 + 8 + BATCH_EFF_PER * sizeof(PairEffect)) + 64;
thread_local BatchSlot g_slots[2];

int host_alloc(void** p, size_t bytes) {
  SIF_HIP(hipHostMalloc(p, bytes, hipHostMallocDefault));
  return SIF_OK;
}

// Both slots at BATCH_ROOM, once per process.
int reserve_slots() {
  for (BatchSlot& B : g_slots) {
    if (B.done)
      continue;
    int e = B.d_flat.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_cx.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_cy.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_cz.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_cnt32.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_off.reserve(BATCH_ROOM * 8);
    if (e == SIF_OK) e = B.d_setmax.reserve(sizeof(float));
    if (e == SIF_OK) e = B.d_cells.reserve(BATCH_ROOM * 4);
    if (e == SIF_OK) e = B.d_out.reserve(BATCH_OUT_BYTES);
    if (e == SIF_OK) e = host_alloc((void**)&B.h_out, BATCH_OUT_BYTES);
    if (e == SIF_OK) e = host_alloc((void**)&B.h_flat, BATCH_ROOM * 4);
    if (e != SIF_OK)
      return e;
    B.edge.assign(BATCH_ROOM, 0);
    SIF_HIP(hipEventCreateWithFlags(&B.done, hipEventDisableTiming));
  }
  return SIF_OK;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
    std::chrono::steady_clock::now().time_since_epoch())
    .count();
}

// The catalogue on the device, appended rung by rung: the overlap index of the
// next rung is built from it and the mask is painted from it.
struct DeviceVoids {
  DeviceBuffer x, y, z, r;
  DeviceBuffer start, packed, occupied;  // the index: cell starts, voids by cell, occupancy
  uint32_t n = 0, cap = 0;
  // Bumped each time the arrays move (reserve_voids): an index built at an
  // earlier generation points at freed memory.
  uint32_t gen = 0;
  // pinned staging for an append
  float* hx = nullptr;
  float* hy = nullptr;
  float* hz = nullptr;
  float* hr = nullptr;
  uint32_t hcap = 0;
  hipEvent_t staged = nullptr;  // after the last append's copies from the staging
  float class_max[VOID_CLASSES] = {-1.0f, -1.0f, -1.0f, -1.0f};
};
thread_local DeviceVoids g_voids;
// the voids the device catalogue is first allocated for (sif_amd_set_void_room)
constexpr uint32_t VOID_ROOM_DEFAULT = 65536;
thread_local uint32_t g_void_room = 0;

int grow_keep(DeviceBuffer& b, size_t used, size_t want) {
  if (want <= b.bytes && b.ptr)
    return SIF_OK;
  DeviceBuffer nb;
  int st = nb.reserve(want);
  if (st != SIF_OK)
    return st;
  if (used)
    SIF_HIP(hipMemcpyAsync(nb.ptr, b.ptr, used, hipMemcpyDeviceToDevice, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  b.release();
  b.ptr = nb.ptr;
  b.bytes = nb.bytes;
  nb.ptr = nullptr;
  nb.bytes = 0;
  return SIF_OK;
}

// Room for `total` voids on the device. Growing moves the arrays (the
// contents kept, the index's own arrays not), which makes every index built
// before stale: so each rung reserves, before its index is built, room for
// every void appended until the next one is (start_filter).
int reserve_voids(uint64_t total) {
  DeviceVoids& D = g_voids;
  if (total <= D.cap)
    return SIF_OK;
  if (total > 0xffffffffull) {
    error_set(SIF_ERR_RANGE, "exodus: room for %llu voids (at most 2^32 - 1)",
      (unsigned long long)total);
    return SIF_ERR_RANGE;
  }
  uint64_t want = D.cap ? D.cap : (g_void_room ? g_void_room : VOID_ROOM_DEFAULT);
  while (want < total)
    want *= 2;
  if (want > 0xffffffffull)
    want = 0xffffffffull;
  int st = grow_keep(D.x, (size_t)D.n * 4, (size_t)want * 4);
  if (st == SIF_OK) st = grow_keep(D.y, (size_t)D.n * 4, (size_t)want * 4);
  if (st == SIF_OK) st = grow_keep(D.z, (size_t)D.n * 4, (size_t)want * 4);
  if (st == SIF_OK) st = grow_keep(D.r, (size_t)D.n * 4, (size_t)want * 4);
  // four entries past the last void: the scans read voids four at a time
  if (st == SIF_OK) st = D.packed.reserve((size_t)(want + 4) * sizeof(float4));
  if (st != SIF_OK)
    return st;
  D.cap = (uint32_t)want;
  D.gen++;
  return SIF_OK;
}

// Appends voids (host arrays) to the device catalogue and paints them into
// the device mask with radius scale * r. The host arrays may change as soon
// as this returns (they are staged in pinned memory; the stream keeps order).
int append_voids(const float* x, const float* y, const float* z, const float* r, uint32_t k,
  float box, float scale, float l, uint32_t grid_n) {
  if (k == 0)
    return SIF_OK;
  DeviceVoids& D = g_voids;
  if ((uint64_t)D.n + k > D.cap) {
    const int st = reserve_voids((uint64_t)D.n + k);
    if (st != SIF_OK)
      return st;
  }
  // the previous append's copies read the staging: let them finish (not the
  // rest of the stream, which may hold a growth launch in flight)
  if (!D.staged)
    SIF_HIP(hipEventCreateWithFlags(&D.staged, hipEventDisableTiming));
  else if (int w = wait_event(D.staged))
    return w;
  if (k > D.hcap) {
    for (float** p : {&D.hx, &D.hy, &D.hz, &D.hr})
      if (*p) {
        (void)hipHostFree(*p);
        *p = nullptr;
      }
    uint32_t want = D.hcap ? D.hcap : 16384;
    while (want < k)
      want *= 2;
    for (float** p : {&D.hx, &D.hy, &D.hz, &D.hr})
      SIF_HIP(hipHostMalloc((void**)p, (size_t)want * 4, hipHostMallocDefault));
    D.hcap = want;
  }
  memcpy(D.hx, x, (size_t)k * 4);
  memcpy(D.hy, y, (size_t)k * 4);
  memcpy(D.hz, z, (size_t)k * 4);
  memcpy(D.hr, r, (size_t)k * 4);
  SIF_HIP(hipMemcpyAsync(D.x.as<float>() + D.n, D.hx, (size_t)k * 4, hipMemcpyHostToDevice,
    stream()));
  SIF_HIP(hipMemcpyAsync(D.y.as<float>() + D.n, D.hy, (size_t)k * 4, hipMemcpyHostToDevice,
    stream()));
  SIF_HIP(hipMemcpyAsync(D.z.as<float>() + D.n, D.hz, (size_t)k * 4, hipMemcpyHostToDevice,
    stream()));
  SIF_HIP(hipMemcpyAsync(D.r.as<float>() + D.n, D.hr, (size_t)k * 4, hipMemcpyHostToDevice,
    stream()));
  SIF_HIP(hipEventRecord(D.staged, stream()));
  const float cell_w = box / (float)VOID_CELLS;
  for (uint32_t i = 0; i < k; i++) {
    const uint32_t cls = r[i] < 0.25f * cell_w ? 0u
                         : r[i] < 0.5f * cell_w ? 1u
                         : r[i] < cell_w        ? 2u
                                                : 3u;
    if (r[i] > D.class_max[cls])
      D.class_max[cls] = r[i];
  }
  int st = paint_voids(D.x.as<float>() + D.n, D.y.as<float>() + D.n, D.z.as<float>() + D.n,
    D.r.as<float>() + D.n, k, scale, l, grid_n, g_dev.mask.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  D.n += k;
  return SIF_OK;
}

// The overlap index of the catalogue on the device, for a rung to begin.
int device_void_index(const Voids& V, VoidIndex* out) {
  DeviceVoids& D = g_voids;
  const uint32_t slots = VOID_CLASSES * VOID_CELLS * VOID_CELLS * VOID_CELLS;
  int st = D.start.reserve((size_t)(slots + 1) * 4);
  if (st == SIF_OK)
    st = D.packed.reserve((size_t)(D.cap + 4) * sizeof(float4));
  if (st == SIF_OK)
    st = D.occupied.reserve((size_t)slots / 32 * 4);
  if (st == SIF_OK && D.n == 0) {  // keep the pointers valid
    st = D.x.reserve(4);
    if (st == SIF_OK) st = D.y.reserve(4);
    if (st == SIF_OK) st = D.z.reserve(4);
    if (st == SIF_OK) st = D.r.reserve(4);
  }
  if (st == SIF_OK)
    st = build_void_index(D.x.as<float>(), D.y.as<float>(), D.z.as<float>(), D.r.as<float>(),
      D.n, V.box, D.start.as<uint32_t>(), D.packed.as<float4>(), D.occupied.as<uint32_t>());
  if (st != SIF_OK)
    return st;
  out->cell_start = D.start.as<uint32_t>();
  out->voids = D.packed.as<float4>();
  out->occupied = D.occupied.as<uint32_t>();
  out->n = D.n;
  out->gen = D.gen;
  out->inv_l = V.inv_l;
  out->max_r = V.max_r;
  for (int c = 0; c < VOID_CLASSES; c++)
    out->class_max_r[c] = D.class_max[c];
  return SIF_OK;
}

int cmp_desc(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x < y) - (x > y);
}

}  // namespace

static int window_order_of(sif_option opt) {
  return (opt & SIF_FINDER_KEEP_CIC_WINDOW) ? 0 : 2;
}

int exodus_prepare_spectrum(const float* d_delta, uint32_t n) {
  const uint64_t sb = (uint64_t)n * spectrum_plane(n) * sizeof(float2);
  int st = g_dev.spec.reserve(sb);
  if (st == SIF_OK)
    st = g_dev.work.reserve(sb);
  if (st != SIF_OK)
    return st;
  return fft_r2c(d_delta, n, g_dev.spec.as<float2>(), g_dev.work.as<float2>());
}

// The top-hat LUT of every rung (fft.c:535-566, float libm on the host, as
// the reference builds it). Host threads build the tables in rung order,
// a chunk of entries at a time, into pinned memory; rung j's table goes to
// the device on the stream of its smoothing as soon as its chunks are done.
// The first smoothing then waits for one table rather than all of them, and
// the rest are built while the device smooths and grows the first rungs. A
// thread that needs a table not yet built builds its chunks itself.
class LutBuilder {
 public:
  int start(uint32_t n, float box, const std::vector<float>& radii) {
    box_ = box;
    radii_ = &radii;
    lsz_ = filter_lut_size(n);
    per_ = (lsz_ + CHUNK - 1) / CHUNK;
    total_ = per_ * radii.size();
    int st = reserve_host(lsz_ * radii.size());
    if (st != SIF_OK)
      return st;
    host_ = host_ptr();  // the workers' own thread_local slot is not this one
    done_ = std::vector<std::atomic<uint64_t>>(radii.size());
    for (auto& d : done_)
      d.store(0, std::memory_order_relaxed);
    next_.store(0, std::memory_order_relaxed);
    stop_.store(false, std::memory_order_relaxed);
    // a few threads: the tables are built far ahead of the rungs that use
    // them, and more would take the cores the launching thread needs
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned workers =
      (unsigned)std::min<uint64_t>(total_, std::min(std::max(1u, hw / 2), 6u));
    for (unsigned w = 0; w < workers; w++)
      pool_.emplace_back([this]() {
        while (!stop_.load(std::memory_order_relaxed) && claim()) {
        }
      });
    return SIF_OK;
  }
  // Rung j's table to d_lut, the copy ordered before stream s's next work.
  int issue(size_t j, float* d_lut, hipStream_t s) {
    while (done_[j].load(std::memory_order_acquire) < per_)
      if (!claim())
        std::this_thread::yield();
    SIF_HIP(hipMemcpyAsync(d_lut, host_ + j * lsz_, lsz_ * sizeof(float),
      hipMemcpyHostToDevice, s));
    return SIF_OK;
  }
  ~LutBuilder() {
    stop_.store(true, std::memory_order_relaxed);
    for (std::thread& t : pool_)
      t.join();
  }

 private:
  static constexpr uint64_t CHUNK = 16384;
  // one chunk of the next table in rung order; false when none is left
  bool claim() {
    const uint64_t c = next_.fetch_add(1, std::memory_order_relaxed);
    if (c >= total_)
      return false;
    const uint64_t j = c / per_, b = (c % per_) * CHUNK;
    const uint64_t e = std::min(b + CHUNK, lsz_);
    build_filter_lut_range(SIF_AMD_FILTER_TOP_HAT, (*radii_)[j], 0.0f, box_, (uint32_t)b,
      (uint32_t)e, host_ + j * lsz_);
    done_[j].fetch_add(1, std::memory_order_release);
    return true;
  }
  // The pinned tables of this host thread's runs, kept between runs. Every
  // copy from them is complete when a run returns (each precedes its rung's
  // smoothing, which the run waits for), so a larger run may replace them.
  static float*& host_ptr() {
    static thread_local float* p = nullptr;
    return p;
  }
  static uint64_t& host_cap() {
    static thread_local uint64_t c = 0;
    return c;
  }
  static int reserve_host(uint64_t floats) {
    if (host_cap() >= floats)
      return SIF_OK;
    if (host_ptr()) {
      SIF_HIP(hipHostFree(host_ptr()));
      host_ptr() = nullptr;
      host_cap() = 0;
    }
    SIF_HIP(hipHostMalloc((void**)&host_ptr(), floats * sizeof(float), hipHostMallocDefault));
    host_cap() = floats;
    return SIF_OK;
  }
  float box_ = 0.0f;
  float* host_ = nullptr;  // the calling thread's pinned tables
  const std::vector<float>* radii_ = nullptr;
  uint64_t lsz_ = 0, per_ = 0, total_ = 0;
  std::vector<std::atomic<uint64_t>> done_;
  std::atomic<uint64_t> next_{0};
  std::atomic<bool> stop_{false};
  std::vector<std::thread> pool_;
};

// The window axis (fft.c:890-975), uploaded once a run.
static int exodus_prepare_axis(uint32_t n, sif_option opt) {
  int st = g_dev.axis.reserve(n * sizeof(double));
  if (st != SIF_OK)
    return st;
  const int order = window_order_of(opt);
  if (order > 0) {
    std::vector<double> axis;
    build_window_axis(n, order, (opt & SIF_FINDER_ALIASING_WINDOW) ? 1 : 0, axis);
    SIF_HIP(hipMemcpy(g_dev.axis.ptr, axis.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
  return SIF_OK;
}

int exodus_smooth(uint32_t n, float box, float radius, sif_option opt, float* d_out) {
  return fft_filter_c2r(g_dev.spec.as<float2>(), n, box, SIF_AMD_FILTER_TOP_HAT, radius,
    0.0f, window_order_of(opt), (opt & SIF_FINDER_ALIASING_WINDOW) ? 1 : 0, d_out,
    g_dev.work.as<float2>());
}

int exodus_run(const RunSpec& R, sif_catalogue_t** out_cat, uint64_t** n_in_out,
  sif_amd_rung_stats_t* stats_out) {
  const uint32_t n = R.n;
  const uint64_t total = (uint64_t)n * n * n;
  const float l = R.box / (float)n;
  const float box_volume = R.box * R.box * R.box;
  const float mean_density = (float)R.n_particles / box_volume;
  const float vol = (4.0f / 3.0f) * SIF_PI * mean_density;
  // exodus_finder.c:1634-1642: the density tests of a weighted mesh use the
  // mean weight density, rounded and divided as the count's is
  MeshWeights mw;
  if (R.d_mw) {
    mw.w = R.d_mw;
    mw.cell = R.d_cw;
    mw.mass_factor = (4.0f / 3.0f) * SIF_PI * ((float)R.total_weight / box_volume);
  }
  const float scale =
    (R.opt & SIF_FINDER_LAGRANGIAN_OVERLAP) ? (float)cbrt(1.0 + (double)R.threshold) : 1.0f;
  const float sf = sif__finder_search_factor(R.opt);

  // SIF_FINDER_EXTEND_REACH (exodus_finder.c:1670-1702): a crossing past a
  // rung's reach is measured again, level by level, the reach growing by
  // EXT_GROWTH from the bottom rung to the largest search radius of the run;
  // level k's core ends EXT_SLACK below level k - 1's reach.
  constexpr float EXT_GROWTH = 1.5f, EXT_SLACK = 1.2f;
  constexpr uint32_t EXT_MAX_LEVELS = 32;
  std::vector<float> ext_reach, ext_rmin;
  if (R.opt & SIF_FINDER_EXTEND_REACH) {
    float r_max = R.radii[0];
    for (float r : R.radii)
      r_max = r > r_max ? r : r_max;
    const float r_ext = r_max + fmaxf((sf - 1.0f) * r_max, 3.0f * l);
    const float r_low = R.radii.back();
    uint32_t n_ext = 0;
    while (n_ext < EXT_MAX_LEVELS &&
           (n_ext == 0 || r_ext / powf(EXT_GROWTH, (float)(n_ext - 1)) > r_low))
      n_ext++;
    ext_reach.resize(n_ext);
    ext_rmin.resize(n_ext);
    for (uint32_t k = 0; k < n_ext; k++)
      ext_reach[k] = r_ext / powf(EXT_GROWTH, (float)(n_ext - 1 - k));
    for (uint32_t k = 0; k < n_ext; k++)
      ext_rmin[k] = (k ? ext_reach[k - 1] : ext_reach[0] / EXT_GROWTH) / EXT_SLACK;
  }
  const uint32_t n_ext = (uint32_t)ext_reach.size();

  // a survey (exodus_finder.c:1649-1653, 1729-1737): the randoms smoothed
  // beside the data each rung, the scanned field their contrast
  const bool survey = R.d_random != nullptr;
  const float survey_K = (1.0f + R.threshold) * R.alpha;
  SurveyMesh data_mesh;
  data_mesh.offsets = R.d_offsets;
  data_mesh.x = R.d_mx;
  data_mesh.y = R.d_my;
  data_mesh.z = R.d_mz;
  data_mesh.w = R.d_mw;
  data_mesh.cw = R.d_cw;
  data_mesh.mesh_n = R.mesh_n;
  int st = exodus_prepare_spectrum(R.d_delta, n);
  if (st == SIF_OK && survey) {
    st = g_dev.rspec.reserve((uint64_t)n * spectrum_plane(n) * sizeof(float2));
    if (st == SIF_OK)
      st = g_dev.rfield.reserve(total * sizeof(float));
    if (st == SIF_OK)
      st = fft_r2c(R.d_random, n, g_dev.rspec.as<float2>(), g_dev.work.as<float2>());
  }
  LutBuilder luts;
  const uint64_t lut_size = filter_lut_size(n);
  if (st == SIF_OK)
    st = luts.start(n, R.box, R.radii);
  if (st == SIF_OK)
    st = g_dev.luts.reserve(R.radii.size() * lut_size * sizeof(float));
  if (st == SIF_OK)
    st = exodus_prepare_axis(n, R.opt);
  if (st != SIF_OK)
    return st;
  const bool windowed = window_order_of(R.opt) > 0;
  // A field that is only scanned for the cells at or below the threshold is
  // not written: the smoothing lists those cells (CellLists), and the scan
  // reads the lists. The grid is then never smoothed in place, and comes
  // back as it was. Minima seeds read each cell's neighbours and a survey
  // divides two fields: those write the fields.
  const bool minima = (R.opt & SIF_FINDER_SEEDS_MINIMA) != 0;
  const bool use_lists = g_cell_lists && !survey && !minima && fft_lists_supported(n);
  CellLists lists[2];
  if (use_lists) {
    const uint64_t blocks = list_blocks(n);
    for (int b = 0; b < 2 && st == SIF_OK; b++) {
      st = g_dev.lkeys[b].reserve(blocks * LIST_BLOCK * sizeof(uint32_t));
      if (st == SIF_OK)
        st = g_dev.llocal[b].reserve(blocks * LIST_BLOCK * sizeof(uint16_t));
      if (st == SIF_OK)
        st = g_dev.lcount[b].reserve(blocks * sizeof(uint32_t));
      lists[b].keys = g_dev.lkeys[b].as<uint32_t>();
      lists[b].local = g_dev.llocal[b].as<uint16_t>();
      lists[b].count = g_dev.lcount[b].as<uint32_t>();
      lists[b].blocks = blocks;
    }
    if (st != SIF_OK)
      return st;
  }
  // rung j's scanned field into out (or its lists into lists[j & 1]), on
  // stream s: the smoothed contrast, or for a survey the smoothed data over
  // the smoothed randoms
  auto smooth = [&](size_t j, float* out, hipStream_t s) -> int {
    float* lut = g_dev.luts.as<float>() + j * lut_size;
    const double t_lut = now_ms();
    int le = luts.issue(j, lut, s);
    if (timing_active())
      timing_add(SIF_AMD_STAGE_HOST, now_ms() - t_lut);
    if (le != SIF_OK)
      return le;
    const double* axis = windowed ? g_dev.axis.as<double>() : nullptr;
    if (use_lists)
      return fft_c2r_lut_lists(g_dev.spec.as<float2>(), n, lut, axis, R.threshold, lists[j & 1],
        g_dev.work.as<float2>(), s);
    int e = fft_c2r_lut(g_dev.spec.as<float2>(), n, lut, axis, out, g_dev.work.as<float2>(), s);
    if (e == SIF_OK && survey)
      e = fft_c2r_lut(g_dev.rspec.as<float2>(), n, lut, axis, g_dev.rfield.as<float>(),
        g_dev.work.as<float2>(), s);
    if (e == SIF_OK && survey)
      e = survey_contrast(out, g_dev.rfield.as<float>(), R.d_occupied, R.alpha, total, s);
    return e;
  };

  // The mask of accepted-void nodes lives on the device: painted there, at
  // the end of each rung, with that rung's voids. Within a rung the host
  // asks only about the voids accepted since it began (RungVoids).
  Voids V;
  V.init(R.box);
  RungVoids RV;
  RV.init(R.box);
  RungVoids CL;  // the voids accepted by the commit under way
  CL.init(R.box);
  CL.columns = false;
  std::vector<uint32_t> taken;  // a commit's accepted candidates, in order
  // A candidate's tests against the rung's voids as its commit begins
  // Two bytes a candidate: the pool threads write them and the ordered pass
  // reads them from those threads' cores, a line at a time.
  struct PreProbe {
    uint16_t node : 1, pole : 1, pole_r : 1;
    uint16_t state : 2, state_r : 2;  // overlap states (overlap.h)
    // the commit's verdict if no earlier candidate of the batch that it
    // accepts affects this one (COMMIT_* below), or COMMIT_IN_ORDER when only
    // the ordered pass can tell (a state 2 sends it to the reference's scan
    // of the catalogue as it stands then)
    uint16_t verdict : 3;
  };
  static_assert(sizeof(PreProbe) == 2, "two bytes a candidate");
  enum : uint8_t {
    COMMIT_ACCEPT = 0,
    COMMIT_NODE,          // masked node
    COMMIT_PROXY,         // pole probe at the rung
    COMMIT_OVERLAP,       // overlap at the rung radius
    COMMIT_RESCALE_FAIL,  // did not grow
    COMMIT_GROWN_PROXY,   // pole probe at the grown radius
    COMMIT_GROWN_OVERLAP, // overlap at the grown radius
    COMMIT_IN_ORDER
  };
  std::vector<PreProbe> pre;
  std::vector<uint8_t> is_taken;  // accepted, by batch place
  st = g_dev.mask.reserve((total + 63) / 64 * sizeof(uint64_t));
  if (st == SIF_OK)
    st = reserve_slots();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_dev.mask.ptr, 0, (total + 63) / 64 * sizeof(uint64_t), stream()));
  g_voids.n = 0;
  for (int c = 0; c < VOID_CLASSES; c++)
    g_voids.class_max[c] = -1.0f;
  const size_t n_rungs = R.radii.size();
  st = g_dev.pf_counts.reserve(n_rungs * 3 * sizeof(uint64_t));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_dev.pf_counts.ptr, 0, n_rungs * 3 * sizeof(uint64_t), stream()));
  std::vector<uint64_t> pf_since(n_rungs * 3, 0);  // prefilter refusals masked since
  st = g_dev.since.reserve(4 * sizeof(uint64_t));
  if (st != SIF_OK)
    return st;
  if (!g_dev.h_since)
    SIF_HIP(hipHostMalloc((void**)&g_dev.h_since, 4 * sizeof(uint64_t), hipHostMallocDefault));
  if (R.prior && R.prior->n_voids) {
    for (uint64_t v = 0; v < R.prior->n_voids; v++)
      V.add(R.prior->cx[v], R.prior->cy[v], R.prior->cz[v], R.prior->radii[v], 0);
    st = append_voids(R.prior->cx, R.prior->cy, R.prior->cz, R.prior->radii,
      (uint32_t)R.prior->n_voids, R.box, scale, l, n);
    if (st != SIF_OK)
      return st;
  }

  std::vector<uint32_t> surv;
  uint64_t cand_cap = 0;
  // under SIF_FINDER_DISPLACE the mask is kept on the host (voids leave it)
  const bool displace = (R.opt & SIF_FINDER_DISPLACE) != 0;
  std::vector<uint64_t> hmask;

  // Two smoothed fields in turn: the caller's grid and a second buffer.
  if (!use_lists) {
    st = g_dev.field2.reserve(total * sizeof(float));
    if (st != SIF_OK)
      return st;
  }
  float* const fields[2] = {R.d_delta, g_dev.field2.as<float>()};
  static thread_local hipEvent_t ev_smoothed[2] = {nullptr, nullptr},
    ev_scanned[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; b++) {
    if (!ev_smoothed[b])
      SIF_HIP(hipEventCreateWithFlags(&ev_smoothed[b], hipEventDisableTiming));
    if (!ev_scanned[b])
      SIF_HIP(hipEventCreateWithFlags(&ev_scanned[b], hipEventDisableTiming));
  }


  // A rung's start -- its candidates scanned, prefiltered against the
  // catalogue on the device, compacted and sorted, the survivors on the host
  // -- runs in steps, so that the next rung's can be issued while this rung's
  // last batches are grown and committed: its waits then fall after those
  // commits instead of before the next rung. The next rung's prefilter sees
  // the catalogue as it stands when it is issued; the voids this rung accepts
  // after that start the next rung's list of newer voids (RV), which the
  // commit's own tests cover as they cover the rung's.
  struct RungStart {
    size_t rung = 0;
    int step = 0;  // 0 none, 1 counting, 2 prefilter issued, 3 survivors on the host
    uint64_t nc = 0, ns = 0;
    uint32_t key_differ = 0xffffffffu;
    VoidIndex VI;
    ScanJob job;
    // Started ahead: the candidates that the voids accepted after its
    // prefilter mask (the reference's scan, after them, never lists them),
    // by prefilter verdict -- kept (the commit refuses them as masked),
    // masked, proxy, overlap. Taken off the rung's counters.
    uint64_t since[4] = {0, 0, 0, 0};
  };
  RungStart rs;
  size_t smoothed_next = R.first_rung + 1;  // the first rung not yet issued to be smoothed
  auto reserve_candidates = [&]() {
    int e = g_dev.keys.reserve(cand_cap * 4);
    if (e == SIF_OK) e = g_dev.flat.reserve(cand_cap * 4);
    if (e == SIF_OK) e = g_dev.keep.reserve(cand_cap * 4);
    if (e == SIF_OK) e = g_dev.surv.reserve(cand_cap * 4);
    if (e == SIF_OK) e = g_dev.skeys.reserve(cand_cap * 4);
    if (e == SIF_OK) e = g_dev.why.reserve(cand_cap);
    return e;
  };
  // step 1: rung j's field (smoothed on the auxiliary stream, the first rung
  // here), and the scan's counting pass
  auto start_scan = [&](size_t j) -> int {
    float* const fj = fields[j & 1];
    if (j == R.first_rung) {
      const int e = smooth(j, fj, stream());
      if (e != SIF_OK)
        return e;
    } else {
      SIF_HIP(hipStreamWaitEvent(stream(), ev_smoothed[j & 1], 0));
    }
    if (cand_cap == 0) {
      // a 64th of the cells, or the room an earlier run left
      const uint64_t left = std::min({g_dev.keys.bytes / 4, g_dev.flat.bytes / 4,
        g_dev.keep.bytes / 4, g_dev.surv.bytes / 4, g_dev.skeys.bytes / 4, g_dev.why.bytes});
      cand_cap = std::max<uint64_t>(total / 64 + 1024, left);
      const int e = reserve_candidates();
      if (e != SIF_OK)
        return e;
    }
    rs = RungStart();
    rs.rung = j;
    rs.job.delta = fj;
    rs.job.lists = use_lists ? &lists[j & 1] : nullptr;
    rs.job.mask = g_dev.mask.as<uint64_t>();
    rs.job.total = total;
    rs.job.t = R.threshold;
    rs.job.minima_n = minima ? n : 0u;
    rs.job.keys = g_dev.keys.as<uint32_t>();
    rs.job.flat = g_dev.flat.as<uint32_t>();
    rs.job.cap = cand_cap;
    const int e = scan_begin(rs.job);
    if (e == SIF_OK)
      rs.step = 1;
    return e;
  };
  // the candidates in flat order (the buffers grown and the scan run again
  // when they are short); the scan was the field's last reader, so the next
  // rung's field is smoothed into the other buffer now
  auto finish_scan = [&]() -> int {
    const size_t j = rs.rung;
    int e = scan_finish(rs.job, g_dev.keys.as<uint32_t>(), g_dev.flat.as<uint32_t>(), cand_cap);
    while (e == SIF_OK && rs.job.count > cand_cap) {
      // the counts stand (the mask is painted only outside a scan's two
      // halves), so only the writing pass is issued again
      cand_cap = (uint64_t)((double)rs.job.count * 1.5) + 1;
      g_dev.keys.release();
      g_dev.flat.release();
      g_dev.keep.release();
      g_dev.surv.release();
      g_dev.skeys.release();
      g_dev.why.release();
      e = reserve_candidates();
      if (e == SIF_OK)
        e = scan_finish(rs.job, g_dev.keys.as<uint32_t>(), g_dev.flat.as<uint32_t>(), cand_cap);
    }
    if (e != SIF_OK)
      return e;
    rs.nc = rs.job.count;
    rs.key_differ = rs.job.key_differ;
    SIF_HIP(hipEventRecord(ev_scanned[j & 1], stream()));
    // The next rungs' fields on the auxiliary stream, two rungs ahead, after
    // this scan: rung j + 2 into the buffer this scan has just read, and rung
    // j + 1, if the previous scan did not issue it, into the other one, whose
    // last reader (rung j - 1's scan) is behind this one. Every earlier use of
    // the work buffer is behind it too. The auxiliary stream then smooths
    // rung after rung while the rungs that find nothing go by.
    SIF_HIP(hipStreamWaitEvent(aux_stream(), ev_scanned[j & 1], 0));
    for (size_t jn = std::max(j + 1, smoothed_next); jn <= j + 2 && jn < R.radii.size();
         jn++) {
      e = smooth(jn, fields[jn & 1], aux_stream());
      if (e != SIF_OK)
        return e;
      SIF_HIP(hipEventRecord(ev_smoothed[jn & 1], aux_stream()));
      smoothed_next = jn + 1;
    }
    return SIF_OK;
  };
  // step 2: the prefilter against the catalogue on the device (the mask and
  // the void index built now), in flat order (the verdicts do not depend on
  // the order; neighbouring candidates share void cells), and the compaction
  auto start_filter = [&]() -> int {
    const size_t j = rs.rung;
    const double th = now_ms();
    // Room for every void appended before the next rung's index is built:
    // the rest of the rung under way (a start that goes ahead comes before
    // its last commit, at most a batch) and this rung's own (at most its
    // candidates). The index then stays valid for this rung's batches.
    int e = reserve_voids((uint64_t)g_voids.n + rs.nc + 2ull * BATCH_ROOM);
    if (e == SIF_OK)
      e = device_void_index(V, &rs.VI);
    if (timing_active())
      timing_add(SIF_AMD_STAGE_HOST, now_ms() - th);
    if (e != SIF_OK)
      return e;
    const int32_t proxy_j = (int32_t)(scale * R.radii[j] * (1.0f - R.f) / l) - 1;
    timing_begin(SIF_AMD_STAGE_PREFILTER);
    e = prefilter(g_dev.flat.as<uint32_t>(), (uint32_t)rs.nc, n, l, R.box,
      g_dev.mask.as<uint64_t>(), proxy_j, R.radii[j], R.f, scale, rs.VI,
      g_dev.keep.as<uint32_t>(), g_dev.pf_counts.as<uint64_t>() + 3 * j,
      g_dev.why.as<uint8_t>());
    if (e == SIF_OK)
      e = compact2_u32_begin(g_dev.flat.as<uint32_t>(), g_dev.keys.as<uint32_t>(),
        g_dev.keep.as<uint32_t>(), rs.nc, g_dev.surv.as<uint32_t>(), g_dev.skeys.as<uint32_t>());
    if (e != SIF_OK)
      return e;
    timing_end(SIF_AMD_STAGE_PREFILTER);
    rs.step = 2;
    return SIF_OK;
  };
  // step 3: the survivors sorted deepest first, and on the host; for a start
  // that went ahead, first the candidates masked since its prefilter (the
  // mask now holds every void the rung starts with)
  auto finish_filter = [&](bool went_ahead) -> int {
    const double th = now_ms();
    if (went_ahead && rs.nc) {
      SIF_HIP(hipMemsetAsync(g_dev.since.ptr, 0, 4 * sizeof(uint64_t), stream()));
      const int e = count_masked_since(g_dev.flat.as<uint32_t>(), g_dev.why.as<uint8_t>(), rs.nc,
        g_dev.mask.as<uint64_t>(), g_dev.since.as<uint64_t>());
      if (e != SIF_OK)
        return e;
      SIF_HIP(hipMemcpyAsync(g_dev.h_since, g_dev.since.ptr, 4 * sizeof(uint64_t),
        hipMemcpyDeviceToHost, stream()));
    }
    int e = compact2_u32_finish(&rs.ns);
    if (e == SIF_OK)  // the survivors are a subset of the scanned candidates
      e = sort_candidates(g_dev.skeys.as<uint32_t>(), g_dev.surv.as<uint32_t>(), rs.ns,
        rs.key_differ);
    if (e != SIF_OK)
      return e;
    surv.resize(rs.ns);
    if (rs.ns)
      SIF_HIP(hipMemcpyAsync(surv.data(), g_dev.surv.ptr, rs.ns * 4, hipMemcpyDeviceToHost,
        stream()));
    e = wait_stream(stream());
    if (e != SIF_OK)
      return e;
    if (went_ahead && rs.nc)
      for (int w = 0; w < 4; w++)
        rs.since[w] = g_dev.h_since[w];
    if (timing_active())
      timing_add(SIF_AMD_STAGE_HOST, now_ms() - th);
    rs.step = 3;
    return SIF_OK;
  };
  // the rung's voids not yet on the device: appended and painted into the mask
  size_t rv_appended = 0;
  auto append_rung_voids = [&]() -> int {
    if (rv_appended >= RV.cx.size())
      return SIF_OK;
    const size_t a = rv_appended;
    rv_appended = RV.cx.size();
    return append_voids(RV.cx.data() + a, RV.cy.data() + a, RV.cz.data() + a, RV.r.data() + a,
      (uint32_t)(rv_appended - a), R.box, scale, l, n);
  };

  uint64_t mismatched = R.mismatched0;
  for (size_t i = R.first_rung; i < R.radii.size() && st == SIF_OK; i++) {
    const float radius = R.radii[i];
    sif_amd_rung_stats_t S;
    memset(&S, 0, sizeof(S));
    S.radius = radius;

    // The rung's start, unless the previous rung issued it ahead.
    if (!(rs.step == 3 && rs.rung == i)) {
      st = start_scan(i);
      if (st == SIF_OK)
        st = finish_scan();
      if (st == SIF_OK && !displace)
        st = start_filter();
      if (st == SIF_OK && !displace)
        st = finish_filter(false);
      if (st != SIF_OK)
        break;
    }
    const uint64_t nc = rs.nc;
    const uint32_t key_differ = rs.key_differ;
    // the reference's counters: the candidates its scan lists, the prefilter's
    // refusals among them, and the commit's (less those masked since, below)
    S.candidates = nc - rs.since[0] - rs.since[1] - rs.since[2] - rs.since[3];
    for (int w = 1; w < 4; w++)
      pf_since[3 * i + w - 1] = rs.since[w];
    const uint64_t masked_since = rs.since[0];

    const float rmin = 1.00f * radius;
    float r_search = radius + fmaxf((sf - 1.0f) * radius, 3.0f * l);
    if (i > 0 && R.radii[i - 1] > r_search)
      r_search = R.radii[i - 1];
    const int32_t proxy = (int32_t)(scale * radius * (1.0f - R.f) / l) - 1;
    // the first extended level this rung enters (exodus_finder.c:1824-1832):
    // the lowest EXT_SLACK past its own reach, or the top if none is
    uint32_t ext0 = n_ext;
    for (uint32_t k = 0; k < n_ext && ext0 == n_ext; k++)
      if (ext_reach[k] >= EXT_SLACK * r_search)
        ext0 = k;
    if (ext0 == n_ext && n_ext > 0 && ext_reach[n_ext - 1] > r_search)
      ext0 = n_ext - 1;

    if (displace) {
      // SIF_FINDER_DISPLACE (exodus_finder.c:1871-1880, 2210-2275): no
      // prefilter; each candidate in order is tested at the rung against the
      // catalogue as it stands, grown, and refused only by a void at least as
      // large, the smaller ones it collides with displaced. The reference
      // ends its batch at each displacement, so its commit is this sequence.
      // Candidates are grown ahead in windows on the device (a growth
      // depends on the candidate alone); a candidate that a displacement
      // frees after its window was formed starts the next window.
      if (hmask.empty()) {  // the mask as the run begins: prior voids painted
        hmask.resize((total + 63) / 64);
        SIF_HIP(hipMemcpyAsync(hmask.data(), g_dev.mask.ptr, hmask.size() * 8,
          hipMemcpyDeviceToHost, stream()));
      }
      st = sort_candidates(g_dev.keys.as<uint32_t>(), g_dev.flat.as<uint32_t>(), nc,
        key_differ);
      if (st != SIF_OK)
        break;
      surv.resize(nc);
      if (nc)
        SIF_HIP(hipMemcpyAsync(surv.data(), g_dev.flat.ptr, nc * 4, hipMemcpyDeviceToHost,
          stream()));
      SIF_HIP(hipStreamSynchronize(stream()));
      const uint32_t nn = n, nn2 = n * n;
      auto masked_at = [&](uint32_t flat) { return (hmask[flat >> 6] >> (flat & 63)) & 1u; };
      // 0 pass, 1 masked, 2 proxy, 3 overlap (exodus_finder.c:2220-2240)
      // (V linked before: the tests only read the catalogue and the mask, and
      // run on the pool while a window forms)
      auto tests = [&](uint32_t flat) -> int {
        if (masked_at(flat))
          return 1;
        const uint32_t ix = flat / nn2, iy = (flat / nn) % nn, iz = flat % nn;
        if (proxy > 0) {  // sif__overlap_quick (utils.c:819-848)
          const int32_t p = proxy, ni = (int32_t)n;
          auto w = [&](int32_t v) { return (uint32_t)(((v % ni) + ni) % ni); };
          const uint32_t probes[6] = {(w((int32_t)ix + p) * nn + iy) * nn + iz,
            (w((int32_t)ix - p) * nn + iy) * nn + iz, (ix * nn + w((int32_t)iy + p)) * nn + iz,
            (ix * nn + w((int32_t)iy - p)) * nn + iz, (ix * nn + iy) * nn + w((int32_t)iz + p),
            (ix * nn + iy) * nn + w((int32_t)iz - p)};
          for (uint32_t q : probes)
            if (masked_at(q))
              return 2;
        }
        if (V.overlap_linked((float)ix * l, (float)iy * l, (float)iz * l, radius, R.f, scale))
          return 3;
        return 0;
      };
      std::vector<uint8_t> why_ahead;
      std::vector<uint32_t> grow_flat, conflicts;
      std::vector<int32_t> slot;  // a window candidate's growth result, or -1
      std::vector<float> hrad;
      std::vector<uint32_t> hreason;
      std::vector<uint64_t> hnin;
      const uint32_t window = g_batch_floor;
      uint64_t k = 0;
      while (st == SIF_OK && k < nc) {
        // the window: candidates from k, those passing as things stand grown;
        // the tests of a stretch of candidates run on the pool at once
        grow_flat.clear();
        slot.clear();
        uint64_t end = k;
        V.link();
        while (end < nc && grow_flat.size() < window) {
          const uint64_t m = std::min<uint64_t>(nc - end, 2 * window);
          why_ahead.resize(m);
          const uint64_t e0 = end;
          host_pool().run(m, [&](uint64_t a, uint64_t b) {
            for (uint64_t t = a; t < b; t++)
              why_ahead[t] = (uint8_t)tests(surv[e0 + t]);
          });
          for (uint64_t t = 0; t < m && grow_flat.size() < window; t++, end++) {
            const int why = why_ahead[t];
            slot.push_back(why ? -1 : (int32_t)grow_flat.size());
            if (!why)
              grow_flat.push_back(surv[end]);
          }
        }
        const uint32_t gc = (uint32_t)grow_flat.size();
        if (gc) {
          st = disp_bufs.reserve(gc);
          if (st != SIF_OK)
            break;
          SIF_HIP(hipMemcpyAsync(disp_bufs.flat.ptr, grow_flat.data(), gc * 4,
            hipMemcpyHostToDevice, stream()));
          st = node_centres(disp_bufs.flat.as<uint32_t>(), gc, n, l, disp_bufs.cx.as<float>(),
            disp_bufs.cy.as<float>(), disp_bufs.cz.as<float>());
          if (st == SIF_OK)
            st = grow_batch(R.d_offsets, R.d_mx, R.d_my, R.d_mz, R.mesh_n, R.box,
              disp_bufs.cx.as<float>(), disp_bufs.cy.as<float>(), disp_bufs.cz.as<float>(), gc,
              rmin, r_search, R.threshold, vol, disp_bufs.rad.as<float>(),
              disp_bufs.reason.as<uint32_t>(), disp_bufs.nin.as<uint64_t>(), nullptr, &mw);
          if (st == SIF_OK && ext0 < n_ext)
            st = grow_extend(R.d_offsets, R.d_mx, R.d_my, R.d_mz, R.mesh_n, R.box,
              disp_bufs.cx.as<float>(), disp_bufs.cy.as<float>(), disp_bufs.cz.as<float>(), gc,
              ext_rmin.data(), ext_reach.data(), ext0, n_ext, R.threshold, vol,
              disp_bufs.rad.as<float>(), disp_bufs.reason.as<uint32_t>(),
              disp_bufs.nin.as<uint64_t>(), &mw);
          if (st != SIF_OK)
            break;
          hrad.resize(gc);
          hreason.resize(gc);
          hnin.resize(gc);
          SIF_HIP(hipMemcpyAsync(hrad.data(), disp_bufs.rad.ptr, gc * 4, hipMemcpyDeviceToHost,
            stream()));
          SIF_HIP(hipMemcpyAsync(hreason.data(), disp_bufs.reason.ptr, gc * 4,
            hipMemcpyDeviceToHost, stream()));
          SIF_HIP(hipMemcpyAsync(hnin.data(), disp_bufs.nin.ptr, gc * 8,
            hipMemcpyDeviceToHost, stream()));
          SIF_HIP(hipStreamSynchronize(stream()));
          S.evaluated += gc;
        }
        // the commit, in order, against the catalogue as it stands
        uint64_t q = k;
        for (; q < end; q++) {
          const uint32_t flat = surv[q];
          V.link();
          const int why = tests(flat);
          if (why) {
            if (why == 1) S.masked++;
            else if (why == 2) S.proxy++;
            else S.overlap++;
            continue;
          }
          const int32_t g = slot[q - k];
          if (g < 0)
            break;  // freed by a displacement since the window formed
          if (hreason[g] == 5) {
            error_set(SIF_ERR_RANGE,
              "exodus: a search shell at rung %g holds more tracers in one histogram "
              "bin than the refinement buffer takes",
              (double)radius);
            st = SIF_ERR_RANGE;
            break;
          }
          const float rr = hrad[g];
          if (!(rr >= 0.0f)) {
            S.rescale_fail++;
            continue;
          }
          const float x = (float)(flat / nn2) * l, y = (float)((flat / nn) % nn) * l,
                      z = (float)(flat % nn) * l;
          const float largest = V.conflicts(x, y, z, rr, R.f, scale, &conflicts);
          if (largest >= rr) {
            S.overlap++;
            continue;
          }
          for (uint32_t j : conflicts) {  // displace_void (:946-1002)
            const float jx = V.cx[j], jy = V.cy[j], jz = V.cz[j], jr = V.r[j];
            V.unlink(j);
            V.r[j] = -jr;
            host_paint(hmask, jx, jy, jz, scale * jr, l, n, false);
            V.visit(jx, jy, jz, scale * (jr + V.max_r), [&](uint32_t m) {
              float dx, dy, dz;
              V.dist(jx, jy, jz, m, &dx, &dy, &dz);
              const float t = scale * (jr + V.r[m]);
              if (dx * dx + dy * dy + dz * dz < t * t)
                host_paint(hmask, V.cx[m], V.cy[m], V.cz[m], scale * V.r[m], l, n, true);
            });
          }
          // accept_void (:656-683)
          V.add(x, y, z, rr, hnin[g]);
          host_paint(hmask, x, y, z, scale * rr, l, n, true);
          if (rr / radius > 2.0f)  // exodus_finder.c:2321-2328
            mismatched++;
          S.accepted++;
        }
        k = q;
      }
      if (st != SIF_OK)
        break;
      if (st == SIF_OK)
        st = all_reads_linked(V);
      if (st != SIF_OK)
        break;
      // the rung's survivors re-listed, the mask to the device for the next scan
      V.compact();
      SIF_HIP(hipMemcpyAsync(g_dev.mask.ptr, hmask.data(), hmask.size() * 8,
        hipMemcpyHostToDevice, stream()));
      SIF_HIP(hipStreamSynchronize(stream()));
      if (stats_out)
        stats_out[i] = S;
      if (R.on_rung)
        R.on_rung((uint32_t)(i + 1), V.cx.data(), V.cy.data(), V.cz.data(), V.r.data(),
          (uint64_t)V.cx.size(), mismatched);
      continue;
    }

    const VoidIndex VI = rs.VI;
    const uint64_t ns = rs.ns;
    double th = 0.0;

    // Speculative batches, sequential commit. While the host commits one
    // batch, the GPU grows the next, assembled from the catalogue as it stood
    // before that commit; every candidate is tested again as it is
    // committed, so the catalogue is that of a sequential pass whatever the
    // batching.
    const uint32_t batch_floor = g_batch_floor, batch_max = g_batch_max;
    uint32_t cap = batch_floor;
    uint64_t k = 0;
    // flat < n^3 <= 2^32: 32-bit arithmetic
    const uint32_t nn = n, nn2 = n * n;
    auto node_of = [&](uint32_t flat, uint32_t* ix, uint32_t* iy, uint32_t* iz) {
      *ix = flat / nn2;
      const uint32_t rest = flat - *ix * nn2;
      *iy = rest / nn;
      *iz = rest - *iy * nn;
    };
    // An exact overlap with the catalogue as it stands (utils.c:850-948),
    // from the overlap state of the voids the rung began with (old_state,
    // from the device) and of those accepted since (new_state, from the
    // host's probe); a state 2 on either side sends the candidate to the
    // reference's own scan of the whole catalogue.
    // The scan links the voids not yet in the cell list: the ordered commit
    // (one thread) may; rung_tests, which runs on the host pool, must not,
    // and reads a list that assemble() linked before starting the pool.
    auto exact_overlap = [&](float cx, float cy, float cz, float rr, uint32_t old_state,
                           int new_state) {
      if (old_state == 1 || (old_state == 0 && new_state == 1))
        return true;
      if (old_state == 0 && new_state == 0)
        return false;
      return V.overlap_exact(cx, cy, cz, rr, R.f, scale);
    };
    auto exact_overlap_linked = [&](float cx, float cy, float cz, float rr,
                                  uint32_t old_state, int new_state) {
      if (old_state == 1 || (old_state == 0 && new_state == 1))
        return true;
      if (old_state == 0 && new_state == 0)
        return false;
      return V.overlap_linked(cx, cy, cz, rr, R.f, scale);
    };
    // The rung-radius tests against the catalogue as it stands: 0 pass,
    // 1 masked, 2 proxy, 3 overlap. The mask and pole probe against the
    // rung's older voids were passed in the prefilter (the device mask does
    // not change within a rung), so only the new voids are probed here.
    // Runs on the host pool: the catalogue is only read.
    auto rung_tests = [&](uint32_t flat, bool edge) -> int {
      uint32_t ix, iy, iz;
      node_of(flat, &ix, &iy, &iz);
      const float cx = (float)ix * l, cy = (float)iy * l, cz = (float)iz * l;
      OverlapQuery q;
      overlap_query(q, cx, cy, cz, radius, V.max_r, R.box, R.f, scale, V.inv_l);
      const RungProbe p = RV.probe(&q, ix, iy, iz, cx, cy, cz, true, proxy, scale, l, n);
      if (p.node)
        return 1;
      if (proxy > 0 && p.pole)
        return 2;
      if (exact_overlap_linked(cx, cy, cz, radius, edge ? 2u : 0u, p.state))
        return 3;
      return 0;
    };
    auto count_refusal = [&](int why) {
      if (why == 1)
        S.masked++;
      else if (why == 2)
        S.proxy++;
      else
        S.overlap++;
    };
    // The next `cap` survivors that pass the rung tests now. The tests of a
    // block of survivors are independent (nothing is accepted meanwhile), so
    // they run on the host pool; the results are taken in order, and those
    // past the last candidate taken are tested again next time.
    std::vector<uint8_t> why_block;
    auto assemble = [&](BatchSlot& B) {
      const double t0 = now_ms();
      B.count = 0;
      while (k < ns && B.count < cap) {
        const uint64_t m = std::min<uint64_t>(ns - k, std::max<uint64_t>(2048, 2 * (cap - B.count)));
        why_block.resize(m);
        const uint64_t k0 = k;
        V.link();  // the pool's scans only read the cell list
        host_pool().run(m, [&](uint64_t a, uint64_t b) {
          for (uint64_t t = a; t < b; t++)
            why_block[t] =
              (uint8_t)rung_tests(surv[k0 + t] & 0x7fffffffu, (surv[k0 + t] >> 31) != 0);
        });
        uint64_t t = 0;
        for (; t < m && B.count < cap; t++) {
          if (why_block[t]) {
            count_refusal(why_block[t]);
          } else {
            B.h_flat[B.count] = surv[k0 + t] & 0x7fffffffu;
            B.edge[B.count] = (surv[k0 + t] >> 31) != 0 ? 1 : 0;
            B.count++;
          }
        }
        k = k0 + t;
      }
      if (timing_active())
        timing_add(SIF_AMD_STAGE_RETEST, now_ms() - t0);
    };
    // growth and the grown-radius overlap state, results to pinned memory
    auto launch = [&](BatchSlot& B) -> int {
      const uint32_t bc = B.count;
      const bool algo = timing_active() && !survey;
      B.place(bc);
      SIF_HIP(hipMemcpyAsync(B.d_flat.ptr, B.h_flat, bc * 4, hipMemcpyHostToDevice, stream()));
      int e = node_centres(B.d_flat.as<uint32_t>(), bc, n, l, B.d_cx.as<float>(),
        B.d_cy.as<float>(), B.d_cz.as<float>());
      if (e == SIF_OK && survey)  // exodus_finder.c:2045-2049
        e = grow_survey(data_mesh, R.rmesh, R.box, B.d_cx.as<float>(), B.d_cy.as<float>(),
          B.d_cz.as<float>(), bc, rmin, r_search, survey_K, B.p_rad, B.p_reason, B.p_nin);
      else if (e == SIF_OK)
        e = grow_batch(R.d_offsets, R.d_mx, R.d_my, R.d_mz, R.mesh_n, R.box,
          B.d_cx.as<float>(), B.d_cy.as<float>(), B.d_cz.as<float>(), bc, rmin, r_search,
          R.threshold, vol, B.p_rad, B.p_reason, B.p_nin, algo ? B.p_algo : nullptr, &mw);
      if (e == SIF_OK && ext0 < n_ext)
        e = grow_extend(R.d_offsets, R.d_mx, R.d_my, R.d_mz, R.mesh_n, R.box,
          B.d_cx.as<float>(), B.d_cy.as<float>(), B.d_cz.as<float>(), bc, ext_rmin.data(),
          ext_reach.data(), ext0, n_ext, R.threshold, vol, B.p_rad, B.p_reason, B.p_nin, &mw);
      if (e == SIF_OK && VI.gen != g_voids.gen) {
        error_set(SIF_ERR_INVALID, "exodus: internal: the rung's void index predates a move "
          "of the device void arrays");
        e = SIF_ERR_INVALID;
      }
      if (e == SIF_OK)
        e = grown_overlap(B.d_cx.as<float>(), B.d_cy.as<float>(), B.d_cz.as<float>(),
          B.d_flat.as<uint32_t>(), B.p_rad, bc, VI, V.max_r, R.box, R.f, scale,
          g_dev.mask.as<uint64_t>(), n, l, B.p_state);
      if (e == SIF_OK) {
        BatchPairParams bp;
        bp.n = n;
        bp.l = l;
        bp.box = R.box;
        bp.f = R.f;
        bp.scale = scale;
        bp.radius = radius;
        bp.cat_max_r = V.max_r;
        bp.inv_l = V.inv_l;
        bp.proxy = proxy;
        BatchPairScratch W;
        W.set_max_r = B.d_setmax.as<float>();
        W.cells = B.d_cells.as<uint32_t>();
        W.cnt32 = B.d_cnt32.as<uint32_t>();
        W.cnt8 = B.p_cnt;
        W.offset = B.d_off.as<uint64_t>();
        W.eff = B.p_eff;
        W.capacity = B.eff_capacity;
        e = batch_pairs(B.d_flat.as<uint32_t>(), B.p_rad, B.p_state, bc, bp, W);
      }
      if (e != SIF_OK)
        return e;
      // the results, one copy
      SIF_HIP(hipMemcpyAsync(B.h_out, B.d_out.ptr, B.out_bytes, hipMemcpyDeviceToHost,
        stream()));
      B.algo = algo;
      SIF_HIP(hipEventRecord(B.done, stream()));
      S.evaluated += bc;
      return SIF_OK;
    };
    // exodus_finder.c:2174-2329, in candidate order. Every test is an OR over
    // voids, so it splits by the voids' age: the rung's older voids (the
    // device answered: prefilter, grown-radius state and pole bit), the rung's
    // new voids as the commit begins, and the voids this commit accepts --
    // the batch's earlier candidates, whose effects on each candidate are
    // listed before the commit (batch_pairs.h). The first two run for every
    // candidate at once on the host pool; the commit in order then ORs the
    // effects of the candidates it has accepted.
    auto combine = [](int a, int b) { return (a == 1 || b == 1) ? 1 : (a | b); };
    auto commit = [&](BatchSlot& B, uint64_t* accepted_out) -> int {
      {
        const int w = wait_event(B.done);
        if (w != SIF_OK)
          return w;
      }
      const double t2 = now_ms();
      for (uint32_t b = 0; b < B.count; b++)
        if (B.h_reason[b] == 5) {
          error_set(SIF_ERR_RANGE,
            "exodus: a search shell at rung %g holds more tracers in one histogram "
            "bin than the refinement buffer takes",
            (double)radius);
          return SIF_ERR_RANGE;
        }
      const float cat_max_r = V.max_r;
      const double t_idx = now_ms();
      // The rung's newer voids (RV) for every candidate at once. Each thread
      // decodes the nodes it needs itself: a table of nodes written by one
      // thread and read by the others costs a cache line's trip between
      // cores per few candidates.
      pre.resize(B.count);
      host_pool().run(B.count, [&](uint64_t b0, uint64_t b1) {
        for (uint64_t b = b0; b < b1; b++) {
          uint32_t ix, iy, iz;
          node_of(B.h_flat[b], &ix, &iy, &iz);
          const float x = (float)ix * l, y = (float)iy * l, z = (float)iz * l;
          OverlapQuery q;
          overlap_query(q, x, y, z, radius, cat_max_r, R.box, R.f, scale, V.inv_l);
          const RungProbe p1 = RV.probe(&q, ix, iy, iz, x, y, z, true, proxy, scale, l, n);
          PreProbe P;
          P.node = p1.node;
          P.pole = p1.pole;
          P.state = (int8_t)p1.state;
          P.pole_r = false;
          P.state_r = 0;
          const float r = B.h_rad[b];
          const int32_t pole_r = r >= 0.0f ? (int32_t)(scale * r * (1.0f - R.f) / l) - 1 : 0;
          const uint32_t gs = B.h_state[b];
          // refused at the rung radius by the older voids alone
          const bool rung_refused = p1.node || (proxy > 0 && p1.pole) || p1.state == 1;
          if (r >= 0.0f && !rung_refused) {
            if (pole_r > 0 && (gs & GROWN_POLE_BIT)) {
              P.pole_r = true;
            } else {
              OverlapQuery q2;
              overlap_query(q2, x, y, z, r, cat_max_r, R.box, R.f, scale, V.inv_l);
              const RungProbe p2 =
                RV.probe((gs & 3u) == 1u ? nullptr : &q2, ix, iy, iz, x, y, z, false, pole_r,
                  scale, l, n);
              P.pole_r = p2.pole;
              P.state_r = (int8_t)p2.state;
            }
          }
          // The ordered pass's tests below with no effect from the batch:
          // exact_overlap decided from the two states alone, or not at all
          // here (a 2 on either side, then only the ordered pass may scan).
          auto decided = [](uint32_t old_state, int new_state) -> int {
            if (old_state == 1 || (old_state == 0 && new_state == 1))
              return 1;
            if (old_state == 0 && new_state == 0)
              return 0;
            return -1;
          };
          const int32_t pole = r >= 0.0f ? pole_r : 0;
          int ov = 0;
          if (P.node)
            P.verdict = COMMIT_NODE;
          else if (proxy > 0 && P.pole)
            P.verdict = COMMIT_PROXY;
          else if ((ov = decided(B.edge[b] ? 2u : 0u, P.state)) != 0)
            P.verdict = ov > 0 ? COMMIT_OVERLAP : COMMIT_IN_ORDER;
          else if (r < 0.0f)
            P.verdict = COMMIT_RESCALE_FAIL;
          else if (pole > 0 && ((gs & GROWN_POLE_BIT) || P.pole_r))
            P.verdict = COMMIT_GROWN_PROXY;
          else if ((ov = decided(gs & 3u, P.state_r)) != 0)
            P.verdict = ov > 0 ? COMMIT_GROWN_OVERLAP : COMMIT_IN_ORDER;
          else
            P.verdict = COMMIT_ACCEPT;
          pre[b] = P;
        }
      });
      // In order. The voids accepted join the catalogue at once; the effects
      // of the batch's earlier candidates come listed from the device
      // (batch_pairs.hip); the list of this commit's voids (CL) is filled only
      // when a candidate with more effects than were listed needs it.
      const double t3 = now_ms();
      CL.clear();
      size_t cl_filled = 0;
      taken.clear();
      is_taken.assign(B.count, 0);
      uint64_t accepted = 0;
      uint64_t eff_at = 0;  // the batch's effects list: this candidate's offset
      auto centre_of = [&](uint32_t b, uint32_t* ix, uint32_t* iy, uint32_t* iz, float* x,
                         float* y, float* z) {
        node_of(B.h_flat[b], ix, iy, iz);
        *x = (float)*ix * l;
        *y = (float)*iy * l;
        *z = (float)*iz * l;
      };
      auto accept = [&](uint32_t b, float x, float y, float z) {
        const float r = B.h_rad[b];
        V.add(x, y, z, r, B.h_nin[b]);
        if (r / radius > 2.0f)  // exodus_finder.c:2321-2328
          mismatched++;
        is_taken[b] = 1;
        taken.push_back(b);
        accepted++;
      };
      for (uint32_t b = 0; b < B.count; b++) {
        const PreProbe P = pre[b];
        const uint32_t n_eff = B.h_cnt[b];
        const uint64_t eff0 = eff_at;
        eff_at += n_eff;
        if (P.node) {
          count_refusal(1);
          continue;
        }
        // No listed effect from a candidate accepted so far: the verdict
        // found for every candidate at once stands (the tests below, every
        // effect empty, give it).
        if (P.verdict != COMMIT_IN_ORDER && n_eff < BATCH_EFF_FULL &&
            eff0 + n_eff <= B.eff_capacity) {
          const PairEffect* eff = B.h_eff + eff0;
          bool touched = false;
          for (uint32_t e = 0; e < n_eff && !touched; e++)
            touched = is_taken[eff[e].a] != 0;
          if (!touched) {
            switch (P.verdict) {
            case COMMIT_PROXY: count_refusal(2); break;
            case COMMIT_OVERLAP: count_refusal(3); break;
            case COMMIT_RESCALE_FAIL: S.rescale_fail++; break;
            case COMMIT_GROWN_PROXY: S.proxy++; break;
            case COMMIT_GROWN_OVERLAP: S.overlap++; break;
            default: {  // COMMIT_ACCEPT
              uint32_t ix, iy, iz;
              float x, y, z;
              centre_of(b, &ix, &iy, &iz, &x, &y, &z);
              accept(b, x, y, z);
            }
            }
            continue;
          }
        }
        uint32_t ix, iy, iz;
        float x, y, z;
        centre_of(b, &ix, &iy, &iz, &x, &y, &z);
        const float r = B.h_rad[b];
        const int32_t pole = r >= 0.0f ? (int32_t)(scale * r * (1.0f - R.f) / l) - 1 : 0;
        // the effects of this commit's voids: listed, or probed when too many
        RungProbe c1, c2;
        if (n_eff < BATCH_EFF_FULL && eff0 + n_eff <= B.eff_capacity) {
          const PairEffect* eff = B.h_eff + eff0;
          for (uint32_t e = 0; e < n_eff; e++) {
            const PairEffect E = eff[e];
            if (!is_taken[E.a])
              continue;
            c1.node |= E.node;
            c1.pole |= E.pole;
            c1.state = combine(c1.state, E.state);
            c2.pole |= E.pole_r;
            c2.state = combine(c2.state, E.state_r);
          }
        } else {
          for (; cl_filled < taken.size(); cl_filled++) {
            const uint32_t a = taken[cl_filled];
            uint32_t jx, jy, jz;
            float ax, ay, az;
            centre_of(a, &jx, &jy, &jz, &ax, &ay, &az);
            CL.add(ax, ay, az, B.h_rad[a]);
          }
          const bool pole_refuses = proxy > 0 && P.pole;
          OverlapQuery q;
          overlap_query(q, x, y, z, radius, V.max_r, R.box, R.f, scale, V.inv_l);
          c1 = CL.probe(pole_refuses || P.state == 1 ? nullptr : &q, ix, iy, iz, x, y, z, true,
            pole_refuses ? 0 : proxy, scale, l, n);
          if (!c1.node && !(proxy > 0 && (P.pole || c1.pole)) && r >= 0.0f) {
            const uint32_t gs = B.h_state[b];
            OverlapQuery q2;
            overlap_query(q2, x, y, z, r, V.max_r, R.box, R.f, scale, V.inv_l);
            c2 = CL.probe((gs & 3u) == 1u || P.state_r == 1 ? nullptr : &q2, ix, iy, iz, x, y,
              z, false, pole, scale, l, n);
          }
        }
        // the rung-radius tests (mask, pole probe, overlap)
        if (c1.node) {
          count_refusal(1);
          continue;
        }
        if (proxy > 0 && (P.pole || c1.pole)) {
          count_refusal(2);
          continue;
        }
        if (exact_overlap(x, y, z, radius, B.edge[b] ? 2u : 0u, combine(P.state, c1.state))) {
          count_refusal(3);
          continue;
        }
        if (r < 0.0f) {
          S.rescale_fail++;
          continue;
        }
        // the tests at the rescaled radius (exodus_finder.c:2290-2311)
        const uint32_t gs = B.h_state[b];
        if (pole > 0 && ((gs & GROWN_POLE_BIT) || P.pole_r || c2.pole)) {
          S.proxy++;
          continue;
        }
        if (exact_overlap(x, y, z, r, gs & 3u, combine(P.state_r, c2.state))) {
          S.overlap++;
          continue;
        }
        accept(b, x, y, z);
      }
      // the rung's list, for the next commits' probes and the mask
      RV.add_many(
        taken.size(),
        [&](uint64_t i, float* x, float* y, float* z, float* r) {
          const uint32_t a = taken[i];
          uint32_t jx, jy, jz;
          centre_of(a, &jx, &jy, &jz, x, y, z);
          *r = B.h_rad[a];
        },
        [&](uint64_t m, const std::function<void(uint64_t, uint64_t)>& fn) {
          host_pool().run(m, fn);
        });
      S.accepted += accepted;
      *accepted_out = accepted;
      if (timing_active()) {
        const double t4 = now_ms();
        timing_add(SIF_AMD_STAGE_COMMIT_INDEX, t_idx - t2);
        timing_add(SIF_AMD_STAGE_COMMIT, t3 - t_idx);
        timing_add(SIF_AMD_STAGE_COMMIT_ORDER, t4 - t3);
        if (B.algo) {
          double bytes = 0.0;
          for (uint32_t b = 0; b < B.count; b++)
            bytes += (double)B.h_algo[b];
          timing_bytes(SIF_AMD_STAGE_GROW, bytes);
        }
      }
      return SIF_OK;
    };

    BatchSlot* cur = &g_slots[0];
    BatchSlot* nxt = &g_slots[1];
    // The next rung's start goes ahead once this rung's last batch is
    // launched: its counting pass beside the next commit, its prefilter
    // beside the one after (rv_mark: the voids its prefilter sees).
    const bool ahead = i + 1 < R.radii.size();
    size_t rv_mark = 0;
    assemble(*cur);
    if (cur->count)
      st = launch(*cur);
    while (st == SIF_OK && cur->count) {
      nxt->count = 0;
      if (k < ns) {
        assemble(*nxt);
        if (nxt->count)
          st = launch(*nxt);
      }
      if (st == SIF_OK && ahead && k >= ns && rs.rung == i) {
        st = append_rung_voids();
        if (st == SIF_OK)
          st = start_scan(i + 1);
      }
      uint64_t accepted = 0;
      if (st == SIF_OK)
        st = commit(*cur, &accepted);
      if (st == SIF_OK && rs.rung == i + 1 && rs.step == 1) {
        st = finish_scan();
        if (st == SIF_OK)
          st = append_rung_voids();
        rv_mark = RV.cx.size();
        if (st == SIF_OK)
          st = start_filter();
      }
      if (st != SIF_OK)
        break;
      // exodus_finder.c:2331-2349: widen after a quiet batch, back off after
      // accepts
      if (accepted == 0) {
        cap += batch_floor;
        if (cap > batch_max)
          cap = batch_max;
      } else {
        cap /= (uint32_t)(accepted + 1);
        if (cap < batch_floor)
          cap = batch_floor;
      }
      std::swap(cur, nxt);
    }
    if (st == SIF_OK)
      st = all_reads_linked(V);
    // the rung's voids into the device mask, for the rungs below
    th = now_ms();
    if (st == SIF_OK)
      st = append_rung_voids();
    if (timing_active())
      timing_add(SIF_AMD_STAGE_HOST, now_ms() - th);
    // the next rung's start, if it went ahead, brought to its end
    const bool started = st == SIF_OK && rs.rung == i + 1 && rs.step == 2;
    if (started)
      st = finish_filter(true);
    S.masked -= masked_since < S.masked ? masked_since : S.masked;
    if (stats_out)
      stats_out[i] = S;
    if (st == SIF_OK && R.on_rung)
      R.on_rung((uint32_t)(i + 1), V.cx.data(), V.cy.data(), V.cz.data(), V.r.data(),
        (uint64_t)V.cx.size(), mismatched);
    // The next rung's newer voids: none, or, when its start went ahead, the
    // voids of this rung its prefilter did not see (already on the device).
    if (started && rv_mark < RV.cx.size()) {
      const std::vector<float> tx(RV.cx.begin() + rv_mark, RV.cx.end()),
        ty(RV.cy.begin() + rv_mark, RV.cy.end()), tz(RV.cz.begin() + rv_mark, RV.cz.end()),
        tr(RV.r.begin() + rv_mark, RV.r.end());
      RV.clear();
      RV.add_many(
        tx.size(),
        [&](uint64_t v, float* x, float* y, float* z, float* r) {
          *x = tx[v];
          *y = ty[v];
          *z = tz[v];
          *r = tr[v];
        },
        [&](uint64_t m, const std::function<void(uint64_t, uint64_t)>& fn) {
          host_pool().run(m, fn);
        });
    } else {
      RV.clear();
    }
    rv_appended = RV.cx.size();
  }

  // the auxiliary stream may still smooth a rung the loop left (or read work)
  (void)hipStreamSynchronize(aux_stream());
  if (R.restore && st == SIF_OK && !use_lists)  // listed rungs left the grid as it was
    st = fft_filter_c2r(g_dev.spec.as<float2>(), n, R.box, SIF_AMD_FILTER_NONE, 0.0f, 0.0f,
      0, 0, R.d_delta, g_dev.work.as<float2>());
  // a failed rung may leave a batch in flight: drain the stream either way
  const hipError_t drained = hipStreamSynchronize(stream());
  if (st != SIF_OK)
    return st;
  SIF_HIP(drained);
  if (stats_out) {  // the prefilter's refusals, counted on the device
    std::vector<uint64_t> pf(n_rungs * 3);
    SIF_HIP(hipMemcpy(pf.data(), g_dev.pf_counts.ptr, pf.size() * sizeof(uint64_t),
      hipMemcpyDeviceToHost));
    for (size_t r = 0; r < n_rungs; r++) {
      stats_out[r].masked += pf[3 * r] - pf_since[3 * r];
      stats_out[r].proxy += pf[3 * r + 1] - pf_since[3 * r + 1];
      stats_out[r].overlap += pf[3 * r + 2] - pf_since[3 * r + 2];
    }
  }

  const uint64_t nv = V.cx.size();
  sif_catalogue_t* cat = sif_catalogue_alloc(nv ? nv : 1);
  if (!cat) {
    error_set(SIF_ERR_ALLOC, "exodus: cannot allocate a catalogue of %llu voids",
      (unsigned long long)nv);
    return SIF_ERR_ALLOC;
  }
  for (uint64_t v = 0; v < nv; v++)
    sif_catalogue_append(cat, V.cx[v], V.cy[v], V.cz[v], V.r[v]);
  // exodus_finder.c:2477-2483: how much of each void the survey observed
  if (survey) {
    st = sif_catalogue_reserve_footprint(cat);
    if (st == SIF_OK)
      st = survey_footprint(R.d_occupied, n, R.box, cat->cx, cat->cy, cat->cz, cat->radii, nv,
        cat->footprint, cat->footprint_shell);
    if (st != SIF_OK) {
      sif_catalogue_free(cat);
      return st;
    }
  }
  if (n_in_out) {
    *n_in_out = (uint64_t*)malloc(sizeof(uint64_t) * (nv ? nv : 1));
    if (*n_in_out)
      memcpy(*n_in_out, V.n_in.data(), sizeof(uint64_t) * nv);
  }
  *out_cat = cat;
  return SIF_OK;
}

int exodus_batch_window(uint32_t* floor, uint32_t* ceiling) {
  *floor = g_batch_floor;
  *ceiling = g_batch_max;
  return SIF_OK;
}

int sorted_ladder(const sif_real* radii, uint32_t n_radii, std::vector<float>& out) {
  out.assign(radii, radii + n_radii);
  qsort(out.data(), out.size(), sizeof(float), cmp_desc);
  return SIF_OK;
}

// sif__finder_check_args (src/finder/utils.c:33-58)
int check_args(const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_real f) {
  if (!radii || n_radii == 0) {
    error_set(SIF_ERR_INVALID, "radii: none given");
    return SIF_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n_radii; i++)
    if (!(radii[i] > 0.0f) || !std::isfinite(radii[i])) {
      error_set(SIF_ERR_INVALID, "radii[%u] = %g: not a positive radius", i, (double)radii[i]);
      return SIF_ERR_INVALID;
    }
  if (!(threshold > -1.0f && threshold < 0.0f)) {
    error_set(SIF_ERR_INVALID, "threshold %g: not a density contrast in (-1, 0)",
      (double)threshold);
    return SIF_ERR_INVALID;
  }
  if (!(f >= 0.0f && f <= 1.0f)) {
    error_set(SIF_ERR_INVALID, "overlap_fraction %g: not in [0, 1]", (double)f);
    return SIF_ERR_INVALID;
  }
  return SIF_OK;
}

// exodus_finder.c:1112-1138: what this build can run. SIF_FINDER_SEEDS_MINIMA
// (minima on their nodes, on the rung's own field), SIF_FINDER_EXTEND_REACH
// and SIF_FINDER_DISPLACE run: the excursion-peak variant whole.
int check_options(sif_option opt) {
  const sif_option refine = opt & SIF__FINDER_REFINE_MASK;
  const char* missing = refine == SIF_FINDER_REFINE_GAUSSIAN    ? "SIF_FINDER_REFINE_GAUSSIAN"
                        : refine == SIF_FINDER_REFINE_TRACERS   ? "SIF_FINDER_REFINE_TRACERS"
                        : refine                                ? "SIF_FINDER_REFINE_HILL"
                                                                : nullptr;
  if (missing) {
    error_set(SIF_ERR_UNSUPPORTED,
      "options: %s is not in this build (seeds stay on their nodes)", missing);
    return SIF_ERR_UNSUPPORTED;
  }
  return SIF_OK;
}

namespace {

enum { WEIGHT_NEGATIVE = 0, WEIGHT_NONFINITE = 1 };

// exodus_finder.c:1039-1060: a weight by its bits -- exponent all ones is not
// finite; the sign bit on anything but a zero is negative (-0 passes).
SIF_HD inline int weight_kind(uint32_t b) {
  if (((b >> 23) & 0xffu) == 0xffu)
    return WEIGHT_NONFINITE;
  return ((b >> 31) && (b << 1)) ? WEIGHT_NEGATIVE : -1;
}

__global__ void weight_check_kernel(const uint32_t* __restrict__ w, uint64_t n,
  unsigned long long* __restrict__ counts) {
  unsigned long long neg = 0, bad = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int k = weight_kind(w[i]);
    neg += k == WEIGHT_NEGATIVE;
    bad += k == WEIGHT_NONFINITE;
  }
  for (int off = 32; off > 0; off >>= 1) {
    neg += __shfl_xor(neg, off, 64);
    bad += __shfl_xor(bad, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && (neg || bad)) {
    atomicAdd(&counts[0], neg);
    atomicAdd(&counts[1], bad);
  }
}

int weights_verdict(uint64_t negative, uint64_t nonfinite, uint64_t n, double total) {
  if (negative || nonfinite) {
    error_set(SIF_ERR_INVALID,
      "mesh: weights must be finite and non-negative (%llu negative, %llu not finite, of "
      "%llu)", (unsigned long long)negative, (unsigned long long)nonfinite,
      (unsigned long long)n);
    return SIF_ERR_INVALID;
  }
  if (!(total > 0.0)) {
    error_set(SIF_ERR_INVALID, "mesh: weights sum to zero (no mean density)");
    return SIF_ERR_INVALID;
  }
  return SIF_OK;
}

thread_local DeviceBuffer g_weight_counts;

}  // namespace

int check_weights_host(const float* w, uint64_t n, const float* cell_w, double total) {
  if (!cell_w) {
    error_set(SIF_ERR_INVALID, "mesh: weights but no per-cell weight table");
    return SIF_ERR_INVALID;
  }
  uint64_t neg = 0, bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint32_t b;
    memcpy(&b, w + i, 4);
    const int k = weight_kind(b);
    neg += k == WEIGHT_NEGATIVE;
    bad += k == WEIGHT_NONFINITE;
  }
  return weights_verdict(neg, bad, n, total);
}

int check_weights_device(const float* d_w, uint64_t n, const float* d_cell_w, double total) {
  if (!d_cell_w) {
    error_set(SIF_ERR_INVALID, "mesh: weights but no per-cell weight table");
    return SIF_ERR_INVALID;
  }
  int st = g_weight_counts.reserve(2 * sizeof(uint64_t));
  if (st != SIF_OK)
    return st;
  uint64_t counts[2] = {0, 0};
  SIF_HIP(hipMemcpyAsync(g_weight_counts.ptr, counts, sizeof(counts), hipMemcpyHostToDevice,
    stream()));
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
  if (n)
    weight_check_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(
      reinterpret_cast<const uint32_t*>(d_w), n,
      reinterpret_cast<unsigned long long*>(g_weight_counts.ptr));
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(counts, g_weight_counts.ptr, sizeof(counts), hipMemcpyDeviceToHost,
    stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  return weights_verdict(counts[0], counts[1], n, total);
}

// exodus_finder.c:1114-1137: a survey runs SIF_FINDER_SEEDS_CELLS on their
// nodes with the rules of 0.4, for now; the rest is refused as invalid.
int survey_check_options(sif_option opt) {
  const sif_option refine = opt & SIF__FINDER_REFINE_MASK;
  const char* box_only = (opt & SIF_FINDER_SEEDS_MINIMA)          ? "SIF_FINDER_SEEDS_MINIMA"
                         : refine == SIF_FINDER_REFINE_GAUSSIAN   ? "SIF_FINDER_REFINE_GAUSSIAN"
                         : refine == SIF_FINDER_REFINE_TRACERS    ? "SIF_FINDER_REFINE_TRACERS"
                         : refine                                 ? "SIF_FINDER_REFINE_HILL"
                         : (opt & SIF_FINDER_EXTEND_REACH)        ? "SIF_FINDER_EXTEND_REACH"
                         : (opt & SIF_FINDER_DISPLACE)            ? "SIF_FINDER_DISPLACE"
                                                                  : nullptr;
  if (box_only) {
    error_set(SIF_ERR_INVALID,
      "options: %s: a survey (expected SIF_FINDER_SEEDS_CELLS on their nodes, for now)",
      box_only);
    return SIF_ERR_INVALID;
  }
  return check_options(opt);
}

namespace {

thread_local DeviceBuffer g_occupied;

// exodus_survey.c:367-385: every tracer at least `margin` from every face
// (a survey too close to its faces is invalid input, exodus_survey.c:520-523).
int survey_check_margin(const float* x, const float* y, const float* z, uint64_t n, float box,
  const char* what, float margin, const char* hint) {
  float lo[3], hi[3];
  int st = mesh_extent(x, y, z, n, lo, hi);
  if (st != SIF_OK)
    return st;
  for (int k = 0; k < 3; k++)
    if (lo[k] < margin || hi[k] > box - margin) {
      error_set(SIF_ERR_INVALID,
        "%s: span [%g, %g] on axis %d of a box of %g, %g from its face (needs %g of padding "
        "on every side%s)", what, (double)lo[k], (double)hi[k], k, (double)box,
        (double)fminf(lo[k], box - hi[k]), (double)margin, hint);
      return SIF_ERR_INVALID;
    }
  return SIF_OK;
}

}  // namespace

// exodus_survey.c:417-535 (sif_finder_exodus_survey_raw, from the checks on
// the meshes on): the meshes' weights and totals, the padding, and the
// survey's own inputs to R -- the observed nodes, alpha, the randoms.
int survey_prepare(RunSpec& R, const sif_amd_mesh_t& data, const sif_amd_mesh_t& randoms,
  const float* d_randoms, const sif_real* radii, uint32_t n_radii, sif_option opt) {
  int st = SIF_OK;
  if (data.weights)
    st = check_weights_device(data.weights, data.n_particles, data.cell_weights,
      data.total_weight);
  if (st == SIF_OK && randoms.weights)
    st = check_weights_device(randoms.weights, randoms.n_particles, randoms.cell_weights,
      randoms.total_weight);
  if (st != SIF_OK)
    return st;
  const double wd = data.weights ? data.total_weight : (double)data.n_particles;
  const double wr = randoms.weights ? randoms.total_weight : (double)randoms.n_particles;
  if (!(wd > 0.0) || !(wr > 0.0)) {
    error_set(SIF_ERR_INVALID, "%s mesh: total weight is zero", wd > 0.0 ? "random" : "data");
    return SIF_ERR_INVALID;
  }
  // exodus_survey.c:511-535: r_search + cell (or more), the hint naming the
  // likeliest reason it falls short
  const float cell = R.box / (float)R.n;
  float r_max = radii[0], r_min = radii[0];
  for (uint32_t i = 1; i < n_radii; i++) {
    r_max = radii[i] > r_max ? radii[i] : r_max;
    r_min = radii[i] < r_min ? radii[i] : r_min;
  }
  const float sf = sif__finder_search_factor(opt);
  const float r_search = r_max + fmaxf((sf - 1.0f) * r_max, 3.0f * cell);
  const float mesh_cell =
    fmaxf(R.box / (float)data.n_cells, R.box / (float)randoms.n_cells);
  const float margin = fmaxf(r_search, 0.5f * (r_search + mesh_cell)) + cell;
  char hint[192] = ", as survey_box sizes it";
  if (cell > r_min)
    snprintf(hint, sizeof(hint),
      "; survey_box allows for grid cells up to the smallest radius, %g, and these are %g",
      (double)r_min, (double)cell);
  st = survey_check_margin(randoms.x, randoms.y, randoms.z, randoms.n_particles, R.box,
    "randoms", margin, hint);
  if (st == SIF_OK)
    st = survey_check_margin(data.x, data.y, data.z, data.n_particles, R.box, "data", margin,
      hint);
  if (st != SIF_OK)
    return st;
  const uint64_t total = (uint64_t)R.n * R.n * R.n;
  if ((st = g_occupied.reserve(total)) != SIF_OK)
    return st;
  int32_t k = 0;
  st = survey_occupancy(randoms.x, randoms.y, randoms.z, randoms.n_particles, R.n, R.box,
    g_occupied.as<uint8_t>(), &k, nullptr, nullptr, nullptr);
  if (st != SIF_OK)
    return st;
  R.d_random = d_randoms;
  R.rmesh.offsets = randoms.offsets;
  R.rmesh.x = randoms.x;
  R.rmesh.y = randoms.y;
  R.rmesh.z = randoms.z;
  R.rmesh.w = randoms.weights;
  R.rmesh.cw = randoms.cell_weights;
  R.rmesh.mesh_n = randoms.n_cells;
  R.d_occupied = g_occupied.as<uint8_t>();
  R.alpha = (float)(wd / wr);  // exodus_finder.c:1651-1652
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

namespace {

sif_catalogue_t* exodus_device(float* d_delta, uint32_t n_cells, float box,
  const uint64_t* d_offsets, const float* d_mx, const float* d_my, const float* d_mz,
  const float* d_mw, const float* d_cw, double total_weight, uint32_t mesh_cells,
  uint64_t n_particles, const sif_real* radii, uint32_t n_radii, sif_real threshold,
  sif_real overlap_fraction, sif_option options, int restore, uint64_t** n_in,
  sif_amd_rung_stats_t* stats) {
  int st = ensure_init();
  if (st == SIF_OK)
    st = check_args(radii, n_radii, threshold, overlap_fraction);
  if (st == SIF_OK && (!d_delta || !d_offsets || !d_mx || !d_my || !d_mz ||
                        n_particles == 0 || mesh_cells == 0 || !(box > 0.0f))) {
    error_set(SIF_ERR_INVALID, "exodus: grid and mesh arrays are required");
    st = SIF_ERR_INVALID;
  }
  if (st == SIF_OK && d_mw)
    st = check_weights_device(d_mw, n_particles, d_cw, total_weight);
  if (st == SIF_OK)
    st = check_options(options);
  if (st != SIF_OK)
    return nullptr;
  RunSpec R;
  R.d_delta = d_delta;
  R.n = n_cells;
  R.box = box;
  R.d_offsets = d_offsets;
  R.d_mx = d_mx;
  R.d_my = d_my;
  R.d_mz = d_mz;
  R.d_mw = d_mw;
  R.d_cw = d_cw;
  R.total_weight = d_mw ? total_weight : (double)n_particles;
  R.mesh_n = mesh_cells;
  R.n_particles = n_particles;
  sorted_ladder(radii, n_radii, R.radii);
  R.threshold = threshold;
  R.f = overlap_fraction;
  R.opt = options;
  R.restore = restore != 0;
  R.prior = nullptr;
  sif_catalogue_t* cat = nullptr;
  st = exodus_run(R, &cat, n_in, stats);
  return st == SIF_OK ? cat : nullptr;
}

}  // namespace

extern "C" {

sif_catalogue_t* sif_amd_exodus_survey_device(float* d_data, const float* d_randoms,
  uint32_t n_cells, float box, const sif_amd_mesh_t* data, const sif_amd_mesh_t* randoms,
  const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_real overlap_fraction,
  sif_option options, int restore, uint64_t** n_in, sif_amd_rung_stats_t* stats) {
  int st = ensure_init();
  if (st != SIF_OK)
    return nullptr;
  const char* missing = !d_data ? "data grid: NULL"
                        : !d_randoms ? "random grid: NULL"
                        : !data ? "data mesh: NULL"
                        : !randoms ? "random mesh: NULL"
                        : !data->x || data->n_particles == 0 ? "data mesh: no particles"
                        : !randoms->x || randoms->n_particles == 0 ? "random mesh: no particles"
                                                                   : nullptr;
  if (missing) {
    error_set(SIF_ERR_INVALID, "%s", missing);
    return nullptr;
  }
  if (n_cells == 0 || !(box > 0.0f) || !data->offsets || !randoms->offsets ||
      data->n_cells == 0 || randoms->n_cells == 0 || !data->y || !data->z || !randoms->y ||
      !randoms->z) {
    error_set(SIF_ERR_INVALID, "sif_amd_exodus_survey_device: invalid arguments");
    return nullptr;
  }
  if ((st = check_args(radii, n_radii, threshold, overlap_fraction)) != SIF_OK ||
      (st = survey_check_options(options)) != SIF_OK)
    return nullptr;
  RunSpec R;
  R.d_delta = d_data;
  R.n = n_cells;
  R.box = box;
  R.d_offsets = data->offsets;
  R.d_mx = data->x;
  R.d_my = data->y;
  R.d_mz = data->z;
  R.d_mw = data->weights;
  R.d_cw = data->weights ? data->cell_weights : nullptr;
  R.total_weight = data->weights ? data->total_weight : (double)data->n_particles;
  R.mesh_n = data->n_cells;
  R.n_particles = data->n_particles;
  sorted_ladder(radii, n_radii, R.radii);
  R.threshold = threshold;
  R.f = overlap_fraction;
  R.opt = options;
  R.restore = restore != 0;
  R.prior = nullptr;
  if ((st = survey_prepare(R, *data, *randoms, d_randoms, radii, n_radii, options)) != SIF_OK)
    return nullptr;
  sif_catalogue_t* cat = nullptr;
  st = exodus_run(R, &cat, n_in, stats);
  return st == SIF_OK ? cat : nullptr;
}

sif_catalogue_t* sif_amd_exodus_device_weighted(float* d_delta, uint32_t n_cells, float box,
  const uint64_t* d_offsets, const float* d_mx, const float* d_my, const float* d_mz,
  const float* d_mw, const float* d_cell_weights, double total_weight, uint32_t mesh_cells,
  uint64_t n_particles, const sif_real* radii, uint32_t n_radii, sif_real threshold,
  sif_real overlap_fraction, sif_option options, int restore, uint64_t** n_in,
  sif_amd_rung_stats_t* stats) {
  if (!d_mw) {
    error_set(SIF_ERR_INVALID, "exodus: weighted: no weights");
    return nullptr;
  }
  return exodus_device(d_delta, n_cells, box, d_offsets, d_mx, d_my, d_mz, d_mw,
    d_cell_weights, total_weight, mesh_cells, n_particles, radii, n_radii, threshold,
    overlap_fraction, options, restore, n_in, stats);
}

sif_catalogue_t* sif_amd_exodus_device(float* d_delta, uint32_t n_cells, float box,
  const uint64_t* d_offsets, const float* d_mx, const float* d_my, const float* d_mz,
  uint32_t mesh_cells, uint64_t n_particles, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_real overlap_fraction, sif_option options, int restore,
  uint64_t** n_in, sif_amd_rung_stats_t* stats) {
  return exodus_device(d_delta, n_cells, box, d_offsets, d_mx, d_my, d_mz, nullptr, nullptr,
    0.0, mesh_cells, n_particles, radii, n_radii, threshold, overlap_fraction, options,
    restore, n_in, stats);
}

int sif_amd_set_void_room(uint32_t voids) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  DeviceVoids& D = g_voids;
  for (DeviceBuffer* b : {&D.x, &D.y, &D.z, &D.r, &D.start, &D.packed, &D.occupied})
    b->release();
  D.n = 0;
  D.cap = 0;
  D.gen++;
  g_void_room = voids;
  return SIF_OK;
}

int sif_amd_set_cell_lists(int on) {
  g_cell_lists = on != 0;
  return SIF_OK;
}

int sif_amd_set_batch_window(uint32_t floor, uint32_t ceiling) {
  if (floor == 0 && ceiling == 0) {
    g_batch_floor = BATCH_FLOOR_DEFAULT;
    g_batch_max = BATCH_ROOM;
    return SIF_OK;
  }
  if (floor < 1 || floor > ceiling || ceiling > BATCH_ROOM) {
    error_set(SIF_ERR_INVALID, "batch window: floor %u, ceiling %u (need 1 <= floor <= "
      "ceiling <= %u)", floor, ceiling, BATCH_ROOM);
    return SIF_ERR_INVALID;
  }
  g_batch_floor = floor;
  g_batch_max = ceiling;
  return SIF_OK;
}

int sif_amd_exodus_fields(const float* d_delta, uint32_t n_cells, float box,
  const sif_real* radii, uint32_t n_radii, sif_option options, float* d_fields) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  std::vector<float> ladder;
  sorted_ladder(radii, n_radii, ladder);
  st = exodus_prepare_spectrum(d_delta, n_cells);
  const uint64_t total = (uint64_t)n_cells * n_cells * n_cells;
  for (uint32_t i = 0; i < n_radii && st == SIF_OK; i++)
    st = exodus_smooth(n_cells, box, ladder[i], options, d_fields + i * total);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
