// The batched form: many independent boxes of uneven size, each a whole
// exodus catalogue (CIC deposit, density contrast, chain mesh, every rung of
// its own ladder), run side by side on one GPU.
//
// A box's catalogue is sequential in its candidates, so boxes, not
// candidates, are the unit of parallelism here: a pool of host workers,
// each with a stream pair and device buffers of its own (the library's state
// is per host thread), takes boxes from one queue, largest first. The
// kernels of different boxes overlap on the device and their commits on the
// host, so the fixed costs of each rung (launches, small kernels, the waits
// between the device and the commit) are paid once per round of boxes
// rather than once per box.
//
// The tunings -- boxes side by side, the CIC tile edge, the growth batch
// window and the wave count a small growth launch splits towards -- change
// no bit of any catalogue: each box's catalogue is that of its plain run.
// sif_amd_batch_autotune() times a few of them on a probe of the boxes and
// returns the fastest.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "internal.h"
#include "sif/utils/logger.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {

constexpr uint32_t MAX_WORKERS = 64;

// One box's buffers on a worker: the grid and its CIC accumulators, the mesh.
struct BoxBuffers {
  DeviceBuffer grid, acc, offsets, mx, my, mz;
};
thread_local BoxBuffers t_box;

int run_box(const sif_amd_box_t& B, sif_real threshold, sif_real f, sif_option opt,
  sif_catalogue_t** out, uint64_t* n_voids) {
  const uint32_t N = B.n_cells;
  const uint64_t total = (uint64_t)N * N * N;
  const uint64_t cells = (uint64_t)B.mesh_cells * B.mesh_cells * B.mesh_cells;
  int st = t_box.grid.reserve(total * sizeof(float));
  if (st == SIF_OK) st = t_box.acc.reserve(total * sizeof(int64_t));
  if (st == SIF_OK) st = t_box.offsets.reserve((cells + 1) * sizeof(uint64_t));
  if (st == SIF_OK) st = t_box.mx.reserve((B.n_particles ? B.n_particles : 1) * sizeof(float));
  if (st == SIF_OK) st = t_box.my.reserve((B.n_particles ? B.n_particles : 1) * sizeof(float));
  if (st == SIF_OK) st = t_box.mz.reserve((B.n_particles ? B.n_particles : 1) * sizeof(float));
  if (st != SIF_OK)
    return st;
  float* grid = t_box.grid.as<float>();
  st = cic_deposit(B.d_x, B.d_y, B.d_z, nullptr, B.n_particles, N, B.box, grid,
    t_box.acc.as<int64_t>());
  double mean = 0.0;
  if (st == SIF_OK)
    st = grid_mean(grid, total, &mean);
  if (st == SIF_OK && !(mean > 0.0)) {
    error_set(SIF_ERR_RANGE, "batch: a box's mean density is %g", mean);
    st = SIF_ERR_RANGE;
  }
  if (st == SIF_OK)
    st = density_contrast(grid, total, mean);
  double tw = 0.0;
  if (st == SIF_OK)
    st = mesh_build(B.d_x, B.d_y, B.d_z, nullptr, B.n_particles, B.mesh_cells, B.box,
      t_box.offsets.as<uint64_t>(), t_box.mx.as<float>(), t_box.my.as<float>(),
      t_box.mz.as<float>(), nullptr, nullptr, nullptr, &tw);
  if (st != SIF_OK)
    return st;
  RunSpec R;
  R.d_delta = grid;
  R.n = N;
  R.box = B.box;
  R.d_offsets = t_box.offsets.as<uint64_t>();
  R.d_mx = t_box.mx.as<float>();
  R.d_my = t_box.my.as<float>();
  R.d_mz = t_box.mz.as<float>();
  R.total_weight = (double)B.n_particles;
  R.mesh_n = B.mesh_cells;
  R.n_particles = B.n_particles;
  sorted_ladder(B.radii, B.n_radii, R.radii);
  R.threshold = threshold;
  R.f = f;
  R.opt = opt;
  R.restore = false;
  R.prior = nullptr;
  sif_catalogue_t* cat = nullptr;
  st = exodus_run(R, &cat, nullptr, nullptr);
  if (st != SIF_OK)
    return st;
  *out = cat;
  *n_voids = cat->n_voids;
  return SIF_OK;
}

// The persistent workers. A job is a list of boxes; the first `active`
// workers take boxes from it until it is empty.
class BatchPool {
 public:
  struct Job {
    const sif_amd_box_t* boxes = nullptr;
    const uint32_t* order = nullptr;
    uint32_t n = 0;
    sif_real threshold = 0, f = 0;
    sif_option opt = 0;
    sif_catalogue_t** out = nullptr;
    uint64_t* n_voids = nullptr;
    std::atomic<uint32_t> next{0};
    std::atomic<int> status{SIF_OK};
    std::mutex err_mu;
    char message[512] = "";
  };

  int run(Job& job, uint32_t active) {
    int st = grow(active);
    if (st != SIF_OK)
      return st;
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &job;
      active_ = active;
      done_ = 0;
      gen_++;
    }
    cv_.notify_all();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&]() { return done_ == active_; });
    job_ = nullptr;
    return SIF_OK;
  }

  ~BatchPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : threads_)
      t.join();
  }

 private:
  int grow(uint32_t want) {
    while (threads_.size() < want) {
      int dev = 0;
      SIF_HIP(hipGetDevice(&dev));
      const uint32_t index = (uint32_t)threads_.size();
      std::atomic<int> ready{0};
      threads_.emplace_back([this, dev, index, &ready]() { loop(dev, index, &ready); });
      while (ready.load() == 0)
        std::this_thread::yield();
      if (ready.load() < 0) {
        error_set(SIF_ERR_ALLOC, "batch: a worker could not create its streams");
        return SIF_ERR_ALLOC;
      }
    }
    return SIF_OK;
  }

  void loop(int dev, uint32_t index, std::atomic<int>* ready) {
    hipStream_t s = nullptr, aux = nullptr;
    const bool ok = hipSetDevice(dev) == hipSuccess && hipStreamCreate(&s) == hipSuccess &&
                    hipStreamCreate(&aux) == hipSuccess;
    ready->store(ok ? 1 : -1);
    if (!ok)
      return;
    set_thread_streams(s, aux);
    uint64_t seen = 0;
    for (;;) {
      Job* job = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
        if (stop_)
          break;
        seen = gen_;
        if (index < active_)
          job = job_;
      }
      if (!job)
        continue;
      for (;;) {
        const uint32_t k = job->next.fetch_add(1);
        if (k >= job->n || job->status.load() != SIF_OK)
          break;
        const uint32_t b = job->order[k];
        int st = run_box(job->boxes[b], job->threshold, job->f, job->opt, &job->out[b],
          &job->n_voids[b]);
        if (st == SIF_OK)
          st = hip_check(hipStreamSynchronize(s), "batch worker sync");
        if (st != SIF_OK) {
          std::lock_guard<std::mutex> lk(job->err_mu);
          if (job->status.load() == SIF_OK) {
            snprintf(job->message, sizeof(job->message), "box %u: %s", b,
              sif_error_message());
            job->status.store(st);
          }
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        done_++;
      }
      done_cv_.notify_all();
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamSynchronize(aux);
  }

  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  Job* job_ = nullptr;
  uint32_t active_ = 0, done_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Held until the process ends, its workers blocked between jobs: the pool is
// never destroyed, so no worker frees device memory while the runtime is
// shutting down.
BatchPool& batch_pool() {
  static BatchPool* pool = new BatchPool;
  return *pool;
}
std::mutex g_batch_mutex;  // one batch at a time: the tunings are process-wide

// The tunings for the call's duration.
struct TuningScope {
  uint32_t tile, split, floor, ceiling;
  int st = SIF_OK;
  explicit TuningScope(const sif_amd_tuning_t& T) {
    tile = cic_tile_edge();
    split = grow_split_waves();
    exodus_batch_window(&floor, &ceiling);
    if (T.cic_tile && (st = set_cic_tile_edge(T.cic_tile)) != SIF_OK)
      return;
    if (T.split_waves && (st = set_grow_split_waves(T.split_waves)) != SIF_OK)
      return;
    if (T.batch_floor || T.batch_ceiling)
      st = sif_amd_set_batch_window(T.batch_floor, T.batch_ceiling);
  }
  ~TuningScope() {
    (void)set_cic_tile_edge(tile);
    (void)set_grow_split_waves(split);
    (void)sif_amd_set_batch_window(floor, ceiling);
  }
};

int check_boxes(const sif_amd_box_t* boxes, uint32_t n_boxes, sif_real threshold,
  sif_real f, sif_option opt) {
  if (!boxes && n_boxes) {
    error_set(SIF_ERR_INVALID, "batch: no boxes");
    return SIF_ERR_INVALID;
  }
  int st = check_options(opt);
  if (st != SIF_OK)
    return st;
  for (uint32_t b = 0; b < n_boxes; b++) {
    const sif_amd_box_t& B = boxes[b];
    st = check_args(B.radii, B.n_radii, threshold, f);
    if (st == SIF_OK && (!B.d_x || !B.d_y || !B.d_z || B.n_particles == 0 || B.n_cells < 2 ||
                          B.mesh_cells == 0 || B.mesh_cells > 1024 || !(B.box > 0.0f))) {
      error_set(SIF_ERR_INVALID, "batch: box %u: particles, grid, mesh and box are required",
        b);
      st = SIF_ERR_INVALID;
    }
    if (st != SIF_OK)
      return st;
  }
  return SIF_OK;
}

int run_batch(const sif_amd_box_t* boxes, uint32_t n_boxes, sif_real threshold, sif_real f,
  sif_option opt, const sif_amd_tuning_t& T, sif_catalogue_t** out, uint64_t* n_voids) {
  const uint32_t workers = std::max(1u, std::min(T.workers ? T.workers : 8u, MAX_WORKERS));
  TuningScope scope(T);
  if (scope.st != SIF_OK)
    return scope.st;
  // largest first, so that the last boxes to start are short
  std::vector<uint32_t> order(n_boxes);
  for (uint32_t b = 0; b < n_boxes; b++)
    order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return boxes[a].n_particles > boxes[b].n_particles;
  });
  for (uint32_t b = 0; b < n_boxes; b++)
    out[b] = nullptr;
  BatchPool::Job job;
  job.boxes = boxes;
  job.order = order.data();
  job.n = n_boxes;
  job.threshold = threshold;
  job.f = f;
  job.opt = opt;
  job.out = out;
  job.n_voids = n_voids;
  int st = batch_pool().run(job, std::min(workers, std::max(n_boxes, 1u)));
  if (st == SIF_OK)
    st = job.status.load();
  if (st != SIF_OK) {
    if (job.message[0])
      error_set(st, "batch: %s", job.message);
    for (uint32_t b = 0; b < n_boxes; b++)
      if (out[b]) {
        sif_catalogue_free(out[b]);
        out[b] = nullptr;
      }
  }
  return st;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
    .count();
}

bool same_catalogue(const sif_catalogue_t* a, const sif_catalogue_t* b) {
  if (a->n_voids != b->n_voids)
    return false;
  const size_t bytes = a->n_voids * sizeof(float);
  return !memcmp(a->cx, b->cx, bytes) && !memcmp(a->cy, b->cy, bytes) &&
         !memcmp(a->cz, b->cz, bytes) && !memcmp(a->radii, b->radii, bytes);
}

}  // namespace

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_exodus_batch(const sif_amd_box_t* boxes, uint32_t n_boxes, sif_real threshold,
  sif_real overlap_fraction, sif_option options, const sif_amd_tuning_t* tuning,
  sif_catalogue_t** catalogues) {
  int st = ensure_init();
  if (st == SIF_OK && !catalogues) {
    error_set(SIF_ERR_INVALID, "batch: no catalogue array");
    st = SIF_ERR_INVALID;
  }
  if (st == SIF_OK)
    st = check_boxes(boxes, n_boxes, threshold, overlap_fraction, options);
  if (st != SIF_OK)
    return st;
  std::lock_guard<std::mutex> lk(g_batch_mutex);
  sif_amd_tuning_t T;
  memset(&T, 0, sizeof(T));
  if (tuning)
    T = *tuning;
  std::vector<uint64_t> n_voids(n_boxes);
  return run_batch(boxes, n_boxes, threshold, overlap_fraction, options, T, catalogues,
    n_voids.data());
}

int sif_amd_batch_autotune(const sif_amd_box_t* boxes, uint32_t n_boxes, sif_real threshold,
  sif_real overlap_fraction, sif_option options, uint32_t probe, sif_amd_tuning_t* best) {
  int st = ensure_init();
  if (st == SIF_OK && !best) {
    error_set(SIF_ERR_INVALID, "batch autotune: no result");
    st = SIF_ERR_INVALID;
  }
  if (st == SIF_OK)
    st = check_boxes(boxes, n_boxes, threshold, overlap_fraction, options);
  if (st != SIF_OK)
    return st;
  std::lock_guard<std::mutex> lk(g_batch_mutex);
  // the probe: boxes spread over the sizes (every k-th of the size order)
  std::vector<uint32_t> order(n_boxes);
  for (uint32_t b = 0; b < n_boxes; b++)
    order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return boxes[a].n_particles > boxes[b].n_particles;
  });
  const uint32_t m = std::max(1u, std::min(probe ? probe : 8u, n_boxes));
  std::vector<sif_amd_box_t> sample;
  for (uint32_t k = 0; k < m; k++)
    sample.push_back(boxes[order[(uint64_t)k * n_boxes / m]]);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  std::vector<sif_catalogue_t*> ref(m, nullptr), cur(m, nullptr);
  std::vector<uint64_t> nv(m);
  auto timed = [&](const sif_amd_tuning_t& T, double* seconds) -> int {
    const double t0 = now_s();
    int e = run_batch(sample.data(), m, threshold, overlap_fraction, options, T, cur.data(),
      nv.data());
    *seconds = now_s() - t0;
    if (e != SIF_OK)
      return e;
    // every tuning must give the same catalogues
    for (uint32_t k = 0; k < m; k++) {
      if (!ref[k]) {
        ref[k] = cur[k];
        cur[k] = nullptr;
      } else {
        const bool same = same_catalogue(ref[k], cur[k]);
        sif_catalogue_free(cur[k]);
        cur[k] = nullptr;
        if (!same) {
          error_set(SIF_ERR_INVALID, "batch autotune: a tuning changed a catalogue");
          return SIF_ERR_INVALID;
        }
      }
    }
    return SIF_OK;
  };
  sif_amd_tuning_t T;
  memset(&T, 0, sizeof(T));
  T.workers = 1;
  T.cic_tile = 8;
  T.split_waves = 4096;
  double t_best = 0.0;
  st = timed(T, &t_best);  // also the warm-up of the first worker's buffers
  if (st == SIF_OK)
    st = timed(T, &t_best);
  sif_amd_tuning_t B = T;
  // one knob at a time: boxes side by side, the wave count a small growth
  // launch splits towards (the chip shared by that many), the CIC tile
  for (uint32_t w : {2u, 4u, 8u, 12u, 16u}) {
    if (st != SIF_OK || w > hw || w > MAX_WORKERS)
      break;
    sif_amd_tuning_t C = B;
    C.workers = w;
    double t;
    st = timed(C, &t);  // the new workers' buffers
    if (st == SIF_OK)
      st = timed(C, &t);
    if (st == SIF_OK && t < t_best) {
      t_best = t;
      B = C;
    }
  }
  for (uint32_t split : {4096u / std::max(1u, B.workers), 1024u, 4096u}) {
    if (st != SIF_OK)
      break;
    sif_amd_tuning_t C = B;
    C.split_waves = std::max(64u, split);
    if (C.split_waves == B.split_waves)
      continue;
    double t;
    st = timed(C, &t);
    if (st == SIF_OK && t < t_best) {
      t_best = t;
      B = C;
    }
  }
  if (st == SIF_OK) {
    sif_amd_tuning_t C = B;
    C.cic_tile = 16;
    double t;
    st = timed(C, &t);
    if (st == SIF_OK && t < t_best) {
      t_best = t;
      B = C;
    }
  }
  for (sif_catalogue_t* c : ref)
    if (c)
      sif_catalogue_free(c);
  if (st != SIF_OK)
    return st;
  *best = B;
  return SIF_OK;
}

}  // extern "C"
