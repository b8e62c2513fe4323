// TEST INFRASTRUCTURE ONLY -- runs sif_amd/csrc/grow_device.h's grow_kernel on
// the CPU under wave_emu.h, for the CPU test of the kernel's text against the
// oracle (tests/test_grow_emulated.py). Built by tests/emu/Makefile.
#include <thread>
#include <vector>

#include "wave_emu.h"

thread_local emu_dim3 threadIdx;
thread_local emu_dim3 blockIdx;
emu_dim3 blockDim, gridDim;

// the block's dynamic LDS (blocks run one at a time)
static std::vector<uint32_t> g_lds;
#define GROW_DYNAMIC_LDS(name) uint32_t* name = g_lds.data()

// a template entry's six fields from lane l: one rendezvous of the wave
#define GROW_ENTRY_FIELDS 1
namespace sif_amd {
struct EntrySlot {
  uint32_t cnt, start, type;
  float x, y, z;
};
static EntrySlot g_entry_slots[64];
inline void entry_fields(int l, uint32_t cnt, uint32_t start, uint32_t type, float ex,
  float ey, float ez, uint32_t* o_cnt, uint32_t* o_start, uint32_t* o_type, float* o_x,
  float* o_y, float* o_z) {
  g_entry_slots[emu_lane()] = EntrySlot{cnt, start, type, ex, ey, ez};
  emu_wave().barrier("entry_fields");
  const EntrySlot e = g_entry_slots[l & 63];
  emu_wave().barrier("entry_fields");
  *o_cnt = e.cnt;
  *o_start = e.start;
  *o_type = e.type;
  *o_x = e.x;
  *o_y = e.y;
  *o_z = e.z;
}
}  // namespace sif_amd

#include "grow_device.h"

static EmuWave g_waves[sif_amd::GROW_WAVES];
static std::atomic<int> g_stuck{0};

EmuWave& emu_wave() { return g_waves[threadIdx.x >> 6]; }

// A generation barrier the lanes spin on, yielding the core: the 64 lane
// threads outnumber the cores, and a rendezvous is far more frequent than a
// sleep and wake-up of every thread can afford.
void EmuWave::barrier(const char* site) {
  std::atomic<uint64_t>& gen_ref = generation;
  std::atomic<int>& arrived_ref = arrived;
  const uint64_t gen = gen_ref.load(std::memory_order_acquire);
  if (arrived_ref.fetch_add(1, std::memory_order_acq_rel) == 63) {
    arrived_ref.store(0, std::memory_order_relaxed);
    gen_ref.store(gen + 1, std::memory_order_release);
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spins = 1; gen_ref.load(std::memory_order_acquire) == gen; spins++) {
    std::this_thread::yield();
    if ((spins & 0xffffu) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
      std::fprintf(stderr,
        "wave_emu: lanes of wave %u, block %u did not all reach %s (%d of 64 arrived)\n",
        threadIdx.x >> 6, blockIdx.x, site, arrived_ref.load());
      g_stuck = 1;
      std::_Exit(3);
    }
  }
}

extern "C" int emu_find_radius(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, int32_t mesh_n, float box, const int32_t* tdx, const int32_t* tdy,
  const int32_t* tdz, const uint8_t* ttype, const float* tmin, const float* tmax,
  uint32_t tcount, float rmin, float r_search, float threshold, float vol_factor,
  const float* cx, const float* cy, const float* cz, uint32_t n, float* radius,
  uint32_t* reason, uint64_t* n_in, uint32_t cache_words, const float* mw,
  const float* cw, float mass_factor, uint32_t slices) {
  // cache_words: the annulus cache a wave holds; 0 sizes it as the library
  // does for the launch. mw non-NULL runs the weighted kernel (its cache is
  // always REFINE_CAP pairs).
  const uint32_t blocks = (n + sif_amd::GROW_WAVES - 1) / sif_amd::GROW_WAVES;
  std::vector<int4> cells(tcount);
  std::vector<float2> bounds(tcount);
  for (uint32_t i = 0; i < tcount; i++) {
    cells[i] = make_int4(tdx[i], tdy[i], tdz[i], ttype[i]);
    bounds[i] = make_float2(tmin[i], tmax[i]);
  }
  sif_amd::GrowParams P;
  P.algo_bytes = nullptr;
  P.offsets = offsets;
  P.mx = mx;
  P.my = my;
  P.mz = mz;
  P.mesh_n = mesh_n;
  P.box = box;
  P.inv_l = 1.0f / (box / (float)mesh_n);
  P.tcell = cells.data();
  P.tbound = bounds.data();
  P.tcount = tcount;
  P.rmin = rmin;
  P.r_search = r_search;
  P.threshold = threshold;
  P.vol_factor = vol_factor;
  P.cx = cx;
  P.cy = cy;
  P.cz = cz;
  P.n_cand = n;
  P.radius = radius;
  P.reason = reason;
  P.n_in = n_in;
  const bool weighted = mw != nullptr;
  P.mw = mw;
  P.cw = cw;
  P.mass_factor = weighted ? mass_factor : vol_factor;
  const uint32_t nb = sif_amd::grow_bin_count(r_search, rmin, vol_factor);
  if (weighted)
    sif_amd::grow_lds_words_weighted(nb, &P.lds_bins, &P.lds_cache);
  else
    sif_amd::grow_lds_words(nb, cache_words, &P.lds_bins, &P.lds_cache);
  // slices > 1 (unweighted): the first pass split over that many waves a
  // candidate, by grow_bin_kernel, before grow_kernel
  std::vector<uint32_t> part_bins;
  std::vector<uint64_t> part_counts;
  P.slices = 1;
  P.part_bins = nullptr;
  P.part_counts = nullptr;
  if (slices > 1 && !weighted) {
    part_bins.assign((size_t)n * slices * nb, 0xdeadbeefu);
    part_counts.assign((size_t)n * slices * sif_amd::PART_COUNTS, 0xdeadbeefdeadbeefull);
    P.slices = slices;
    P.part_bins = part_bins.data();
    P.part_counts = part_counts.data();
  }
  // whole words, 8-byte aligned for the weighted bins
  g_lds.assign((size_t)sif_amd::GROW_WAVES *
                   sif_amd::grow_wave_words(P.lds_bins, P.lds_cache, weighted) + 2,
    0xdeadbeefu);
  if (P.slices > 1) {
    blockDim.x = 64;
    gridDim.x = n * P.slices;
    for (uint32_t b = 0; b < n * P.slices; b++) {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < 64; t++)
        th.emplace_back([&, t, b] {
          threadIdx.x = t;
          blockIdx.x = b;
          sif_amd::grow_bin_kernel(P);
        });
      for (auto& x : th)
        x.join();
    }
  }
  blockDim.x = 64 * sif_amd::GROW_WAVES;
  gridDim.x = blocks;
  for (uint32_t b = 0; b < blocks; b++) {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < blockDim.x; t++)
      th.emplace_back([&, t, b] {
        threadIdx.x = t;
        blockIdx.x = b;
        if (weighted)
          sif_amd::grow_kernel<true>(P);
        else
          sif_amd::grow_kernel<false>(P);
      });
    for (auto& x : th)
      x.join();
  }
  return g_stuck.load();
}
