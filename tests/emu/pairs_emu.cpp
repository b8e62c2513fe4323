// TEST INFRASTRUCTURE ONLY -- runs sif_amd/csrc/batch_pairs_device.h's
// batch_pairs_kernel on the CPU under wave_emu.h, beside the host's
// pair_effects() (batch_pairs.h) on the same batch, for the CPU test that the
// device lists each pair's effects as the host does (tests/test_batch_pairs.py).
// Effects come out packed: a * 256 + node + 2 pole + 4 pole_r + 8 state +
// 32 state_r. Built by tests/emu/Makefile.
#include <algorithm>
#include <thread>
#include <vector>

#include "wave_emu.h"

thread_local emu_dim3 threadIdx;
thread_local emu_dim3 blockIdx;
emu_dim3 blockDim, gridDim;

#include "batch_pairs_device.h"

static EmuWave g_waves[sif_amd::BP_WAVES];
static std::atomic<int> g_stuck{0};

EmuWave& emu_wave() { return g_waves[threadIdx.x >> 6]; }

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
      std::fprintf(stderr, "wave_emu: lanes of wave %u, block %u did not all reach %s\n",
        threadIdx.x >> 6, blockIdx.x, site);
      g_stuck = 1;
      std::_Exit(3);
    }
  }
}

using sif_amd::PairEffect;

static uint32_t pack(const PairEffect& e) {
  return e.a * 256u + (uint32_t)e.node + 2u * e.pole + 4u * e.pole_r + 8u * e.state +
         32u * e.state_r;
}

static sif_amd::BatchPairParams params(uint32_t n, float l, float box, float f, float scale,
  float radius, float cat_max_r, int32_t proxy) {
  sif_amd::BatchPairParams P;
  P.n = n;
  P.l = l;
  P.box = box;
  P.f = f;
  P.scale = scale;
  P.radius = radius;
  P.cat_max_r = cat_max_r;
  P.inv_l = (float)sif_amd::VOID_CELLS / box;
  P.proxy = proxy;
  return P;
}

template <bool WRITE>
static void run_pass(const uint32_t* flat, const float* rad, const uint32_t* gs,
  uint32_t count, const sif_amd::BatchPairParams& P, const float* bmax, const uint32_t* cells,
  uint32_t* cnt32,
  uint8_t* cnt8, const uint64_t* offset, PairEffect* eff, uint64_t capacity) {
  blockDim.x = 64 * sif_amd::BP_WAVES;
  const uint32_t blocks = (count + sif_amd::BP_WAVES - 1) / sif_amd::BP_WAVES;
  gridDim.x = blocks;
  for (uint32_t b = 0; b < blocks; b++) {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < blockDim.x; t++)
      th.emplace_back([&, t, b] {
        threadIdx.x = t;
        blockIdx.x = b;
        sif_amd::batch_pairs_kernel<WRITE>(flat, rad, gs, count, P, bmax, cells, cnt32, cnt8,
          offset, eff, capacity);
      });
    for (auto& x : th)
      x.join();
  }
}

// The device text: counts (capped at BATCH_EFF_FULL) and, for the candidates
// whose list fits, their effects packed into out[b * cap ...], in the device's
// order; n_out[b] the count, -1 for a candidate the commit would probe.
extern "C" int emu_batch_pairs(const uint32_t* flat, const float* rad, const uint32_t* gs,
  uint32_t count, uint32_t n, float l, float box, float f, float scale, float radius,
  float cat_max_r, int32_t proxy, uint64_t capacity, int32_t* n_out, uint32_t* out,
  int cap) {
  const sif_amd::BatchPairParams P = params(n, l, box, f, scale, radius, cat_max_r, proxy);
  float bmax = -1.0f;  // batch_prep_kernel's answers
  std::vector<uint32_t> cells(count);
  for (uint32_t i = 0; i < count; i++) {
    bmax = fmaxf(bmax, rad[i]);
    cells[i] = sif_amd::batch_cell_code(flat[i], rad[i], P);
  }
  std::vector<uint32_t> cnt32(count, 0xdeadbeefu);
  std::vector<uint8_t> cnt8(count, 0xaa);
  std::vector<uint64_t> offset(count);
  std::vector<PairEffect> eff(capacity + 1);
  run_pass<false>(flat, rad, gs, count, P, &bmax, cells.data(), cnt32.data(), cnt8.data(),
    nullptr, nullptr, 0);
  uint64_t at = 0;
  for (uint32_t b = 0; b < count; b++) {
    if (cnt32[b] != cnt8[b])
      return -1;
    offset[b] = at;
    at += cnt32[b];
  }
  run_pass<true>(flat, rad, gs, count, P, &bmax, cells.data(), nullptr, cnt8.data(),
    offset.data(), eff.data(), capacity);
  for (uint32_t b = 0; b < count; b++) {
    const uint32_t c = cnt8[b];
    if (c >= sif_amd::BATCH_EFF_FULL || offset[b] + c > capacity || (int)c > cap) {
      n_out[b] = -1;
      continue;
    }
    n_out[b] = (int32_t)c;
    for (uint32_t e = 0; e < c; e++)
      out[(uint64_t)b * cap + e] = pack(eff[offset[b] + e]);
  }
  return g_stuck.load();
}

// The host's pair_effects() with the tests the device lists (every one a
// candidate may still need, as far as the device can tell).
extern "C" int host_batch_pairs(const uint32_t* flat, const float* rad, const uint32_t* gs,
  uint32_t count, uint32_t n, float l, float box, float f, float scale, float radius,
  float cat_max_r, int32_t proxy, int32_t* n_out, uint32_t* out, int cap) {
  using namespace sif_amd;
  RungVoids BI;
  BI.init(box);
  BI.columns = false;
  std::vector<uint32_t> place;
  const uint32_t nn = n * n;
  for (uint32_t b = 0; b < count; b++)
    if (rad[b] >= 0.0f) {
      const uint32_t ix = flat[b] / nn, iy = (flat[b] / n) % n, iz = flat[b] % n;
      BI.add((float)ix * l, (float)iy * l, (float)iz * l, rad[b]);
      place.push_back(b);
    }
  std::vector<PairEffect> buf(cap);
  for (uint32_t b = 0; b < count; b++) {
    PairQuery q;
    q.b = b;
    q.ix = flat[b] / nn;
    q.iy = (flat[b] / n) % n;
    q.iz = flat[b] % n;
    q.x = (float)q.ix * l;
    q.y = (float)q.iy * l;
    q.z = (float)q.iz * l;
    q.r = rad[b];
    q.pole = proxy;
    q.pole_r = q.r >= 0.0f ? (int32_t)(scale * q.r * (1.0f - f) / l) - 1 : 0;
    q.need = NEED_NODE | NEED_STATE | (proxy > 0 ? NEED_POLE : 0);
    if (q.r >= 0.0f && q.pole_r > 0 && !(gs[b] & GROWN_POLE_BIT))
      q.need |= NEED_POLE_R;
    if (q.r >= 0.0f && (gs[b] & 3u) == 0u)
      q.need |= NEED_STATE_R;
    const int c = pair_effects(BI, place.data(), q, radius, cat_max_r, f, scale, l, n,
      buf.data(), cap);
    n_out[b] = c;
    for (int e = 0; e < c; e++)
      out[(uint64_t)b * cap + e] = pack(buf[e]);
  }
  return 0;
}
