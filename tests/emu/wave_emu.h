// TEST INFRASTRUCTURE ONLY -- a CPU emulation of the HIP features the
// sphere-growth kernel (sif_amd/csrc/grow_device.h) uses, so its exact text
// can run without a GPU: one std::thread per lane, a block's threads alive
// together, every cross-lane operation (shuffle, ballot, wave barrier) a
// rendezvous of the wave's 64 threads. A lane that reaches a different
// sequence of cross-lane operations than its wave deadlocks the rendezvous,
// which the barrier reports and aborts on instead of hanging.
#pragma once
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

struct emu_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local emu_dim3 threadIdx;
extern thread_local emu_dim3 blockIdx;
extern emu_dim3 blockDim, gridDim;

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

struct int4 {
  int x, y, z, w;
};
struct float2 {
  float x, y;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
inline float2 make_float2(float a, float b) { return float2{a, b}; }

struct EmuWave {
  std::atomic<int> arrived{0};
  std::atomic<uint64_t> generation{0};
  uint64_t slot[64];
  const char* where = "";
  void barrier(const char* site);
};
EmuWave& emu_wave();

inline int emu_lane() { return (int)(threadIdx.x & 63); }

template <class T> inline uint64_t emu_bits(T v) {
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T> inline T emu_from(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}

template <class T> inline T emu_exchange(T v, int src, const char* site) {
  EmuWave& w = emu_wave();
  w.slot[emu_lane()] = emu_bits(v);
  w.barrier(site);
  const uint64_t r = w.slot[src & 63];
  w.barrier(site);
  return emu_from<T>(r);
}

template <class T> inline T __shfl(T v, int src, int width = 64) {
  (void)width;
  return emu_exchange(v, src, "__shfl");
}
template <class T> inline T __shfl_up(T v, unsigned d, int width = 64) {
  (void)width;
  const int src = emu_lane() - (int)d;
  T r = emu_exchange(v, src < 0 ? emu_lane() : src, "__shfl_up");
  return r;
}
template <class T> inline T __shfl_down(T v, unsigned d, int width = 64) {
  (void)width;
  const int src = emu_lane() + (int)d;
  return emu_exchange(v, src > 63 ? emu_lane() : src, "__shfl_down");
}
template <class T> inline T __shfl_xor(T v, int m, int width = 64) {
  (void)width;
  return emu_exchange(v, emu_lane() ^ m, "__shfl_xor");
}
inline unsigned long long __ballot(int pred) {
  EmuWave& w = emu_wave();
  w.slot[emu_lane()] = pred ? 1 : 0;
  w.barrier("__ballot");
  unsigned long long mask = 0;
  for (int i = 0; i < 64; i++)
    if (w.slot[i])
      mask |= 1ull << i;
  w.barrier("__ballot");
  return mask;
}

// The first active lane is lane 0: every call site runs with the whole wave.
inline int __builtin_amdgcn_readfirstlane(int v) { return emu_exchange(v, 0, "readfirstlane"); }
inline int __builtin_amdgcn_readlane(int v, int l) { return emu_exchange(v, l, "readlane"); }
inline int __float_as_int(float v) { return emu_from<int>(emu_bits(v)); }
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_wave_barrier() emu_wave().barrier("wave_barrier")

inline uint32_t atomicAdd(uint32_t* p, uint32_t v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
// The emulated lanes are threads: the order of their additions to one bin is
// the scheduler's, so a sum is order-independent only where it is exact.
inline double atomicAdd(double* p, double v) {
  double old, want;
  __atomic_load(p, &old, __ATOMIC_SEQ_CST);
  do {
    want = old + v;
  } while (!__atomic_compare_exchange(p, &old, &want, false, __ATOMIC_SEQ_CST,
    __ATOMIC_SEQ_CST));
  return old;
}

inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline float __fdiv_rn(float a, float b) { return a / b; }
inline float __fsqrt_rn(float a) { return std::sqrt(a); }
inline float __int_as_float(int b) { return emu_from<float>((uint32_t)b); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline unsigned __float_as_uint(float v) { return emu_from<unsigned>(emu_bits(v)); }
inline float __uint_as_float(unsigned b) { return emu_from<float>(b); }
