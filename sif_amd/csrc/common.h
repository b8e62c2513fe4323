// sif-amd internal: HIP plumbing shared by the kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstdio>

#include "sif/core/macros.h"

namespace sif_amd {

// The error record (logger.c of the reference keeps it thread-local too).
void error_set(int status, const char* fmt, ...)
  __attribute__((format(printf, 2, 3)));

// Library stream, created by sif_system_init(); blocking with respect to the
// null stream so work queued by other users of the device orders with it.
hipStream_t stream();
// A second stream (same default priority) for work that may overlap the
// library stream's: the next rung's smoothing while a rung is committed.
hipStream_t aux_stream();
bool initialised();
int ensure_init();  // lazy device init for the sif_amd_* device entry points
// The calling thread's own stream pair (null, null: the library's), and
// whether it has one.
void set_thread_streams(hipStream_t s, hipStream_t aux);
bool on_own_streams();

// Per-stage device timing (HIP events on the library stream).
void timing_begin(int stage);
void timing_end(int stage);
void timing_begin(int stage, hipStream_t s);
void timing_end(int stage, hipStream_t s);
bool timing_active();
void timing_add(int stage, double ms);  // host-side stages (the commit)
void timing_bytes(int stage, double bytes);  // algorithmic bytes of a stage

// Checks a HIP call; on failure records SIF_ERR_ALLOC/INVALID and returns it.
int hip_check(hipError_t e, const char* what);

// The host's waits on the rung loop's critical path: the event (or all work
// enqueued so far on the stream) polled until it completes, which wakes the
// host as soon as the device is done instead of after a blocking wait's
// sleep. The caller's thread spins meanwhile.
int wait_event(hipEvent_t e);
int wait_stream(hipStream_t s);

#define SIF_HIP(call)                                                          \
  do {                                                                         \
    int _s = ::sif_amd::hip_check((call), #call);                              \
    if (_s != SIF_OK)                                                          \
      return _s;                                                               \
  } while (0)

// Device scratch that grows on demand and is reused across calls.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t want);
  void release();
  ~DeviceBuffer() { release(); }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// The order-preserving integer image of a float (the candidate sort's key)
// and back: a lower key is a lower value, -0 just below +0.
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  return (unsigned)(g > 0x7fffffffull ? 0x7fffffffull : (g ? g : 1));
}

}  // namespace sif_amd

// Stage ids for the per-stage timers.
enum {
  SIF_AMD_STAGE_CIC = 0,
  SIF_AMD_STAGE_CONTRAST,
  SIF_AMD_STAGE_MESH,
  SIF_AMD_STAGE_FFT_FORWARD,
  SIF_AMD_STAGE_SMOOTH,
  SIF_AMD_STAGE_SCAN,
  SIF_AMD_STAGE_PREFILTER,
  SIF_AMD_STAGE_GROW,
  SIF_AMD_STAGE_COMMIT,
  SIF_AMD_STAGE_RETEST,
  SIF_AMD_STAGE_SORT,
  SIF_AMD_STAGE_HOST,
  SIF_AMD_STAGE_PROFILES,
  SIF_AMD_STAGE_COMMIT_ORDER,
  SIF_AMD_STAGE_COMMIT_INDEX,
  SIF_AMD_N_STAGES
};
