// sif-amd runtime: error record, system init, device memory helpers and the
// per-stage timers.
//
// Replaces include/sif/utils/logger.h:297-349 (error record) and
// include/sif/core/system.h (sif_system_init/finalise) of sif v0.5.0 for the
// exodus path.
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "sif/core/system.h"
#include "sif/utils/logger.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {
thread_local int t_status = SIF_OK;
thread_local char t_message[512] = "";

std::mutex g_mutex;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
hipStream_t g_aux = nullptr;

// Stage timing: a begin and an end event per bracketed launch, recorded on
// the library stream and resolved only when read (or when the pending list
// fills), so timing adds no synchronisation to the run it measures.
bool g_timing = false;
struct Pending {
  int stage;
  hipEvent_t start, stop;
};
std::vector<hipEvent_t> g_free_events;
std::vector<Pending> g_pending;
hipEvent_t g_open_start[SIF_AMD_N_STAGES];
bool g_open[SIF_AMD_N_STAGES];
double g_ms[SIF_AMD_N_STAGES];
double g_bytes[SIF_AMD_N_STAGES];
uint64_t g_count[SIF_AMD_N_STAGES];
constexpr size_t PENDING_MAX = 8192;
}  // namespace

void error_set(int status, const char* fmt, ...) {
  t_status = status;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_message, sizeof(t_message), fmt, ap);
  va_end(ap);
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess)
    return SIF_OK;
  const int st =
    (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SIF_ERR_ALLOC
                                                                 : SIF_ERR_INVALID;
  error_set(st, "HIP: %s failed: %s", what, hipGetErrorString(e));
  return st;
}

static int init_locked() {
  if (g_init)
    return SIF_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    error_set(SIF_ERR_UNSUPPORTED, "sif_system_init: no HIP device is visible");
    return SIF_ERR_UNSUPPORTED;
  }
  SIF_HIP(hipGetDevice(&g_device));
  SIF_HIP(hipStreamCreate(&g_stream));
  SIF_HIP(hipStreamCreate(&g_aux));
  g_init = true;
  return SIF_OK;
}

int ensure_init() {
  std::lock_guard<std::mutex> lock(g_mutex);
  return init_locked();
}

// A thread may run the library on streams of its own (the batched form's
// workers): its calls then use them, and stage timing is off there.
thread_local hipStream_t t_stream = nullptr, t_aux = nullptr;

bool initialised() { return g_init; }
hipStream_t stream() { return t_stream ? t_stream : g_stream; }
hipStream_t aux_stream() { return t_aux ? t_aux : g_aux; }
bool on_own_streams() { return t_stream != nullptr; }
void set_thread_streams(hipStream_t s, hipStream_t aux) {
  t_stream = s;
  t_aux = aux;
}

int DeviceBuffer::reserve(size_t want) {
  if (want <= bytes && ptr)
    return SIF_OK;
  release();
  if (want == 0)
    want = 256;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    error_set(SIF_ERR_ALLOC, "device allocation of %zu bytes failed: %s", want,
      hipGetErrorString(e));
    return SIF_ERR_ALLOC;
  }
  bytes = want;
  return SIF_OK;
}

void DeviceBuffer::release() {
  if (ptr)
    (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

static hipEvent_t take_event() {
  if (!g_free_events.empty()) {
    hipEvent_t e = g_free_events.back();
    g_free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess)
    return nullptr;
  return e;
}

// Waits for the recorded events and adds their intervals to the stages.
static void timing_resolve() {
  if (g_pending.empty())
    return;
  for (const Pending& p : g_pending) {
    (void)hipEventSynchronize(p.stop);  // stages on two streams end out of order
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      g_ms[p.stage] += ms;
      g_count[p.stage]++;
    }
    g_free_events.push_back(p.start);
    g_free_events.push_back(p.stop);
  }
  g_pending.clear();
}

void timing_begin(int s) { timing_begin(s, stream()); }
void timing_end(int s) { timing_end(s, stream()); }

void timing_begin(int s, hipStream_t st) {
  if (!g_timing || t_stream)
    return;
  hipEvent_t e = take_event();
  if (!e)
    return;
  (void)hipEventRecord(e, st);
  if (g_open[s])
    g_free_events.push_back(g_open_start[s]);
  g_open_start[s] = e;
  g_open[s] = true;
}

void timing_end(int s, hipStream_t st) {
  if (!g_timing || t_stream || !g_open[s])
    return;
  hipEvent_t e = take_event();
  if (!e)
    return;
  (void)hipEventRecord(e, st);
  g_pending.push_back(Pending{s, g_open_start[s], e});
  g_open[s] = false;
  if (g_pending.size() >= PENDING_MAX)
    timing_resolve();
}

bool timing_active() { return g_timing && !t_stream; }

void timing_add(int s, double ms) {
  if (t_stream)
    return;
  g_ms[s] += ms;
  g_count[s]++;
}

void timing_bytes(int s, double bytes) {
  if (!t_stream)
    g_bytes[s] += bytes;
}

// grid-stride copy, one 16-byte element per thread per step
__global__ void __launch_bounds__(256) copy_kernel(uint4* __restrict__ dst,
  const uint4* __restrict__ src, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

int wait_event(hipEvent_t e) {
  for (;;) {
    const hipError_t r = hipEventQuery(e);
    if (r == hipSuccess)
      return SIF_OK;
    if (r != hipErrorNotReady)
      return hip_check(r, "hipEventQuery");
    __builtin_ia32_pause();
  }
}

int wait_stream(hipStream_t s) {
  thread_local hipEvent_t done = nullptr;
  if (!done)
    SIF_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  SIF_HIP(hipEventRecord(done, s));
  return wait_event(done);
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

const char* sif_error_message(void) { return t_message; }
int sif_error_status(void) { return t_status; }
int sif_error_errno(void) { return 0; }
void sif_error_clear(void) {
  t_status = SIF_OK;
  t_message[0] = '\0';
}

int sif_system_init(const sif_config_t* config) {
  (void)config;
  std::lock_guard<std::mutex> lock(g_mutex);
  return init_locked();
}

void sif_system_finalise(void) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_init)
    return;
  (void)hipStreamSynchronize(g_stream);
  (void)hipStreamSynchronize(g_aux);
  timing_resolve();
  for (int s = 0; s < SIF_AMD_N_STAGES; s++)
    if (g_open[s]) {
      g_free_events.push_back(g_open_start[s]);
      g_open[s] = false;
    }
  for (hipEvent_t e : g_free_events)
    (void)hipEventDestroy(e);
  g_free_events.clear();
  (void)hipStreamDestroy(g_stream);
  (void)hipStreamDestroy(g_aux);
  g_stream = nullptr;
  g_aux = nullptr;
  g_init = false;
}

int sif_system_is_initialised(void) { return g_init ? 1 : 0; }

int sif_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
    return 0;
  return n;
}

int sif_amd_malloc(void** ptr, uint64_t bytes) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMalloc(ptr, bytes ? bytes : 256));
  return SIF_OK;
}

int sif_amd_free(void* ptr) {
  if (ptr)
    SIF_HIP(hipFree(ptr));
  return SIF_OK;
}

int sif_amd_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  SIF_HIP(hipStreamSynchronize(g_stream));
  return SIF_OK;
}

int sif_amd_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  SIF_HIP(hipStreamSynchronize(g_stream));
  return SIF_OK;
}

int sif_amd_memset(void* dst, int value, uint64_t bytes) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(dst, value, bytes, g_stream));
  SIF_HIP(hipStreamSynchronize(g_stream));
  return SIF_OK;
}

int sif_amd_copy(void* dst, const void* src, uint64_t bytes) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if ((bytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u)) {
    error_set(SIF_ERR_INVALID, "sif_amd_copy: bytes and both pointers must be 16-byte aligned");
    return SIF_ERR_INVALID;
  }
  if (bytes == 0)
    return SIF_OK;
  const uint64_t n = bytes / 16u;
  copy_kernel<<<grid_for(n, 256) < 8192 ? grid_for(n, 256) : 8192, 256, 0, g_stream>>>(
    static_cast<uint4*>(dst), static_cast<const uint4*>(src), n);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipStreamSynchronize(g_stream));
  return SIF_OK;
}

int sif_amd_synchronize(void) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(g_stream));
  SIF_HIP(hipStreamSynchronize(g_aux));
  return SIF_OK;
}

int sif_amd_device_memory(uint64_t* used, uint64_t* total) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!used || !total) {
    error_set(SIF_ERR_INVALID, "sif_amd_device_memory: NULL argument");
    return SIF_ERR_INVALID;
  }
  size_t free_b = 0, total_b = 0;
  SIF_HIP(hipMemGetInfo(&free_b, &total_b));
  *used = (uint64_t)(total_b - free_b);
  *total = (uint64_t)total_b;
  return SIF_OK;
}

int sif_amd_timing_enable(int on) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!on)
    timing_resolve();
  g_timing = on != 0;
  return SIF_OK;
}

void sif_amd_timing_reset(void) {
  timing_resolve();
  for (int s = 0; s < SIF_AMD_N_STAGES; s++) {
    g_ms[s] = 0.0;
    g_bytes[s] = 0.0;
    g_count[s] = 0;
  }
}

int sif_amd_timing_read(double* ms, uint64_t* launches, int n_stages) {
  timing_resolve();
  for (int s = 0; s < n_stages && s < SIF_AMD_N_STAGES; s++) {
    if (ms)
      ms[s] = g_ms[s];
    if (launches)
      launches[s] = g_count[s];
  }
  return SIF_AMD_N_STAGES;
}

int sif_amd_timing_bytes(double* bytes, int n_stages) {
  timing_resolve();
  for (int s = 0; s < n_stages && s < SIF_AMD_N_STAGES; s++)
    bytes[s] = g_bytes[s];
  return SIF_AMD_N_STAGES;
}

}  // extern "C"
