// 3D real FFT for the exodus smoothing: r2c once per run, then per rung a
// filtered c2r, in FP32, written for CDNA4.
//
// Replaces the FFTW plane + pencil plans of src/math/fft.c:378-434, 790-888
// and the filter/window multiply of fft.c:698-788 of sif v0.5.0.
//
// Layout is FFTW's: real field [n][n][n] (z fastest), spectrum
// [n][n][n/2+1] complex. Each axis is a batch of 1D Stockham radix-2
// transforms staged in LDS (ping-pong buffers, natural-order output, no bit
// reversal). The x and y passes load tiles of consecutive kz columns so that
// every global access is a contiguous run; the z pass works on whole rows.
// The per-mode smoothing factor (float)(LUT[k^2] * w(kx) w(ky) w(kz)) of
// fft.c:753-780 is applied as the spectrum is first read by the inverse x
// pass, so filtering costs no extra pass over memory. Power-of-two n from 16
// to 1024.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"
#include "sif_amd.h"

namespace sif_amd {

constexpr int FFT_THREADS = 256;
constexpr int FFT_TILE = 4096;  // complex elements per block (per buffer)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// P pencils of length n = 2^log2n in buffer a (pencil-major). Returns the
// buffer holding the natural-order result. tw[m] = exp(sign 2 pi i m / n).
__device__ float2* stockham(float2* a, float2* b, int log2n, int P,
  const float2* __restrict__ tw) {
  const int n = 1 << log2n;
  const int half = n >> 1;
  const int total = P * half;
  for (int ls = 0; ls < log2n; ls++) {
    const int Ns = 1 << ls;
    for (int t = threadIdx.x; t < total; t += FFT_THREADS) {
      const int p = t >> (log2n - 1);
      const int j = t & (half - 1);
      const int k = j & (Ns - 1);
      const float2 w = tw[k << (log2n - 1 - ls)];
      const float2 x0 = a[p * n + j];
      const float2 x1 = cmul(a[p * n + j + half], w);
      const int dst = p * n + ((j >> ls) << (ls + 1)) + k;
      b[dst] = make_float2(x0.x + x1.x, x0.y + x1.y);
      b[dst + Ns] = make_float2(x0.x - x1.x, x0.y - x1.y);
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
  }
  return a;
}

// --- z pass: whole rows ---

__global__ void __launch_bounds__(FFT_THREADS) r2c_rows_kernel(const float* __restrict__ in,
  float2* __restrict__ spec, int log2n, uint64_t rows, const float2* __restrict__ tw_g) {
  __shared__ float2 A[FFT_TILE], B[FFT_TILE];
  __shared__ float2 tw[512];
  const int n = 1 << log2n, nz = n / 2 + 1;
  const int P = FFT_TILE / n;
  for (int i = threadIdx.x; i < n / 2; i += FFT_THREADS)
    tw[i] = tw_g[i];
  const uint64_t row0 = (uint64_t)blockIdx.x * P;
  for (int t = threadIdx.x; t < P * n; t += FFT_THREADS) {
    const uint64_t row = row0 + t / n;
    A[t] = make_float2(row < rows ? in[row * n + (t % n)] : 0.0f, 0.0f);
  }
  __syncthreads();
  const float2* r = stockham(A, B, log2n, P, tw);
  for (int t = threadIdx.x; t < P * nz; t += FFT_THREADS) {
    const int p = t / nz, k = t % nz;
    const uint64_t row = row0 + p;
    if (row < rows)
      spec[row * nz + k] = r[p * n + k];
  }
}

__global__ void __launch_bounds__(FFT_THREADS) c2r_rows_kernel(const float2* __restrict__ spec,
  float* __restrict__ out, int log2n, uint64_t rows, float norm,
  const float2* __restrict__ tw_g) {
  __shared__ float2 A[FFT_TILE], B[FFT_TILE];
  __shared__ float2 tw[512];
  const int n = 1 << log2n, nz = n / 2 + 1;
  const int P = FFT_TILE / n;
  for (int i = threadIdx.x; i < n / 2; i += FFT_THREADS)
    tw[i] = tw_g[i];
  const uint64_t row0 = (uint64_t)blockIdx.x * P;
  for (int t = threadIdx.x; t < P * nz; t += FFT_THREADS) {
    const int p = t / nz, k = t % nz;
    const uint64_t row = row0 + p;
    const float2 v = row < rows ? spec[row * nz + k] : make_float2(0.0f, 0.0f);
    A[p * n + k] = v;
    if (k > 0 && k < n - k)
      A[p * n + n - k] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  const float2* r = stockham(A, B, log2n, P, tw);
  for (int t = threadIdx.x; t < P * n; t += FFT_THREADS) {
    const uint64_t row = row0 + t / n;
    if (row < rows)
      out[row * n + (t % n)] = r[t].x * norm;
  }
}

// --- x and y passes: tiles of TK consecutive kz columns ---
//
// Element i of pencil (a, kz) sits at base(a) + i * stride_i + kz with
// base(a) = a * stride_a. With `filter`, the element read from `in` is
// multiplied by the smoothing factor of its mode; the pass index i is then the
// x index, a the y index.
__global__ void __launch_bounds__(FFT_THREADS) cols_kernel(const float2* __restrict__ in,
  float2* __restrict__ out, int log2n, int nz, uint64_t stride_a,
  uint64_t stride_i, const float2* __restrict__ tw_g,
  const float* __restrict__ lut, const double* __restrict__ axis) {
  __shared__ float2 A[FFT_TILE], B[FFT_TILE];
  __shared__ float2 tw[512];
  const int n = 1 << log2n;
  int TK = FFT_TILE / n;
  if (TK > nz)
    TK = nz;
  const int tiles = (nz + TK - 1) / TK;
  const int a = blockIdx.x / tiles;
  const int kz0 = (blockIdx.x % tiles) * TK;
  const int cols = nz - kz0 < TK ? nz - kz0 : TK;
  for (int i = threadIdx.x; i < n / 2; i += FFT_THREADS)
    tw[i] = tw_g[i];
  const uint64_t base = (uint64_t)a * stride_a + kz0;
  const int half = n >> 1;
  for (int t = threadIdx.x; t < n * TK; t += FFT_THREADS) {
    const int i = t / TK, c = t % TK;
    float2 v = make_float2(0.0f, 0.0f);
    if (c < cols) {
      v = in[base + (uint64_t)i * stride_i + c];
      if (lut) {
        // i = x index, a = y index, kz0 + c = z index (fft.c:753-780)
        const int kx = i > half ? i - n : i;
        const int ky = a > half ? a - n : a;
        const int kz = kz0 + c;
        const uint32_t k2 = (uint32_t)(kx * kx) + (uint32_t)(ky * ky) + (uint32_t)(kz * kz);
        float s = lut[k2];
        if (axis) {
          const double cxy = axis[i] * axis[a];
          s = (float)((double)s * cxy * axis[kz]);
        }
        v = make_float2(v.x * s, v.y * s);
      }
    }
    A[c * n + i] = v;
  }
  __syncthreads();
  const float2* r = stockham(A, B, log2n, TK, tw);
  for (int t = threadIdx.x; t < n * TK; t += FFT_THREADS) {
    const int i = t / TK, c = t % TK;
    if (c < cols)
      out[base + (uint64_t)i * stride_i + c] = r[c * n + i];
  }
}

// --- host side ---

namespace {

struct Twiddles {
  int n = 0;
  DeviceBuffer fwd, inv;
};
Twiddles g_tw;
DeviceBuffer g_lut, g_axis;
std::vector<float> g_lut_host;
std::vector<double> g_axis_host;

int log2_exact(uint32_t n) {
  int l = 0;
  while ((1u << l) < n)
    l++;
  return (1u << l) == n ? l : -1;
}

int twiddles(uint32_t n) {
  if (g_tw.n == (int)n)
    return SIF_OK;
  std::vector<float2> f(n / 2), b(n / 2);
  for (uint32_t m = 0; m < n / 2; m++) {
    const double ang = 2.0 * 3.14159265358979323846 * (double)m / (double)n;
    f[m] = make_float2((float)cos(ang), (float)-sin(ang));
    b[m] = make_float2((float)cos(ang), (float)sin(ang));
  }
  int st = g_tw.fwd.reserve(n / 2 * sizeof(float2));
  if (st == SIF_OK)
    st = g_tw.inv.reserve(n / 2 * sizeof(float2));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpy(g_tw.fwd.ptr, f.data(), n / 2 * sizeof(float2), hipMemcpyHostToDevice));
  SIF_HIP(hipMemcpy(g_tw.inv.ptr, b.data(), n / 2 * sizeof(float2), hipMemcpyHostToDevice));
  g_tw.n = (int)n;
  return SIF_OK;
}

int check_n(uint32_t n, int* log2n) {
  *log2n = log2_exact(n);
  if (*log2n < 4 || n > 1024) {
    error_set(SIF_ERR_UNSUPPORTED,
      "fft: n = %u cells per side (this build transforms powers of two from 16 to 1024)",
      n);
    return SIF_ERR_UNSUPPORTED;
  }
  return SIF_OK;
}

}  // namespace

// src/math/fft.c:521-568, in float, on the host (the same libm as the
// reference's own table), indexed by the integer |k|^2.
void build_filter_lut(int filter, float r, float sigma, float box, uint32_t n,
  std::vector<float>& lut) {
  const uint32_t h = n >> 1;
  const uint32_t max_k2 = 3 * h * h;
  lut.resize(max_k2 + 1);
  const float factor = (float)(r * 2.0 * SIF_PI / box);
  const float factor2 = factor * factor;
  const float gfactor = (float)(sigma * 2.0 * SIF_PI / box);
  const float gfactor2 = gfactor * gfactor;
  for (uint32_t k2 = 0; k2 <= max_k2; k2++) {
    float v;
    if (k2 == 0) {
      v = 1.0f;
    } else if (filter == SIF_AMD_FILTER_TOP_HAT) {
      const float kr = factor * sqrtf((float)k2);
      v = kr > 1e-4f ? 3.0f * (sinf(kr) - kr * cosf(kr)) / (kr * kr * kr) : 1.0f;
    } else if (filter == SIF_AMD_FILTER_GAUSSIAN) {
      v = expf(-0.5f * factor2 * (float)k2);
    } else {
      v = 1.0f;
    }
    if (sigma > 0.0f && k2 > 0)
      v *= expf(-0.5f * gfactor2 * (float)k2);
    lut[k2] = v;
  }
}

// fft.c:890-975: per-axis window correction, 1/W (full) or W/S (aliasing).
void build_window_axis(uint32_t n, int order, int mode, std::vector<double>& axis) {
  axis.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t h = n >> 1;
    const int32_t k = i > h ? (int32_t)i - (int32_t)n : (int32_t)i;
    const double x = (double)SIF_PI * (double)k / (double)n;
    const double s = (x < 1e-9 && x > -1e-9) ? 1.0 : sin(x) / x;
    double w = 1.0;
    for (int j = 0; j < order; j++)
      w *= s;
    double S = 1.0;
    if (mode) {
      const double s2 = sin(x) * sin(x);
      switch (order) {
      case 1:
        S = 1.0;
        break;
      case 2:
        S = 1.0 - 2.0 * s2 / 3.0;
        break;
      case 3:
        S = 1.0 - s2 + 2.0 * s2 * s2 / 15.0;
        break;
      default:
        S = 1.0 - 4.0 * s2 / 3.0 + 2.0 * s2 * s2 / 5.0 - 4.0 * s2 * s2 * s2 / 315.0;
      }
    }
    axis[i] = order == 0 ? 1.0 : (mode ? w / S : 1.0 / w);
  }
}

int fft_r2c(const float* values, uint32_t n, float2* spec, float2* work) {
  int log2n;
  int st = check_n(n, &log2n);
  if (st == SIF_OK)
    st = twiddles(n);
  if (st != SIF_OK)
    return st;
  const int nz = n / 2 + 1;
  const uint64_t rows = (uint64_t)n * n;
  const int P = FFT_TILE / n;
  timing_begin(SIF_AMD_STAGE_FFT_FORWARD);
  r2c_rows_kernel<<<grid_for(rows, P), FFT_THREADS, 0, stream()>>>(
    values, work, log2n, rows, g_tw.fwd.as<float2>());
  SIF_HIP(hipGetLastError());
  int TK = FFT_TILE / n;
  if (TK > nz)
    TK = nz;
  const int tiles = (nz + TK - 1) / TK;
  // y pass: pencil a = x, i = y
  cols_kernel<<<n * tiles, FFT_THREADS, 0, stream()>>>(work, work, log2n, nz,
    (uint64_t)n * nz, (uint64_t)nz, g_tw.fwd.as<float2>(), nullptr, nullptr);
  SIF_HIP(hipGetLastError());
  // x pass: pencil a = y, i = x
  cols_kernel<<<n * tiles, FFT_THREADS, 0, stream()>>>(work, spec, log2n, nz,
    (uint64_t)nz, (uint64_t)n * nz, g_tw.fwd.as<float2>(), nullptr, nullptr);
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_FFT_FORWARD);
  if (timing_active())  // the grid once, the spectrum once
    timing_bytes(SIF_AMD_STAGE_FFT_FORWARD, 4.0 * (double)rows * n + 8.0 * (double)rows * nz);
  return SIF_OK;
}

// The backward stage of fft.c:703-888 for one filter: spectrum (unchanged)
// -> smoothed real field, normalised by (1.0f / n^3) as fft.c:832 does.
int fft_filter_c2r(const float2* spec, uint32_t n, float box, int filter,
  float r, float sigma, int window_order, int window_mode, float* out,
  float2* work) {
  int log2n;
  int st = check_n(n, &log2n);
  if (st == SIF_OK)
    st = twiddles(n);
  if (st != SIF_OK)
    return st;
  build_filter_lut(filter, r, sigma, box, n, g_lut_host);
  st = g_lut.reserve(g_lut_host.size() * sizeof(float));
  if (st != SIF_OK)
    return st;
  const bool windowed = window_order > 0 && filter != SIF_AMD_FILTER_NONE;
  if (windowed) {
    build_window_axis(n, window_order, window_mode, g_axis_host);
    st = g_axis.reserve(n * sizeof(double));
    if (st != SIF_OK)
      return st;
  }
  timing_begin(SIF_AMD_STAGE_SMOOTH);
  SIF_HIP(hipMemcpyAsync(g_lut.ptr, g_lut_host.data(), g_lut_host.size() * sizeof(float),
    hipMemcpyHostToDevice, stream()));
  if (windowed)
    SIF_HIP(hipMemcpyAsync(g_axis.ptr, g_axis_host.data(), n * sizeof(double),
      hipMemcpyHostToDevice, stream()));
  const int nz = n / 2 + 1;
  int TK = FFT_TILE / n;
  if (TK > nz)
    TK = nz;
  const int tiles = (nz + TK - 1) / TK;
  const bool filtered = filter != SIF_AMD_FILTER_NONE;
  // x pass with the filter fused: pencil a = y, i = x
  cols_kernel<<<n * tiles, FFT_THREADS, 0, stream()>>>(spec, work, log2n, nz,
    (uint64_t)nz, (uint64_t)n * nz, g_tw.inv.as<float2>(),
    filtered ? g_lut.as<float>() : nullptr, windowed ? g_axis.as<double>() : nullptr);
  SIF_HIP(hipGetLastError());
  // y pass: pencil a = x, i = y
  cols_kernel<<<n * tiles, FFT_THREADS, 0, stream()>>>(work, work, log2n, nz,
    (uint64_t)n * nz, (uint64_t)nz, g_tw.inv.as<float2>(), nullptr, nullptr);
  SIF_HIP(hipGetLastError());
  const uint64_t rows = (uint64_t)n * n;
  const int P = FFT_TILE / n;
  const float norm = 1.0f / (float)((uint64_t)n * n * n);
  c2r_rows_kernel<<<grid_for(rows, P), FFT_THREADS, 0, stream()>>>(
    work, out, log2n, rows, norm, g_tw.inv.as<float2>());
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_SMOOTH);
  if (timing_active())  // the spectrum once, the field once
    timing_bytes(SIF_AMD_STAGE_SMOOTH, 8.0 * (double)rows * nz + 4.0 * (double)rows * n);
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

uint64_t sif_amd_spectrum_bytes(uint32_t n) {
  return (uint64_t)n * n * (n / 2 + 1) * 2 * sizeof(float);
}

int sif_amd_fft_r2c(const float* values, uint32_t n, float* spectrum, void* work) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  st = fft_r2c(values, n, (float2*)spectrum, (float2*)work);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_fft_filter_c2r(const float* spectrum, uint32_t n, float box,
  int filter, float r, float sigma, int window_order, int window_mode,
  float* out, void* work) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  st = fft_filter_c2r((const float2*)spectrum, n, box, filter, r, sigma,
    window_order, window_mode, out, (float2*)work);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
