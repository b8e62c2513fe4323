// The survey half of the exodus finder on the device: the footprint the
// randoms mark out, the contrast scanned against them, and the observed
// fraction of each void.
//
// Replaces sif__exodus_survey_occupancy (src/finder/exodus_survey.c:148-207)
// with its closing (85-143), sif__exodus_survey_contrast (223-234),
// sif__exodus_survey_footprint (247-318) and mesh_extent (321-342) of sif
// v0.5.0. The observed nodes are a mask of one byte a grid node (0 or 1,
// flat index (ix n + iy) n + iz): every pass over it is a coalesced byte
// stream, and the closing's box windows are plain loads of neighbours.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"
#include "sif_amd.h"

namespace sif_amd {

// exodus_survey.c:40-44: the node nearest p on an axis, clamped (p >= 0).
__device__ __forceinline__ uint32_t survey_node(float p, float inv_l, uint32_t n) {
  const uint32_t i = (uint32_t)__fadd_rn(__fmul_rn(p, inv_l), 0.5f);
  return i < n ? i : n - 1;
}

__global__ void __launch_bounds__(256) survey_mark_kernel(const float* __restrict__ x,
  const float* __restrict__ y, const float* __restrict__ z, uint64_t n, float inv_l,
  uint32_t nn, uint8_t* __restrict__ occ) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t f = ((uint64_t)survey_node(x[p], inv_l, nn) * nn +
                         survey_node(y[p], inv_l, nn)) * nn +
                       survey_node(z[p], inv_l, nn);
    occ[f] = 1;  // every writer stores the same value
  }
}

// The set bytes of a 0/1 mask, summed into *out.
__global__ void __launch_bounds__(256) byte_count_kernel(const uint8_t* __restrict__ m,
  uint64_t total, unsigned long long* __restrict__ out) {
  unsigned long long c = 0;
  const uint64_t words = total / 4;
  const uint32_t* w = reinterpret_cast<const uint32_t*>(m);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += __popc(w[i]);  // each byte is 0 or 1: one bit a set byte
  for (uint64_t i = words * 4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += m[i];
  for (int off = 32; off > 0; off >>= 1)
    c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c)
    atomicAdd(out, c);
}

// exodus_survey.c:85-122: one axis of the closing, node by node. A dilation
// sets a node when any node within k along the axis (inside the grid) is set;
// an erosion when all 2k + 1 are, nodes past a face counting as unset.
__global__ void __launch_bounds__(256) survey_close_kernel(const uint8_t* __restrict__ src,
  uint8_t* __restrict__ dst, uint32_t nn, int axis, int32_t k, int erode) {
  const uint64_t n2 = (uint64_t)nn * nn, total = n2 * nn;
  const uint64_t stride = axis == 0 ? n2 : axis == 1 ? (uint64_t)nn : 1u;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < total;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t i = (int32_t)((c / stride) % nn);
    uint8_t v;
    if (erode) {
      v = (i - k >= 0 && i + k < (int32_t)nn) ? 1 : 0;
      for (int32_t j = -k; j <= k && v; j++)
        v &= src[(uint64_t)((int64_t)c + (int64_t)j * (int64_t)stride)];
    } else {
      v = 0;
      const int32_t j0 = i - k < 0 ? -i : -k;
      const int32_t j1 = i + k >= (int32_t)nn ? (int32_t)nn - 1 - i : k;
      for (int32_t j = j0; j <= j1 && !v; j++)
        v |= src[(uint64_t)((int64_t)c + (int64_t)j * (int64_t)stride)];
    }
    dst[c] = v;
  }
}

// exodus_survey.c:223-234
__global__ void __launch_bounds__(256) survey_contrast_kernel(float* __restrict__ data,
  const float* __restrict__ randoms, const uint8_t* __restrict__ occ, float alpha,
  uint64_t total) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < total;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const float expected = __fmul_rn(alpha, randoms[c]);
    data[c] = (occ[c] && expected > 0.0f) ? __fsub_rn(__fdiv_rn(data[c], expected), 1.0f)
                                          : SURVEY_OUTSIDE;
  }
}

// exodus_survey.c:247-303: one workgroup a void. Lattice point t of the
// (2 half)^3 cube around the centre, spacing h = max(r / 8, l / 2); a point
// outside [0, box)^3 is unobserved. The four counts are integers, summed in
// any order.
constexpr int FOOT_THREADS = 256;
__global__ void __launch_bounds__(FOOT_THREADS) survey_footprint_kernel(
  const uint8_t* __restrict__ occ, uint32_t nn, float l, float inv_l, float box,
  const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
  const float* __restrict__ vr, float* __restrict__ inside, float* __restrict__ shell) {
  __shared__ unsigned long long part[4][FOOT_THREADS / 64];
  const uint64_t v = blockIdx.x;
  const float cx = vx[v], cy = vy[v], cz = vz[v], r = vr[v];
  float h = __fmul_rn(0.125f, r);
  if (h < __fmul_rn(0.5f, l))
    h = __fmul_rn(0.5f, l);
  const int32_t half = (int32_t)__fdiv_rn(__fmul_rn(2.0f, r), h) + 1;
  const float r2 = __fmul_rn(r, r);
  const float r2_out = __fmul_rn(4.0f, r2);
  const uint64_t side = 2u * (uint64_t)half, pts = side * side * side;
  unsigned long long n_in = 0, occ_in = 0, n_sh = 0, occ_sh = 0;
  for (uint64_t t = threadIdx.x; t < pts; t += FOOT_THREADS) {
    const int32_t a = (int32_t)(t / (side * side)) - half;
    const int32_t b = (int32_t)((t / side) % side) - half;
    const int32_t c = (int32_t)(t % side) - half;
    const float dx = __fmul_rn(__fadd_rn((float)a, 0.5f), h);
    const float dy = __fmul_rn(__fadd_rn((float)b, 0.5f), h);
    const float dz = __fmul_rn(__fadd_rn((float)c, 0.5f), h);
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    if (d2 > r2_out)
      continue;
    const float px = __fadd_rn(cx, dx), py = __fadd_rn(cy, dy), pz = __fadd_rn(cz, dz);
    unsigned long long o = 0;
    if (px >= 0.0f && px < box && py >= 0.0f && py < box && pz >= 0.0f && pz < box)
      o = occ[((uint64_t)survey_node(px, inv_l, nn) * nn + survey_node(py, inv_l, nn)) * nn +
              survey_node(pz, inv_l, nn)];
    if (d2 <= r2) {
      n_in++;
      occ_in += o;
    } else {
      n_sh++;
      occ_sh += o;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    n_in += __shfl_xor(n_in, off, 64);
    occ_in += __shfl_xor(occ_in, off, 64);
    n_sh += __shfl_xor(n_sh, off, 64);
    occ_sh += __shfl_xor(occ_sh, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][w] = n_in;
    part[1][w] = occ_in;
    part[2][w] = n_sh;
    part[3][w] = occ_sh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s[4] = {0, 0, 0, 0};
    for (int q = 0; q < 4; q++)
      for (int k = 0; k < FOOT_THREADS / 64; k++)
        s[q] += part[q][k];
    inside[v] = s[0] ? (float)((double)s[1] / (double)s[0]) : 0.0f;
    shell[v] = s[2] ? (float)((double)s[3] / (double)s[2]) : 0.0f;
  }
}

// mesh_extent (exodus_survey.c:321-342): the bounding box of positions, as
// ordered integers (a float's bits, flipped so that integer order is float
// order), reduced with atomics; slots 0-2 the minima, 3-5 the maxima.
__device__ __forceinline__ uint32_t float_order(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float order_float(uint32_t u) {
  const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

__global__ void __launch_bounds__(256) extent_kernel(const float* __restrict__ x,
  const float* __restrict__ y, const float* __restrict__ z, uint64_t n,
  uint32_t* __restrict__ out) {
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v[3] = {float_order(x[i]), float_order(y[i]), float_order(z[i])};
    for (int a = 0; a < 3; a++) {
      lo[a] = v[a] < lo[a] ? v[a] : lo[a];
      hi[a] = v[a] > hi[a] ? v[a] : hi[a];
    }
  }
  for (int a = 0; a < 3; a++)
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t l2 = __shfl_xor(lo[a], off, 64), h2 = __shfl_xor(hi[a], off, 64);
      lo[a] = l2 < lo[a] ? l2 : lo[a];
      hi[a] = h2 > hi[a] ? h2 : hi[a];
    }
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; a++) {
      atomicMin(&out[a], lo[a]);
      atomicMax(&out[3 + a], hi[a]);
    }
}

namespace {
thread_local DeviceBuffer g_tmp_occ, g_count, g_extent, g_foot;

unsigned launch_blocks(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

int count_set(const uint8_t* occ, uint64_t total, uint64_t* out) {
  int st = g_count.reserve(8);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_count.ptr, 0, 8, stream()));
  byte_count_kernel<<<launch_blocks(total / 4 + 1), 256, 0, stream()>>>(occ, total,
    g_count.as<unsigned long long>());
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(out, g_count.ptr, 8, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}
}  // namespace

// exodus_survey.c:62-76: randoms per node from m, randoms per marked node
// (m = lambda / (1 - exp(-lambda)) for Poisson counts), by 64 bisections.
double survey_randoms_per_node(double m) {
  if (!(m > 1.0))
    return 0.0;
  double lo = 0.0, hi = m;
  for (int it = 0; it < 64; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid < m * (1.0 - std::exp(-mid)))
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

int survey_occupancy(const float* x, const float* y, const float* z, uint64_t n, uint32_t nn,
  float box, uint8_t* occ, int32_t* k_out, double* lambda_out, uint64_t* marked_out,
  uint64_t* observed_out) {
  const uint64_t total = (uint64_t)nn * nn * nn;
  const float l = box / (float)nn;
  int st = g_tmp_occ.reserve(total);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(occ, 0, total, stream()));
  if (n) {
    survey_mark_kernel<<<launch_blocks(n), 256, 0, stream()>>>(x, y, z, n, 1.0f / l, nn, occ);
    SIF_HIP(hipGetLastError());
  }
  uint64_t marked = 0;
  if ((st = count_set(occ, total, &marked)) != SIF_OK)
    return st;
  int32_t k = 0;
  double lambda = 0.0;
  uint64_t observed = 0;
  if (marked) {
    // exodus_survey.c:176-181: the smallest window expecting 128 randoms
    lambda = survey_randoms_per_node((double)n / (double)marked);
    k = 1;
    while (k < 4 && lambda * (2 * k + 1) * (2 * k + 1) * (2 * k + 1) < 128.0)
      k++;
    uint8_t* tmp = g_tmp_occ.as<uint8_t>();
    const unsigned g = launch_blocks(total);
    // dilate x, y, z, then erode x, y, z (exodus_survey.c:130-143)
    survey_close_kernel<<<g, 256, 0, stream()>>>(occ, tmp, nn, 0, k, 0);
    survey_close_kernel<<<g, 256, 0, stream()>>>(tmp, occ, nn, 1, k, 0);
    survey_close_kernel<<<g, 256, 0, stream()>>>(occ, tmp, nn, 2, k, 0);
    survey_close_kernel<<<g, 256, 0, stream()>>>(tmp, occ, nn, 0, k, 1);
    survey_close_kernel<<<g, 256, 0, stream()>>>(occ, tmp, nn, 1, k, 1);
    survey_close_kernel<<<g, 256, 0, stream()>>>(tmp, occ, nn, 2, k, 1);
    SIF_HIP(hipGetLastError());
    if ((st = count_set(occ, total, &observed)) != SIF_OK)
      return st;
  }
  if (k_out)
    *k_out = k;
  if (lambda_out)
    *lambda_out = lambda;
  if (marked_out)
    *marked_out = marked;
  if (observed_out)
    *observed_out = observed;
  return SIF_OK;
}

int survey_contrast(float* data, const float* randoms, const uint8_t* occ, float alpha,
  uint64_t total, hipStream_t s) {
  survey_contrast_kernel<<<launch_blocks(total), 256, 0, s>>>(data, randoms, occ, alpha, total);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int survey_footprint(const uint8_t* occ, uint32_t nn, float box, const float* cx,
  const float* cy, const float* cz, const float* r, uint64_t nv, float* inside, float* shell) {
  if (nv == 0)
    return SIF_OK;
  if (nv > 0x7fffffffull) {
    error_set(SIF_ERR_RANGE, "survey footprint: %llu voids (at most 2^31 - 1)",
      (unsigned long long)nv);
    return SIF_ERR_RANGE;
  }
  int st = g_foot.reserve(nv * 6 * sizeof(float));
  if (st != SIF_OK)
    return st;
  float* d = g_foot.as<float>();
  SIF_HIP(hipMemcpyAsync(d, cx, nv * 4, hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipMemcpyAsync(d + nv, cy, nv * 4, hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipMemcpyAsync(d + 2 * nv, cz, nv * 4, hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipMemcpyAsync(d + 3 * nv, r, nv * 4, hipMemcpyHostToDevice, stream()));
  const float l = box / (float)nn;
  survey_footprint_kernel<<<(unsigned)nv, FOOT_THREADS, 0, stream()>>>(occ, nn, l, 1.0f / l, box,
    d, d + nv, d + 2 * nv, d + 3 * nv, d + 4 * nv, d + 5 * nv);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(inside, d + 4 * nv, nv * 4, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipMemcpyAsync(shell, d + 5 * nv, nv * 4, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int mesh_extent(const float* x, const float* y, const float* z, uint64_t n, float lo[3],
  float hi[3]) {
  int st = g_extent.reserve(6 * 4);
  if (st != SIF_OK)
    return st;
  const uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  SIF_HIP(hipMemcpyAsync(g_extent.ptr, init, sizeof(init), hipMemcpyHostToDevice, stream()));
  if (n) {
    extent_kernel<<<launch_blocks(n), 256, 0, stream()>>>(x, y, z, n, g_extent.as<uint32_t>());
    SIF_HIP(hipGetLastError());
  }
  uint32_t got[6];
  SIF_HIP(hipMemcpyAsync(got, g_extent.ptr, sizeof(got), hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  for (int a = 0; a < 3; a++) {
    lo[a] = order_float(got[a]);
    hi[a] = order_float(got[3 + a]);
  }
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_survey_occupancy(const float* d_x, const float* d_y, const float* d_z, uint64_t n,
  uint32_t n_cells, float box, uint8_t* d_occupied, int32_t* k, double* lambda,
  uint64_t* marked, uint64_t* observed) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_occupied || n_cells == 0 || n_cells > 4096 || !(box > 0.0f) ||
      (n && (!d_x || !d_y || !d_z))) {
    error_set(SIF_ERR_INVALID, "sif_amd_survey_occupancy: invalid arguments");
    return SIF_ERR_INVALID;
  }
  return survey_occupancy(d_x, d_y, d_z, n, n_cells, box, d_occupied, k, lambda, marked,
    observed);
}

int sif_amd_survey_contrast(float* d_data, const float* d_randoms, const uint8_t* d_occupied,
  float alpha, uint64_t total_cells) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_data || !d_randoms || !d_occupied) {
    error_set(SIF_ERR_INVALID, "sif_amd_survey_contrast: invalid arguments");
    return SIF_ERR_INVALID;
  }
  st = survey_contrast(d_data, d_randoms, d_occupied, alpha, total_cells, stream());
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_survey_footprint(const uint8_t* d_occupied, uint32_t n_cells, float box,
  const float* cx, const float* cy, const float* cz, const float* r, uint64_t n_voids,
  float* inside, float* shell) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_occupied || n_cells == 0 || !(box > 0.0f) ||
      (n_voids && (!cx || !cy || !cz || !r || !inside || !shell))) {
    error_set(SIF_ERR_INVALID, "sif_amd_survey_footprint: invalid arguments");
    return SIF_ERR_INVALID;
  }
  return survey_footprint(d_occupied, n_cells, box, cx, cy, cz, r, n_voids, inside, shell);
}

}  // extern "C"
