// Spherically averaged density profiles of a density-contrast grid, their
// transpose, and the transpose of the CIC deposit: the forward and adjoint
// kernels of the chain positions -> CIC -> smoothing -> profiles.
//
// The profiles follow sif_profiles (src/measure/profiles.c:354-673), density
// part, with the grid's nodes as the mass elements: node g at position g*l
// carries mass (1 + delta[g]) rho_bar l^3. A node is binned as the reference
// bins a tracer (profiles.c:202-223): nearest-image separation, r^2 <=
// (ext rv)^2, bin = (uint32_t)(sqrt(r^2) / (rv bin_width)), kept below n_bins.
// The cumulative profile (profiles.c:332-350, 619-635) is then
//
//     p[c, j] = l^3 sum_{bins <= j} (1 + delta) / (vol_j rv^3) - 1,
//     vol_j = (4/3) pi ((j + 1) bin_width)^3   (in units of rv),
//
// and a centre with no radius keeps a zero row (profiles.c:523-527).
//
// p is linear in delta, so its transpose spreads dL/dp[c, j] back to every
// node of bin b <= j as w_c(b) = sum_{j >= b} dL/dp[c, j] l^3 / (vol_j rv^3).
// The forward gathers one centre per workgroup; the transpose pulls, one
// 16^3 tile of nodes per workgroup, over the centres whose cube of nodes
// touches the tile, in centre order -- no float atomics, the same bits for
// every launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {

constexpr uint32_t PROFILE_THREADS = 256;
constexpr uint32_t PROFILE_TILE = 16;  // adjoint tile side, in nodes
constexpr uint32_t PROFILE_CHUNK = 32;  // centres staged in LDS at a time

// One centre's walk: the node cube around it, as the reference bounds a
// void's cells, at most one full row per axis.
struct Cube {
  int32_t lo[3];
  int32_t span[3];
};

__host__ __device__ inline int32_t floor_div(int32_t a, int32_t b) {
  const int32_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__host__ __device__ inline Cube centre_cube(float cx, float cy, float cz, float r_max, float l,
  uint32_t N) {
  Cube k;
  const float c[3] = {cx, cy, cz};
  for (int a = 0; a < 3; a++) {
    const int32_t lo = (int32_t)floorf((c[a] - r_max) / l);
    const int32_t hi = (int32_t)ceilf((c[a] + r_max) / l);
    int32_t span = hi - lo + 1;
    if (span > (int32_t)N)
      span = (int32_t)N;
    k.lo[a] = lo;
    k.span[a] = span;
  }
  return k;
}

__device__ __forceinline__ uint32_t wrap_node(int32_t g, int32_t N) {
  int32_t w = g % N;
  return (uint32_t)(w < 0 ? w + N : w);
}

// profiles.c:206-215: the branchless nearest image of a separation
__device__ __forceinline__ float nearest(float d, float box, float half) {
  return d - (box * (float)(d > half) - box * (float)(d < -half));
}

// vol_j of fill_edges (profiles.c:332-345), cumulative, in float
__device__ __forceinline__ float bin_volume(uint32_t j, float bw) {
  const float r_outer = (float)(j + 1) * bw;
  return (float)(4.0 / 3.0) * (float)SIF_PI * (r_outer * r_outer * r_outer);
}

// One workgroup per centre: count and delta sum per bin over the nodes of
// its sphere, then the cumulative profile.
template <int NB>
__global__ void __launch_bounds__(PROFILE_THREADS) profile_kernel(
  const float* __restrict__ delta, uint32_t N, float l, float box,
  const float* __restrict__ cxs, const float* __restrict__ cys, const float* __restrict__ czs,
  const float* __restrict__ rvs, float ext, uint32_t n_bins, float bw,
  float* __restrict__ out) {
  __shared__ double s_sum[PROFILE_THREADS / 64][NB];
  __shared__ uint32_t s_cnt[PROFILE_THREADS / 64][NB];
  const uint32_t c = blockIdx.x;
  const float rv = rvs[c];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!(rv > 0.0f)) {
    for (uint32_t j = threadIdx.x; j < n_bins; j += blockDim.x)
      out[(uint64_t)c * n_bins + j] = 0.0f;
    return;
  }
  const float cx = cxs[c], cy = cys[c], cz = czs[c];
  const float r_max = rv * ext;
  const float r_max_sq = r_max * r_max;
  const float inv_bin = 1.0f / (rv * bw);
  const float half = box * 0.5f;
  const Cube k = centre_cube(cx, cy, cz, r_max, l, N);
  double sum[NB];
  uint32_t cnt[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    sum[j] = 0.0;
    cnt[j] = 0;
  }
  const uint32_t rows = (uint32_t)k.span[0] * (uint32_t)k.span[1];
  for (uint32_t row = wave; row < rows; row += PROFILE_THREADS / 64) {
    const uint32_t gx = wrap_node(k.lo[0] + (int32_t)(row / (uint32_t)k.span[1]), (int32_t)N);
    const uint32_t gy = wrap_node(k.lo[1] + (int32_t)(row % (uint32_t)k.span[1]), (int32_t)N);
    const float dx = nearest((float)gx * l - cx, box, half);
    const float dy = nearest((float)gy * l - cy, box, half);
    const float dxy = dx * dx + dy * dy;
    if (dxy > r_max_sq)
      continue;
    const float* rowp = delta + ((uint64_t)gx * N + gy) * N;
    for (int32_t z = lane; z < k.span[2]; z += 64) {
      const uint32_t gz = wrap_node(k.lo[2] + z, (int32_t)N);
      const float dz = nearest((float)gz * l - cz, box, half);
      const float r2 = dxy + dz * dz;
      if (r2 <= r_max_sq) {
        const uint32_t bin = (uint32_t)(__builtin_sqrtf(r2) * inv_bin);
        const double v = (double)rowp[gz];
#pragma unroll
        for (int j = 0; j < NB; j++) {
          const bool hit = bin == (uint32_t)j;
          sum[j] += hit ? v : 0.0;
          cnt[j] += hit ? 1u : 0u;
        }
      }
    }
  }
  // wave sums, then the workgroup's in wave order
#pragma unroll
  for (int j = 0; j < NB; j++) {
    double s = sum[j];
    uint32_t n = cnt[j];
    for (int off = 32; off > 0; off >>= 1) {
      s += __shfl_xor(s, off, 64);
      n += __shfl_xor(n, off, 64);
    }
    if (lane == 0) {
      s_sum[wave][j] = s;
      s_cnt[wave][j] = n;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double l3 = (double)l * (double)l * (double)l;
    const double rv3 = (double)rv * (double)rv * (double)rv;
    double cum_sum = 0.0;
    uint64_t cum_cnt = 0;
    for (uint32_t j = 0; j < n_bins; j++) {
      double s = 0.0;
      uint64_t n = 0;
      for (int w = 0; w < (int)(PROFILE_THREADS / 64); w++) {
        s += s_sum[w][j];
        n += s_cnt[w][j];
      }
      cum_sum += s;
      cum_cnt += n;
      const double mass = l3 * ((double)cum_cnt + cum_sum);
      out[(uint64_t)c * n_bins + j] =
        (float)(mass / ((double)bin_volume(j, bw) * rv3) - 1.0);
    }
  }
}

// per centre: how many adjoint tiles its cube touches
__global__ void profile_pair_count_kernel(const float* __restrict__ cxs,
  const float* __restrict__ cys, const float* __restrict__ czs, const float* __restrict__ rvs,
  uint32_t nc, float ext, float l, uint32_t N, uint32_t nt, uint32_t* __restrict__ count) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc)
    return;
  const float rv = rvs[c];
  if (!(rv > 0.0f)) {
    count[c] = 0;
    return;
  }
  const Cube k = centre_cube(cxs[c], cys[c], czs[c], rv * ext, l, N);
  uint32_t total = 1;
  for (int a = 0; a < 3; a++) {
    const int32_t t0 = floor_div(k.lo[a], (int32_t)PROFILE_TILE);
    const int32_t t1 = floor_div(k.lo[a] + k.span[a] - 1, (int32_t)PROFILE_TILE);
    uint32_t n = (uint32_t)(t1 - t0 + 1);
    total *= n < nt ? n : nt;
  }
  count[c] = total;
}

// per centre: its (tile, centre) pairs at its offset, tiles in x, y, z order
__global__ void profile_pair_emit_kernel(const float* __restrict__ cxs,
  const float* __restrict__ cys, const float* __restrict__ czs, const float* __restrict__ rvs,
  uint32_t nc, float ext, float l, uint32_t N, uint32_t nt, const uint64_t* __restrict__ offset,
  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc)
    return;
  const float rv = rvs[c];
  if (!(rv > 0.0f))
    return;
  const Cube k = centre_cube(cxs[c], cys[c], czs[c], rv * ext, l, N);
  int32_t t0[3];
  uint32_t n[3];
  for (int a = 0; a < 3; a++) {
    t0[a] = floor_div(k.lo[a], (int32_t)PROFILE_TILE);
    const int32_t t1 = floor_div(k.lo[a] + k.span[a] - 1, (int32_t)PROFILE_TILE);
    const uint32_t m = (uint32_t)(t1 - t0[a] + 1);
    n[a] = m < nt ? m : nt;
  }
  uint64_t o = offset[c];
  for (uint32_t a = 0; a < n[0]; a++) {
    const uint32_t tx = wrap_node(t0[0] + (int32_t)a, (int32_t)nt);
    for (uint32_t b = 0; b < n[1]; b++) {
      const uint32_t ty = wrap_node(t0[1] + (int32_t)b, (int32_t)nt);
      for (uint32_t d = 0; d < n[2]; d++) {
        const uint32_t tz = wrap_node(t0[2] + (int32_t)d, (int32_t)nt);
        keys[o] = (tx * nt + ty) * nt + tz;
        vals[o] = c;
        o++;
      }
    }
  }
}

// sorted keys -> [start, end) of every tile's run
__global__ void profile_tile_runs_kernel(const uint32_t* __restrict__ keys, uint32_t n_pairs,
  uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs)
    return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k)
    start[k] = i;
  if (i + 1 == n_pairs || keys[i + 1] != k)
    end[k] = i + 1;
}

// One workgroup per 16^3 tile; thread (a, b) owns the column of 16 nodes
// (x0 + a, y0 + b, z0 ..). The centres whose cube touches the tile are taken
// in centre order, PROFILE_CHUNK at a time through LDS.
template <int NB>
__global__ void __launch_bounds__(PROFILE_THREADS) profile_adjoint_kernel(
  const float* __restrict__ grad_p, uint32_t N, float l, float box,
  const float* __restrict__ cxs, const float* __restrict__ cys, const float* __restrict__ czs,
  const float* __restrict__ rvs, float ext, uint32_t n_bins, float bw, uint32_t nt,
  const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ tile_end,
  const uint32_t* __restrict__ centres, float* __restrict__ grad_delta) {
  __shared__ float s_c[PROFILE_CHUNK][4];  // cx, cy, cz, r_max^2
  __shared__ float s_inv[PROFILE_CHUNK];
  __shared__ double s_w[PROFILE_CHUNK][NB];
  const uint32_t t = blockIdx.x;
  const uint32_t tx = t / (nt * nt), ty = (t / nt) % nt, tz = t % nt;
  const uint32_t gx = tx * PROFILE_TILE + threadIdx.x / PROFILE_TILE;
  const uint32_t gy = ty * PROFILE_TILE + threadIdx.x % PROFILE_TILE;
  const uint32_t z0 = tz * PROFILE_TILE;
  const float half = box * 0.5f;
  const float px = (float)gx * l, py = (float)gy * l;
  float pz[PROFILE_TILE];
#pragma unroll
  for (uint32_t z = 0; z < PROFILE_TILE; z++)
    pz[z] = (float)(z0 + z) * l;
  double acc[PROFILE_TILE];
#pragma unroll
  for (uint32_t z = 0; z < PROFILE_TILE; z++)
    acc[z] = 0.0;
  const uint32_t s0 = tile_start[t], s1 = tile_end[t];
  const double l3 = (double)l * (double)l * (double)l;
  for (uint32_t base = s0; base < s1; base += PROFILE_CHUNK) {
    const uint32_t m = s1 - base < PROFILE_CHUNK ? s1 - base : PROFILE_CHUNK;
    __syncthreads();
    if (threadIdx.x < m) {
      const uint32_t c = centres[base + threadIdx.x];
      const float rv = rvs[c];
      const float r_max = rv * ext;
      s_c[threadIdx.x][0] = cxs[c];
      s_c[threadIdx.x][1] = cys[c];
      s_c[threadIdx.x][2] = czs[c];
      s_c[threadIdx.x][3] = r_max * r_max;
      s_inv[threadIdx.x] = 1.0f / (rv * bw);
      // w(b) = sum_{j >= b} dL/dp[c, j] l^3 / (vol_j rv^3)
      const double rv3 = (double)rv * (double)rv * (double)rv;
      double run = 0.0;
      for (int j = NB - 1; j >= 0; j--) {
        if ((uint32_t)j < n_bins)
          run += (double)grad_p[(uint64_t)c * n_bins + j] * l3 /
                 ((double)bin_volume((uint32_t)j, bw) * rv3);
        s_w[threadIdx.x][j] = (uint32_t)j < n_bins ? run : 0.0;
      }
    }
    __syncthreads();
    for (uint32_t i = 0; i < m; i++) {
      const float r_max_sq = s_c[i][3];
      const float dx = nearest(px - s_c[i][0], box, half);
      const float dy = nearest(py - s_c[i][1], box, half);
      const float dxy = dx * dx + dy * dy;
      if (dxy > r_max_sq)
        continue;
      const float cz = s_c[i][2], inv_bin = s_inv[i];
#pragma unroll
      for (uint32_t z = 0; z < PROFILE_TILE; z++) {
        const float dz = nearest(pz[z] - cz, box, half);
        const float r2 = dxy + dz * dz;
        if (r2 <= r_max_sq) {
          const uint32_t bin = (uint32_t)(__builtin_sqrtf(r2) * inv_bin);
          if (bin < n_bins)
            acc[z] += s_w[i][bin];
        }
      }
    }
  }
  float* col = grad_delta + ((uint64_t)gx * N + gy) * N + z0;
#pragma unroll
  for (uint32_t z = 0; z < PROFILE_TILE; z++)
    col[z] = (float)acc[z];
}

__device__ __forceinline__ float fold_pos(float v, float box) {
  // src/structures/field_internal.h:58-64, as the deposit folds
  if (v >= box)
    v -= box;
  else if (v < 0.0f)
    v += box;
  return v >= box ? 0.0f : v;
}

// The transpose of the deposit (grid.c:585-650, cic.hip): for
// rho[g] = sum_p w_p(g) / l^3, d(sum_g G[g] rho[g]) / dx_p with the weights
// (1 - t, t) per axis and dt/dx = n / L.
__global__ void __launch_bounds__(256) cic_adjoint_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, uint64_t n, uint32_t N, float idx,
  float box, float scale, const float* __restrict__ G, float* __restrict__ gxs,
  float* __restrict__ gys, float* __restrict__ gzs) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const float ux = fold_pos(xs[p], box) * idx, uy = fold_pos(ys[p], box) * idx,
              uz = fold_pos(zs[p], box) * idx;
  const uint32_t rx = (uint32_t)ux, ry = (uint32_t)uy, rz = (uint32_t)uz;
  const uint32_t ix = rx >= N ? 0u : rx, iy = ry >= N ? 0u : ry, iz = rz >= N ? 0u : rz;
  const float tx = ux - (float)rx, ty = uy - (float)ry, tz = uz - (float)rz;
  const uint32_t jx = ix + 1 == N ? 0u : ix + 1, jy = iy + 1 == N ? 0u : iy + 1,
                 jz = iz + 1 == N ? 0u : iz + 1;
  const uint64_t NN = (uint64_t)N * N;
  const uint64_t X[2] = {ix * NN, jx * NN}, Y[2] = {(uint64_t)iy * N, (uint64_t)jy * N};
  const uint32_t Z[2] = {iz, jz};
  float g[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++)
        g[a][b][c] = G[X[a] + Y[b] + Z[c]];
  const float wx[2] = {1.0f - tx, tx}, wy[2] = {1.0f - ty, ty}, wz[2] = {1.0f - tz, tz};
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const float v = g[a][b][c];
        sx += (a ? v : -v) * wy[b] * wz[c];
        sy += (b ? v : -v) * wx[a] * wz[c];
        sz += (c ? v : -v) * wx[a] * wy[b];
      }
  gxs[p] = sx * scale;
  gys[p] = sy * scale;
  gzs[p] = sz * scale;
}

DeviceBuffer g_pair_count, g_pair_offset, g_pair_keys, g_pair_vals, g_tile_start, g_tile_end;

bool bins_supported(uint32_t n_bins) { return n_bins >= 1 && n_bins <= 32; }

}  // namespace

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_grid_profiles(const float* d_delta, uint32_t n_cells, float box, const float* d_cx,
  const float* d_cy, const float* d_cz, const float* d_rv, uint32_t n_centres, float ext,
  uint32_t n_bins, float* d_profiles) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_delta || !d_cx || !d_cy || !d_cz || !d_rv || !d_profiles || n_cells == 0 ||
      !(box > 0.0f) || !(ext > 0.0f) || !bins_supported(n_bins)) {
    error_set(SIF_ERR_INVALID, "sif_amd_grid_profiles: invalid arguments (1 <= n_bins <= 32)");
    return SIF_ERR_INVALID;
  }
  if (n_centres == 0)
    return SIF_OK;
  const float l = box / (float)n_cells;
  const float bw = ext / (float)n_bins;
  if (n_bins <= 16)
    profile_kernel<16><<<n_centres, PROFILE_THREADS, 0, stream()>>>(d_delta, n_cells, l, box,
      d_cx, d_cy, d_cz, d_rv, ext, n_bins, bw, d_profiles);
  else
    profile_kernel<32><<<n_centres, PROFILE_THREADS, 0, stream()>>>(d_delta, n_cells, l, box,
      d_cx, d_cy, d_cz, d_rv, ext, n_bins, bw, d_profiles);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int sif_amd_grid_profiles_adjoint(const float* d_grad_profiles, uint32_t n_cells, float box,
  const float* d_cx, const float* d_cy, const float* d_cz, const float* d_rv, uint32_t n_centres,
  float ext, uint32_t n_bins, float* d_grad_delta) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_grad_profiles || !d_cx || !d_cy || !d_cz || !d_rv || !d_grad_delta || n_cells == 0 ||
      n_cells % PROFILE_TILE != 0 || !(box > 0.0f) || !(ext > 0.0f) ||
      !bins_supported(n_bins)) {
    error_set(SIF_ERR_INVALID,
      "sif_amd_grid_profiles_adjoint: invalid arguments (n_cells a multiple of 16, "
      "1 <= n_bins <= 32)");
    return SIF_ERR_INVALID;
  }
  const float l = box / (float)n_cells;
  const float bw = ext / (float)n_bins;
  const uint32_t nt = n_cells / PROFILE_TILE;
  const uint32_t n_tiles = nt * nt * nt;
  st = g_tile_start.reserve((uint64_t)n_tiles * 4);
  if (st == SIF_OK)
    st = g_tile_end.reserve((uint64_t)n_tiles * 4);
  if (st == SIF_OK)
    st = g_pair_count.reserve((uint64_t)(n_centres ? n_centres : 1) * 4);
  if (st == SIF_OK)
    st = g_pair_offset.reserve((uint64_t)(n_centres + 1) * 8);
  if (st != SIF_OK)
    return st;
  // tiles with no centre see start == end
  SIF_HIP(hipMemsetAsync(g_tile_start.ptr, 0, (uint64_t)n_tiles * 4, stream()));
  SIF_HIP(hipMemsetAsync(g_tile_end.ptr, 0, (uint64_t)n_tiles * 4, stream()));
  uint64_t n_pairs = 0;
  if (n_centres) {
    profile_pair_count_kernel<<<grid_for(n_centres, 256), 256, 0, stream()>>>(d_cx, d_cy, d_cz,
      d_rv, n_centres, ext, l, n_cells, nt, g_pair_count.as<uint32_t>());
    SIF_HIP(hipGetLastError());
    st = exclusive_scan_u32(g_pair_count.as<uint32_t>(), n_centres, g_pair_offset.as<uint64_t>(),
      g_pair_offset.as<uint64_t>() + n_centres);
    if (st != SIF_OK)
      return st;
    SIF_HIP(hipMemcpyAsync(&n_pairs, g_pair_offset.as<uint64_t>() + n_centres, 8,
      hipMemcpyDeviceToHost, stream()));
    SIF_HIP(hipStreamSynchronize(stream()));
    if (n_pairs >= 0xffffffffull) {
      error_set(SIF_ERR_RANGE, "sif_amd_grid_profiles_adjoint: %llu (tile, centre) pairs",
        (unsigned long long)n_pairs);
      return SIF_ERR_RANGE;
    }
  }
  if (n_pairs) {
    st = g_pair_keys.reserve(n_pairs * 4);
    if (st == SIF_OK)
      st = g_pair_vals.reserve(n_pairs * 4);
    if (st != SIF_OK)
      return st;
    profile_pair_emit_kernel<<<grid_for(n_centres, 256), 256, 0, stream()>>>(d_cx, d_cy, d_cz,
      d_rv, n_centres, ext, l, n_cells, nt, g_pair_offset.as<uint64_t>(),
      g_pair_keys.as<uint32_t>(), g_pair_vals.as<uint32_t>());
    SIF_HIP(hipGetLastError());
    int bits = 1;
    while ((1u << bits) < n_tiles)
      bits++;
    st = radix_sort_u32(g_pair_keys.as<uint32_t>(), g_pair_vals.as<uint32_t>(),
      (uint32_t)n_pairs, bits);
    if (st != SIF_OK)
      return st;
    profile_tile_runs_kernel<<<grid_for(n_pairs, 256), 256, 0, stream()>>>(
      g_pair_keys.as<uint32_t>(), (uint32_t)n_pairs, g_tile_start.as<uint32_t>(),
      g_tile_end.as<uint32_t>());
    SIF_HIP(hipGetLastError());
  }
  if (n_bins <= 16)
    profile_adjoint_kernel<16><<<n_tiles, PROFILE_THREADS, 0, stream()>>>(d_grad_profiles,
      n_cells, l, box, d_cx, d_cy, d_cz, d_rv, ext, n_bins, bw, nt, g_tile_start.as<uint32_t>(),
      g_tile_end.as<uint32_t>(), g_pair_vals.as<uint32_t>(), d_grad_delta);
  else
    profile_adjoint_kernel<32><<<n_tiles, PROFILE_THREADS, 0, stream()>>>(d_grad_profiles,
      n_cells, l, box, d_cx, d_cy, d_cz, d_rv, ext, n_bins, bw, nt, g_tile_start.as<uint32_t>(),
      g_tile_end.as<uint32_t>(), g_pair_vals.as<uint32_t>(), d_grad_delta);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int sif_amd_cic_adjoint(const float* d_x, const float* d_y, const float* d_z, uint64_t n,
  uint32_t n_cells, float box, const float* d_grad_rho, float* d_gx, float* d_gy, float* d_gz) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_x || !d_y || !d_z || !d_grad_rho || !d_gx || !d_gy || !d_gz || n_cells == 0 ||
      !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "sif_amd_cic_adjoint: invalid arguments");
    return SIF_ERR_INVALID;
  }
  if (n == 0)
    return SIF_OK;
  const float l = box / (float)n_cells;
  const float idx = (float)n_cells / box;
  // d rho / d x: the weight's slope n / L over the cell volume l^3
  const float scale = idx / (l * l * l);
  cic_adjoint_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(d_x, d_y, d_z, n, n_cells, idx,
    box, scale, d_grad_rho, d_gx, d_gy, d_gz);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // extern "C"
