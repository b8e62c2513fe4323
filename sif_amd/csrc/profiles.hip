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
#include <vector>

#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {

constexpr uint32_t PROFILE_THREADS = 256;
constexpr uint32_t PROFILE_TILE = 16;  // adjoint tile side, in nodes
constexpr uint32_t PROFILE_CHUNK = 32;  // centres staged in LDS at a time
constexpr uint32_t PROFILE_ROWS = 8;    // rows a wave of the gather has in flight

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

// One workgroup per centre: the mass (1 + delta) per bin over the nodes of
// its sphere, then the cumulative profile. Each lane keeps its own histogram
// in LDS (bin-major, so the lanes of a wave hit distinct words) and updates it
// with plain loads and stores: no lane writes another's words, so no atomics.
// A lane's float sum covers the few thousand nodes it visits (its rounding
// stays near 1e-7 of the sum); the lanes' sums are added in double, in lane
// order.
template <int NB>
__global__ void __launch_bounds__(PROFILE_THREADS) profile_kernel(
  const float* __restrict__ delta, uint32_t N, float l, float box,
  const float* __restrict__ cxs, const float* __restrict__ cys, const float* __restrict__ czs,
  const float* __restrict__ rvs, float ext, uint32_t n_bins, float bw,
  float* __restrict__ out) {
  __shared__ float s_mass[NB][PROFILE_THREADS];
  __shared__ double s_tot[NB];
  const uint32_t c = blockIdx.x;
  const float rv = rvs[c];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!(rv > 0.0f)) {
    for (uint32_t j = threadIdx.x; j < n_bins; j += blockDim.x)
      out[(uint64_t)c * n_bins + j] = 0.0f;
    return;
  }
#pragma unroll
  for (int j = 0; j < NB; j++)
    s_mass[j][threadIdx.x] = 0.0f;
  const float cx = cxs[c], cy = cys[c], cz = czs[c];
  const float r_max = rv * ext;
  const float r_max_sq = r_max * r_max;
  const float inv_bin = 1.0f / (rv * bw);
  const float half = box * 0.5f;
  const Cube k = centre_cube(cx, cy, cz, r_max, l, N);
  // the lane's z offsets in the row, wrapped once (span <= N: one step)
  const uint32_t z_base = wrap_node(k.lo[2], (int32_t)N);
  const uint32_t rows = (uint32_t)k.span[0] * (uint32_t)k.span[1];
  const uint32_t gx0 = wrap_node(k.lo[0], (int32_t)N), gy0 = wrap_node(k.lo[1], (int32_t)N);
  float* my_mass = &s_mass[0][threadIdx.x];
  // PROFILE_ROWS rows of the cube per wave at a time, their loads issued
  // together; the waves take interleaved groups of rows
  const uint32_t span_y = (uint32_t)k.span[1];
  const uint32_t wave_u = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
  for (uint32_t row0 = wave_u * PROFILE_ROWS; row0 < rows;
       row0 += (PROFILE_THREADS / 64) * PROFILE_ROWS) {
    uint64_t rowo[PROFILE_ROWS];  // row offsets into delta (kept global loads)
    float dxy[PROFILE_ROWS];
    bool live = false;
    // the group's first (x, y) node once, then step along y (uniform values:
    // scalar arithmetic)
    uint32_t ax = row0 / span_y, ay = row0 - ax * span_y;
#pragma unroll
    for (uint32_t r = 0; r < PROFILE_ROWS; r++) {
      const uint32_t row = row0 + r;
      dxy[r] = r_max_sq + 1.0f;  // past the sphere: skipped
      rowo[r] = 0;
      if (row < rows) {
        uint32_t gx = gx0 + ax, gy = gy0 + ay;
        gx = gx >= N ? gx - N : gx;
        gy = gy >= N ? gy - N : gy;
        const float dx = nearest((float)gx * l - cx, box, half);
        const float dy = nearest((float)gy * l - cy, box, half);
        dxy[r] = dx * dx + dy * dy;
        rowo[r] = ((uint64_t)gx * N + gy) * N;
        live |= !(dxy[r] > r_max_sq);
      }
      if (++ay == span_y) {
        ay = 0;
        ax++;
      }
    }
    if (!live)
      continue;
    for (int32_t z0 = 0; z0 < k.span[2]; z0 += 64) {
      const int32_t z = z0 + lane;
      uint32_t gz = z_base + (uint32_t)z;
      gz = gz >= N ? gz - N : gz;
      const bool in_row = z < k.span[2];
      const float dz = nearest((float)gz * l - cz, box, half);
      const float dz2 = dz * dz;
      // every lane loads (an in-bounds node where it has none), so the loads
      // issue together without a branch between them
      const uint32_t gzl = in_row ? gz : 0u;
      float v[PROFILE_ROWS];
#pragma unroll
      for (uint32_t r = 0; r < PROFILE_ROWS; r++)
        v[r] = delta[rowo[r] + gzl];
#pragma unroll
      for (uint32_t r = 0; r < PROFILE_ROWS; r++) {
        const float r2 = dxy[r] + dz2;
        if (in_row && r2 <= r_max_sq) {
          const uint32_t bin = (uint32_t)(__builtin_sqrtf(r2) * inv_bin);
          if (bin < n_bins)
            my_mass[bin * PROFILE_THREADS] += 1.0f + v[r];
        }
      }
    }
  }
  __syncthreads();
  // bin j: the 256 lanes' sums in double, in lane order
  if (threadIdx.x < NB) {
    const uint32_t j = threadIdx.x;
    double m = 0.0;
    for (uint32_t t = 0; t < PROFILE_THREADS; t++)
      m += (double)s_mass[j][t];
    s_tot[j] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double l3 = (double)l * (double)l * (double)l;
    const double rv3 = (double)rv * (double)rv * (double)rv;
    double cum = 0.0;
    for (uint32_t j = 0; j < n_bins; j++) {
      cum += s_tot[j];
      const double mass = l3 * cum;
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

// loss = sum (p - t)^2 and its gradient scale * 2 (p - t): per-block sums in
// double in a fixed tree, added in block order on the host
constexpr uint32_t RESID_THREADS = 256, RESID_PER_BLOCK = 8192;

__global__ void __launch_bounds__(RESID_THREADS) residual_kernel(const float* __restrict__ p,
  const float* __restrict__ t, uint64_t n, float scale, float* __restrict__ grad,
  double* __restrict__ block_sums) {
  __shared__ double s[RESID_THREADS];
  const uint64_t b0 = (uint64_t)blockIdx.x * RESID_PER_BLOCK;
  double acc = 0.0;
  for (uint32_t k = threadIdx.x; k < RESID_PER_BLOCK; k += RESID_THREADS) {
    const uint64_t i = b0 + k;
    if (i < n) {
      const float d = p[i] - t[i];
      acc += (double)d * (double)d;
      if (grad)
        grad[i] = scale * (2.0f * d);
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t off = RESID_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    block_sums[blockIdx.x] = s[0];
}

thread_local DeviceBuffer g_pair_count, g_pair_offset, g_pair_keys, g_pair_vals, g_tile_start, g_tile_end;
thread_local DeviceBuffer g_resid;

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

int sif_amd_squared_residual(const float* d_p, const float* d_target, uint64_t n, float scale,
  float* d_grad, double* loss) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_p || !d_target || !loss) {
    error_set(SIF_ERR_INVALID, "sif_amd_squared_residual: invalid arguments");
    return SIF_ERR_INVALID;
  }
  *loss = 0.0;
  if (n == 0)
    return SIF_OK;
  const uint64_t blocks = (n + RESID_PER_BLOCK - 1) / RESID_PER_BLOCK;
  st = g_resid.reserve(blocks * sizeof(double));
  if (st != SIF_OK)
    return st;
  residual_kernel<<<(unsigned)blocks, RESID_THREADS, 0, stream()>>>(d_p, d_target, n, scale,
    d_grad, g_resid.as<double>());
  SIF_HIP(hipGetLastError());
  std::vector<double> sums(blocks);
  SIF_HIP(hipMemcpyAsync(sums.data(), g_resid.ptr, blocks * sizeof(double),
    hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  double total = 0.0;
  for (double v : sums)
    total += v;
  *loss = total;
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
