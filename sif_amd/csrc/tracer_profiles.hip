// Stacked radial profiles of voids over the tracers of a chain mesh:
// sif_profiles (src/measure/profiles.c:354-673) on the GPU.
//
// One wave per void. The wave walks the mesh cells of the sphere's bounding
// box 64 at a time (profiles.c:545-606): each lane places one cell (the
// wrapped index, the image of the centre nearest the cell and the gap from
// the centre to the cell's slab on each axis, profiles.c:261-300) and drops
// it by the same three gap tests the reference applies to the slab, the row
// and the cell. The cells left are read one at a time, 64 tracers a step,
// with the cell's fields uniform over the wave. Each tracer is binned with
// the reference's float operations (profiles.c:184-246), so a tracer lands
// in the same bin as there.
//
// Sums: each lane adds into its own column of LDS (counts as integers,
// weights and velocities in double), in an order fixed by the cell walk,
// and a bin's columns are added lane by lane. A count is exact, so an
// unweighted density profile is the reference's to the bit (its float sum
// of ones is exact below 2^24 tracers a bin); weighted sums and velocities
// differ from the reference's running float sums by their rounding.
//
// A second kernel turns the sums into profiles with the reference's float
// formulas, bin by bin (profiles.c:608-645).
#include <cstdint>

#include "common.h"
#include "internal.h"
#include "sif_amd.h"
#include "sif/measure/profiles.h"

namespace sif_amd {

namespace {

constexpr uint32_t TP_BINS = 32;  // bins a launch accumulates
constexpr int TP_STEPS = 4;       // 64-tracer steps a wave has in flight

struct TracerParams {
  const uint64_t* offsets;
  const float* mx;
  const float* my;
  const float* mz;
  const float* mw;  // NULL: unweighted
  const float* vx;  // NULL: no velocities
  const float* vy;
  const float* vz;
  uint32_t mesh_n;
  float box, half_box, inv_box, cell_length, half_cell, inv_cell_len;
  int periodic;
  const float* cx;
  const float* cy;
  const float* cz;
  const float* rv;
  uint32_t n_voids;
  float ext, bin_width;
  uint32_t n_bins, b0, nbc;
  int vel_weighted;
  uint32_t* count;  // [void][bin]
  double* mass;     // [void][bin], weighted only
  double* vrad;     // [void][bin], velocities only
};

__device__ __forceinline__ void tp_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t tp_uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ float tp_unif(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ uint32_t tp_lane_u32(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float tp_lane_f32(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// profiles.c:302-330: the inclusive cell range of the sphere on one axis,
// in double, at most one full row when periodic, clipped to the box when
// open.
__device__ __forceinline__ void axis_range(float centre, float r_max, float inv_cell_len,
  uint32_t n, int periodic, int32_t* lo_out, int32_t* hi_out) {
  double lo = floor(((double)centre - (double)r_max) * (double)inv_cell_len);
  double hi = floor(((double)centre + (double)r_max) * (double)inv_cell_len);
  if (periodic) {
    if (hi - lo + 1.0 >= (double)n) {
      lo = 0.0;
      hi = (double)n - 1.0;
    }
  } else {
    if (lo < 0.0)
      lo = 0.0;
    if (hi > (double)n - 1.0)
      hi = (double)n - 1.0;
  }
  *lo_out = (int32_t)lo;
  *hi_out = (int32_t)hi;
}

// profiles.c:261-300: cell idx (off by at most one row) on one axis: its
// wrapped index, the image of the centre it is measured from, and the gap
// from the centre to its slab (0 inside it).
__device__ __forceinline__ uint32_t cell_image(int32_t idx, float centre, const TracerParams& P,
  float* image_out, float* gap_out) {
  int32_t w = idx;
  if (P.periodic) {
    if (w < 0)
      w += (int32_t)P.mesh_n;
    else if (w >= (int32_t)P.mesh_n)
      w -= (int32_t)P.mesh_n;
  }
  float gap = __fsub_rn(__fmul_rn(__fadd_rn((float)w, 0.5f), P.cell_length), centre);
  float image = centre;
  if (P.periodic) {
    const float k = roundf(__fmul_rn(gap, P.inv_box));
    gap = __fsub_rn(gap, __fmul_rn(k, P.box));
    image = __fadd_rn(image, __fmul_rn(k, P.box));
  }
  gap = __fsub_rn(fabsf(gap), P.half_cell);
  *image_out = image;
  *gap_out = gap > 0.0f ? gap : 0.0f;
  return (uint32_t)w;
}

// profiles.c:211-219: the branchless nearest-image wrap of one separation.
__device__ __forceinline__ float wrap_sep(float d, float box, float half_box) {
  const float up = d > half_box ? 1.0f : 0.0f;
  const float down = d < -half_box ? 1.0f : 0.0f;
  return __fsub_rn(d, __fsub_rn(__fmul_rn(box, up), __fmul_rn(box, down)));
}

template <bool WT, bool VEL>
__global__ void __launch_bounds__(64) tracer_profile_kernel(TracerParams P) {
  extern __shared__ __align__(16) uint32_t tp_lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t c = blockIdx.x;
  if (c >= P.n_voids)
    return;
  const float rv = P.rv[c];
  if (!(rv > 0.0f))  // profiles.c:505-510: no radius, no profile
    return;
  const uint32_t nbc = P.nbc;
  uint32_t* s_count = tp_lds;                                  // [nbc][64]
  double* s_mass = reinterpret_cast<double*>(tp_lds + nbc * 64);  // [nbc][64]
  double* s_vrad = s_mass + (WT ? nbc * 64 : 0);                // [nbc][64]
  for (uint32_t j = 0; j < nbc; j++) {
    s_count[j * 64 + lane] = 0;
    if constexpr (WT)
      s_mass[j * 64 + lane] = 0.0;
    if constexpr (VEL)
      s_vrad[j * 64 + lane] = 0.0;
  }

  const float ccx = P.cx[c], ccy = P.cy[c], ccz = P.cz[c];
  const float r_max = __fmul_rn(rv, P.ext);
  const float r_max_sq = __fmul_rn(r_max, r_max);
  const float inv_bin = __fdiv_rn(1.0f, __fmul_rn(rv, P.bin_width));
  // profiles.c:529-530: the centre placed once per cell, or the separation
  // wrapped once per tracer when the box is too small for that
  const bool per_cell_image = !P.periodic || __fadd_rn(r_max, P.cell_length) <= P.half_box;
  int32_t lo[3], hi[3];
  axis_range(ccx, r_max, P.inv_cell_len, P.mesh_n, P.periodic, &lo[0], &hi[0]);
  axis_range(ccy, r_max, P.inv_cell_len, P.mesh_n, P.periodic, &lo[1], &hi[1]);
  axis_range(ccz, r_max, P.inv_cell_len, P.mesh_n, P.periodic, &lo[2], &hi[2]);
  const uint32_t nx = (uint32_t)(hi[0] - lo[0] + 1), ny = (uint32_t)(hi[1] - lo[1] + 1),
                 nz = (uint32_t)(hi[2] - lo[2] + 1);
  const uint64_t n_box = (uint64_t)nx * ny * nz;
  const uint32_t N = P.mesh_n;

  for (uint64_t base = 0; base < n_box; base += 64) {
    const uint64_t t = base + (uint64_t)lane;
    uint32_t cnt = 0, start = 0;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (t < n_box) {
      const uint32_t kz = (uint32_t)(t % nz);
      const uint64_t rest = t / nz;
      const uint32_t ky = (uint32_t)(rest % ny), kx = (uint32_t)(rest / ny);
      float gx, gy, gz;
      const uint32_t wx = cell_image(lo[0] + (int32_t)kx, ccx, P, &ex, &gx);
      const uint32_t wy = cell_image(lo[1] + (int32_t)ky, ccy, P, &ey, &gy);
      const uint32_t wz = cell_image(lo[2] + (int32_t)kz, ccz, P, &ez, &gz);
      // profiles.c:565-591: the slab, row and cell tests, in that order
      const float gx2 = __fmul_rn(gx, gx);
      const float gxy2 = __fadd_rn(gx2, __fmul_rn(gy, gy));
      const bool in = !(gx2 > r_max_sq) && !(gxy2 > r_max_sq) &&
                      !(__fadd_rn(gxy2, __fmul_rn(gz, gz)) > r_max_sq);
      if (in) {
        const uint64_t flat = ((uint64_t)wx * N + wy) * N + wz;
        const uint64_t s = P.offsets[flat], e = P.offsets[flat + 1];
        start = (uint32_t)s;
        cnt = (uint32_t)(e - s);
      }
      if (!per_cell_image) {  // the raw centre, wrapped per tracer
        ex = ccx;
        ey = ccy;
        ez = ccz;
      }
    }
    uint64_t live = (uint64_t)__ballot(cnt != 0);
    live = ((uint64_t)tp_uni((uint32_t)(live >> 32)) << 32) | (uint64_t)tp_uni((uint32_t)live);
    int cur = live ? __builtin_ctzll(live) : 64;
    uint32_t k = 0;
    uint32_t c_cnt = 0, c_start = 0;
    float c_x = 0.0f, c_y = 0.0f, c_z = 0.0f;
    auto take = [&]() {
      if (cur < 64) {
        c_cnt = tp_lane_u32(cnt, cur);
        c_start = tp_lane_u32(start, cur);
        c_x = tp_lane_f32(ex, cur);
        c_y = tp_lane_f32(ey, cur);
        c_z = tp_lane_f32(ez, cur);
      }
    };
    take();
    while (cur < 64) {
      uint32_t st_start[TP_STEPS], st_n[TP_STEPS];
      float st_x[TP_STEPS], st_y[TP_STEPS], st_z[TP_STEPS];
#pragma unroll
      for (int u = 0; u < TP_STEPS; u++) {
        st_n[u] = 0;
        st_start[u] = 0;
        st_x[u] = st_y[u] = st_z[u] = 0.0f;
        if (cur < 64) {
          st_start[u] = c_start + k;
          st_n[u] = c_cnt - k < 64u ? c_cnt - k : 64u;
          st_x[u] = c_x;
          st_y[u] = c_y;
          st_z[u] = c_z;
          k += 64;
          if (k >= c_cnt) {
            live &= live - 1;
            cur = live ? __builtin_ctzll(live) : 64;
            k = 0;
            take();
          }
        }
      }
      float px[TP_STEPS], py[TP_STEPS], pz[TP_STEPS], pw[TP_STEPS];
      float qx[TP_STEPS], qy[TP_STEPS], qz[TP_STEPS];
#pragma unroll
      for (int u = 0; u < TP_STEPS; u++) {
        const uint32_t off = (uint32_t)lane < st_n[u] ? (uint32_t)lane : st_n[u] - 1u;
        px[u] = py[u] = pz[u] = pw[u] = qx[u] = qy[u] = qz[u] = 0.0f;
        if (st_n[u]) {
          px[u] = (P.mx + st_start[u])[off];
          py[u] = (P.my + st_start[u])[off];
          pz[u] = (P.mz + st_start[u])[off];
          if constexpr (WT)
            pw[u] = (P.mw + st_start[u])[off];
          if constexpr (VEL) {
            qx[u] = (P.vx + st_start[u])[off];
            qy[u] = (P.vy + st_start[u])[off];
            qz[u] = (P.vz + st_start[u])[off];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < TP_STEPS; u++) {
        if (!st_n[u] || (uint32_t)lane >= st_n[u])
          continue;
        float dx = __fsub_rn(px[u], st_x[u]);
        float dy = __fsub_rn(py[u], st_y[u]);
        float dz = __fsub_rn(pz[u], st_z[u]);
        if (!per_cell_image) {
          dx = wrap_sep(dx, P.box, P.half_box);
          dy = wrap_sep(dy, P.box, P.half_box);
          dz = wrap_sep(dz, P.box, P.half_box);
        }
        const float r2 =
          __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
        if (!(r2 <= r_max_sq))
          continue;
        const float r = __builtin_sqrtf(r2);  // correctly rounded, as sqrtf
        const uint32_t bin = (uint32_t)__fmul_rn(r, inv_bin);
        if (bin >= P.n_bins || bin < P.b0 || bin - P.b0 >= nbc)
          continue;
        const uint32_t j = (bin - P.b0) * 64 + (uint32_t)lane;
        s_count[j] += 1u;
        if constexpr (WT)
          s_mass[j] += (double)pw[u];
        if constexpr (VEL) {
          const float inv_r = r > 0.0f ? __fdiv_rn(1.0f, r) : 0.0f;
          const float vr = __fmul_rn(
            __fadd_rn(__fadd_rn(__fmul_rn(qx[u], dx), __fmul_rn(qy[u], dy)), __fmul_rn(qz[u], dz)),
            inv_r);
          float add = vr;
          if constexpr (WT)
            add = P.vel_weighted ? __fmul_rn(pw[u], vr) : vr;
          s_vrad[j] += (double)add;
        }
      }
    }
  }
  tp_sync();
  // a bin's columns, lane by lane: lane j adds up bin b0 + j
  for (uint32_t j = (uint32_t)lane; j < nbc; j += 64) {
    uint32_t n = 0;
    double m = 0.0, v = 0.0;
    for (int l = 0; l < 64; l++) {
      n += s_count[j * 64 + l];
      if constexpr (WT)
        m += s_mass[j * 64 + l];
      if constexpr (VEL)
        v += s_vrad[j * 64 + l];
    }
    const uint64_t o = (uint64_t)c * P.n_bins + P.b0 + j;
    P.count[o] = n;
    if constexpr (WT)
      P.mass[o] = m;
    if constexpr (VEL)
      P.vrad[o] = v;
  }
}

// profiles.c:608-645, one thread a void: the profiles from the sums, bin by
// bin, with the reference's float operations.
__global__ void tracer_profile_finish_kernel(const float* __restrict__ rv, uint32_t n_voids,
  uint32_t n_bins, const uint32_t* __restrict__ count, const double* __restrict__ mass,
  const double* __restrict__ vrad, const float* __restrict__ bin_vols, float mean_dens,
  int differential, int weighted, int vel_weighted, float* __restrict__ dens,
  float* __restrict__ vel) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_voids)
    return;
  const float r = rv[c];
  if (!(r > 0.0f))
    return;
  const float rv3 = __fmul_rn(__fmul_rn(r, r), r);
  float cumulative = 0.0f;
  for (uint32_t j = 0; j < n_bins; j++) {
    const uint64_t o = (uint64_t)c * n_bins + j;
    const float local_mass = weighted ? (float)mass[o] : (float)count[o];
    if (dens) {
      cumulative = __fadd_rn(cumulative, local_mass);
      const float weight = differential ? local_mass : cumulative;
      const float raw = __fdiv_rn(weight, __fmul_rn(bin_vols[j], rv3));
      dens[o] = __fsub_rn(__fdiv_rn(raw, mean_dens), 1.0f);
    }
    if (vel) {
      const float v = (float)vrad[o];
      if (vel_weighted)
        vel[o] = local_mass > 0.0f ? __fdiv_rn(v, local_mass) : 0.0f;
      else
        vel[o] = count[o] > 0 ? __fdiv_rn(v, (float)count[o]) : 0.0f;
    }
  }
}

thread_local DeviceBuffer g_tp_count, g_tp_mass, g_tp_vrad, g_tp_vols;

}  // namespace

void tracer_profile_spec(TracerProfileSpec* S, uint32_t mesh_n, float box, double total_weight,
  float ext, uint32_t n_bins, sif_option opt, bool weighted, bool want_vel) {
  // profiles.c:382-383: anything but a positive ext asks for the default
  if (!(ext > 0.0f))
    ext = SIF_PROFILES_DEFAULT_EXT;
  S->mesh_n = mesh_n;
  S->box = box;
  S->cell_length = box / (float)mesh_n;
  S->periodic = (opt & SIF__PBC_MASK) == SIF_PBC_PERIODIC;
  S->ext = ext;
  S->n_bins = n_bins;
  S->bin_width = ext / (float)n_bins;
  S->differential = (opt & SIF__PROFILES_BIN_MASK) == SIF_PROFILES_DIFFERENTIAL;
  // profiles.c:448-451: the weighted mean differs only where there are weights
  S->vel_weighted = want_vel && weighted &&
                    (opt & SIF__PROFILES_VELOCITY_MASK) == SIF_PROFILES_VELOCITY_WEIGHTED;
  // profiles.c:144-150: the mesh's total weight over the box, in double
  const double box_vol = (double)box * (double)box * (double)box;
  S->mean_dens = (float)(total_weight / box_vol);
  // profiles.c:332-352: edges in void radii, and the volume a bin's weight is
  // spread through -- the sphere to its outer edge, or its shell
  S->edges.assign(n_bins + 1, 0.0f);
  S->vols.assign(n_bins, 0.0f);
  for (uint32_t j = 0; j < n_bins; j++) {
    S->edges[j] = (float)j * S->bin_width;
    const float outer = (float)(j + 1) * S->bin_width;
    const float inner = S->differential ? S->edges[j] : 0.0f;
    S->vols[j] = (float)(4.0 / 3.0) * SIF_PI * (outer * outer * outer - inner * inner * inner);
  }
  S->edges[n_bins] = ext;
  S->bin_vols = S->vols.data();
}

int tracer_profiles(const TracerProfileSpec& S) {
  const uint64_t cells = (uint64_t)S.n_voids * S.n_bins;
  int st = g_tp_count.reserve((cells ? cells : 1) * sizeof(uint32_t));
  if (st == SIF_OK && S.mw)
    st = g_tp_mass.reserve((cells ? cells : 1) * sizeof(double));
  if (st == SIF_OK && S.vx)
    st = g_tp_vrad.reserve((cells ? cells : 1) * sizeof(double));
  if (st == SIF_OK)
    st = g_tp_vols.reserve(S.n_bins * sizeof(float));
  if (st != SIF_OK)
    return st;
  if (S.n_voids == 0)
    return SIF_OK;
  SIF_HIP(hipMemcpyAsync(g_tp_vols.ptr, S.bin_vols, S.n_bins * sizeof(float),
    hipMemcpyHostToDevice, stream()));
  TracerParams P;
  P.offsets = S.offsets;
  P.mx = S.mx;
  P.my = S.my;
  P.mz = S.mz;
  P.mw = S.mw;
  P.vx = S.vx;
  P.vy = S.vy;
  P.vz = S.vz;
  P.mesh_n = S.mesh_n;
  P.box = S.box;
  P.half_box = S.box * 0.5f;
  P.inv_box = 1.0f / S.box;
  P.cell_length = S.cell_length;
  P.half_cell = S.cell_length * 0.5f;
  P.inv_cell_len = 1.0f / S.cell_length;
  P.periodic = S.periodic;
  P.cx = S.cx;
  P.cy = S.cy;
  P.cz = S.cz;
  P.rv = S.rv;
  P.n_voids = S.n_voids;
  P.ext = S.ext;
  P.bin_width = S.bin_width;
  P.n_bins = S.n_bins;
  P.vel_weighted = S.vel_weighted;
  P.count = g_tp_count.as<uint32_t>();
  P.mass = S.mw ? g_tp_mass.as<double>() : nullptr;
  P.vrad = S.vx ? g_tp_vrad.as<double>() : nullptr;
  timing_begin(SIF_AMD_STAGE_PROFILES);
  for (uint32_t b0 = 0; b0 < S.n_bins; b0 += TP_BINS) {
    P.b0 = b0;
    P.nbc = S.n_bins - b0 < TP_BINS ? S.n_bins - b0 : TP_BINS;
    const size_t lds = (size_t)P.nbc * 64 *
                       (sizeof(uint32_t) + (S.mw ? sizeof(double) : 0) + (S.vx ? sizeof(double) : 0));
    if (S.mw && S.vx)
      tracer_profile_kernel<true, true><<<S.n_voids, 64, lds, stream()>>>(P);
    else if (S.mw)
      tracer_profile_kernel<true, false><<<S.n_voids, 64, lds, stream()>>>(P);
    else if (S.vx)
      tracer_profile_kernel<false, true><<<S.n_voids, 64, lds, stream()>>>(P);
    else
      tracer_profile_kernel<false, false><<<S.n_voids, 64, lds, stream()>>>(P);
    SIF_HIP(hipGetLastError());
  }
  tracer_profile_finish_kernel<<<grid_for(S.n_voids, 256), 256, 0, stream()>>>(S.rv, S.n_voids,
    S.n_bins, P.count, P.mass, P.vrad, g_tp_vols.as<float>(), S.mean_dens, S.differential,
    S.mw != nullptr, S.vel_weighted, S.dens, S.vel);
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_PROFILES);
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_tracer_profiles(const uint64_t* d_offsets, const float* d_mx, const float* d_my,
  const float* d_mz, const float* d_mw, const float* d_vx, const float* d_vy,
  const float* d_vz, uint32_t mesh_cells, float box, double total_weight, const float* d_cx,
  const float* d_cy, const float* d_cz, const float* d_rv, uint32_t n_voids, float ext,
  uint32_t n_bins, sif_option options, float* d_dens, float* d_vel) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_offsets || !d_mx || !d_my || !d_mz || mesh_cells == 0 || !(box > 0.0f) ||
      n_bins == 0 || (!d_dens && !d_vel) || (n_voids && (!d_cx || !d_cy || !d_cz || !d_rv)) ||
      (d_vel && (!d_vx || !d_vy || !d_vz)) || !(total_weight > 0.0)) {
    error_set(SIF_ERR_INVALID, "sif_amd_tracer_profiles: invalid arguments");
    return SIF_ERR_INVALID;
  }
  TracerProfileSpec S;
  tracer_profile_spec(&S, mesh_cells, box, total_weight, ext, n_bins, options,
    d_mw != nullptr, d_vel != nullptr);
  S.offsets = d_offsets;
  S.mx = d_mx;
  S.my = d_my;
  S.mz = d_mz;
  S.mw = d_mw;
  S.vx = d_vel ? d_vx : nullptr;
  S.vy = d_vel ? d_vy : nullptr;
  S.vz = d_vel ? d_vz : nullptr;
  S.cx = d_cx;
  S.cy = d_cy;
  S.cz = d_cz;
  S.rv = d_rv;
  S.n_voids = n_voids;
  S.dens = d_dens;
  S.vel = d_vel;
  st = tracer_profiles(S);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
