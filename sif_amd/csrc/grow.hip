// Sphere growth (the exodus radius rescaling) and the candidate prefilter.
//
// Replaces sif__exodus_find_radius / find_exact_radius_impl / resolve_bin /
// query_cell (src/finder/exodus_rescale.c:295-369, 391-507, 642-1012) and the
// rung-radius overlap tests of prefilter_cells (src/finder/exodus_finder.c:
// 871-944, src/finder/utils.c:819-960) of sif v0.5.0.
//
// One 64-lane wavefront grows one candidate; four per workgroup, each with its
// own LDS region and no workgroup barriers, so waves retire independently and
// take the next candidate from a global counter. The per-candidate arithmetic
// (distances, bins, the skip test, the crossing test, the early refusal) is the
// reference's, float for float, with no contraction; integer counts are exact.
//
// Particle distribution: the template is walked 64 cells at a time. Each lane
// resolves one cell to its particle range; a wave-wide prefix sum of the
// counts then hands particle j of the chunk to lane j % 64, which finds its
// cell by binary search over the prefix -- every lane busy, and consecutive
// lanes reading consecutive particles of a cell.
//
// The crossing is resolved over a window of bins whose particles fit the LDS
// buffer (the reference resolves a window of W bins too, and its result does
// not depend on W). A single bin too full for the buffer is split into 64
// sub-ranges of squared distance, resolved top down.
#include <memory>
#include <vector>

#include <atomic>

#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"
#include "grow_device.h"
#include "grow_survey.h"
#include "overlap.h"
#include "paint.h"

namespace sif_amd {

// --- prefilter: the rung-radius tests of exodus_finder.c:871-944 ---

__device__ __forceinline__ uint32_t wrap_pbc(int32_t v, uint32_t n) {
  const int32_t w = v % (int32_t)n;
  return (uint32_t)(w < 0 ? w + (int32_t)n : w);
}

__device__ __forceinline__ bool mask_bit(const uint64_t* m, uint64_t i) {
  return (m[i >> 6] >> (i & 63)) & 1ull;
}

// utils.c:819-848: any of the six nodes r away along the axes in the mask.
// The six words are loaded together (the mask is far larger than the caches,
// so each is a trip to memory) and then tested; the answer is their OR.
__device__ bool overlap_quick(const uint64_t* mask, uint32_t n, uint32_t ix,
  uint32_t iy, uint32_t iz, uint32_t r) {
  const uint64_t nn = n;
  const uint64_t node[6] = {
    ((uint64_t)wrap_pbc((int32_t)(ix + r), n) * nn + iy) * nn + iz,
    ((uint64_t)wrap_pbc((int32_t)ix - (int32_t)r, n) * nn + iy) * nn + iz,
    ((uint64_t)ix * nn + wrap_pbc((int32_t)(iy + r), n)) * nn + iz,
    ((uint64_t)ix * nn + wrap_pbc((int32_t)iy - (int32_t)r, n)) * nn + iz,
    ((uint64_t)ix * nn + iy) * nn + wrap_pbc((int32_t)(iz + r), n),
    ((uint64_t)ix * nn + iy) * nn + wrap_pbc((int32_t)iz - (int32_t)r, n)};
  uint64_t w[6];
#pragma unroll
  for (int p = 0; p < 6; p++)
    w[p] = mask[node[p] >> 6];
  uint64_t hit = 0;
#pragma unroll
  for (int p = 0; p < 6; p++)
    hit |= (w[p] >> (node[p] & 63)) & 1ull;
  return hit != 0;
}

// The overlap state (overlap.h) of a query against the CSR void index: the
// index lists each void under the reference's cell, so a scan of the
// superset cells sees every void that could collide.
// The cells of one wrapped row (32 along z) that the z range [lo, hi] of a
// scan covers, as a mask over the row's occupancy word (hi - lo < 32, or
// the whole axis).
__device__ __forceinline__ uint32_t row_zmask(int32_t lo, int32_t hi) {
  const uint32_t span = (uint32_t)(hi - lo);
  if (span + 1 >= (uint32_t)VOID_CELLS)
    return 0xffffffffu;
  const uint64_t m = ((1ull << (span + 1)) - 1ull) << void_wrap(lo);
  return (uint32_t)(m | (m >> VOID_CELLS));
}

// The index lists a row's 32 cells along z one after another, so the voids of
// the cells a scan covers in a row are one stretch of the index (two where
// the z range wraps): a scan reads the row's occupancy word, skips an empty
// row, and otherwise reads that stretch four voids at a time. A void's z cell
// is that of its centre (void_axis_cell), the cell the index lists it under.
// The state is an OR over voids, so the order they come in changes nothing.
struct RowSpans {
  uint32_t a0, b0, a1, b1;  // [a0, b0) and [a1, b1) of the row's z cells
};
__device__ __forceinline__ RowSpans row_spans(int32_t lo, int32_t hi) {
  RowSpans r{0, 0, 0, 0};
  const uint32_t span = (uint32_t)(hi - lo);
  if (span + 1 >= (uint32_t)VOID_CELLS) {
    r.b0 = VOID_CELLS;
    return r;
  }
  const uint32_t w = void_wrap(lo);
  r.a0 = w;
  if (w + span < (uint32_t)VOID_CELLS) {
    r.b0 = w + span + 1;
  } else {
    r.b0 = VOID_CELLS;
    r.b1 = w + span + 1 - VOID_CELLS;
  }
  return r;
}

// The voids [k0, k1) of the index against q (row wa, wb): 1 at the first
// state 1, else the OR of their states.
__device__ __forceinline__ int stretch_state(const VoidIndex& V, const OverlapQuery& q,
  uint32_t wa, uint32_t wb, uint32_t k0, uint32_t k1) {
  int state = 0;
  for (uint32_t k = k0; k < k1; k += 4) {
    float4 v[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; u++)
      v[u] = k + u < k1 ? V.voids[k + u] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
      if (k + u < k1) {
        const uint32_t wc = void_axis_cell(v[u].z, V.inv_l);
        const int s = pair_state(q, wa, wb, wc, v[u].x, v[u].y, v[u].z, v[u].w);
        if (s == 1)
          return 1;
        state |= s;
      }
    }
  }
  return state;
}

__device__ int overlap_state(const VoidIndex& V, const OverlapQuery& q) {
  int state = 0;
  for (int cls = VOID_CLASSES - 1; cls >= 0; cls--) {
    if (V.class_max_r[cls] < 0.0f)
      continue;
    int32_t lo[3], hi[3];
    superset_scan(q, V.class_max_r[cls], V.inv_l, lo, hi);
    const uint32_t* start = V.cell_start + (uint32_t)cls * VOID_CELLS * VOID_CELLS * VOID_CELLS;
    const uint32_t* occ = V.occupied + (uint32_t)cls * VOID_CELLS * VOID_CELLS;
    const uint32_t zmask = row_zmask(lo[2], hi[2]);
    const RowSpans z = row_spans(lo[2], hi[2]);
    const uint32_t nb = (uint32_t)(hi[1] - lo[1] + 1);
    const uint32_t rows = (uint32_t)(hi[0] - lo[0] + 1) * nb;
    // four rows at a time: their occupancy words, then the stretch bounds of
    // the occupied ones, each set loaded together, then the stretches
    for (uint32_t r0 = 0; r0 < rows; r0 += 4) {
      uint32_t wa[4], wb[4], o[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t idx = r0 + u < rows ? r0 + u : r0;
        wa[u] = void_wrap(lo[0] + (int32_t)(idx / nb));
        wb[u] = void_wrap(lo[1] + (int32_t)(idx % nb));
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; u++)
        o[u] = r0 + u < rows ? occ[wa[u] * VOID_CELLS + wb[u]] & zmask : 0u;
      uint32_t k0[4], k1[4], k2[4], k3[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t* rs = start + (wa[u] * VOID_CELLS + wb[u]) * VOID_CELLS;
        k0[u] = o[u] ? rs[z.a0] : 0u;
        k1[u] = o[u] ? rs[z.b0] : 0u;
        k2[u] = o[u] && z.b1 > z.a1 ? rs[z.a1] : 0u;
        k3[u] = o[u] && z.b1 > z.a1 ? rs[z.b1] : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        if (!o[u])
          continue;
        const int s0 = stretch_state(V, q, wa[u], wb[u], k0[u], k1[u]);
        if (s0 == 1)
          return 1;
        state |= s0;
        const int s1 = stretch_state(V, q, wa[u], wb[u], k2[u], k3[u]);
        if (s1 == 1)
          return 1;
        state |= s1;
      }
    }
  }
  return state;
}

// overlap_state() by a whole wavefront for one query: the lanes split the
// rows of each class's superset cells. The wave stops at the first state 1.
// Every lane returns the same state.
__device__ int wave_overlap_state(const VoidIndex& V, const OverlapQuery& q, uint32_t lane) {
  int st = 0;
  for (int cls = VOID_CLASSES - 1; cls >= 0; cls--) {
    if (V.class_max_r[cls] < 0.0f)
      continue;
    int32_t lo[3], hi[3];
    superset_scan(q, V.class_max_r[cls], V.inv_l, lo, hi);
    const uint32_t na = (uint32_t)(hi[0] - lo[0] + 1), nb = (uint32_t)(hi[1] - lo[1] + 1);
    const uint32_t* start = V.cell_start + (uint32_t)cls * VOID_CELLS * VOID_CELLS * VOID_CELLS;
    const uint32_t* occ = V.occupied + (uint32_t)cls * VOID_CELLS * VOID_CELLS;
    const uint32_t zmask = row_zmask(lo[2], hi[2]);
    const RowSpans z = row_spans(lo[2], hi[2]);
    for (uint32_t base = 0; base < na * nb; base += 64) {
      const uint32_t idx = base + lane;
      if (idx < na * nb && st != 1) {
        const uint32_t wa = void_wrap(lo[0] + (int32_t)(idx / nb));
        const uint32_t wb = void_wrap(lo[1] + (int32_t)(idx % nb));
        const uint32_t row = wa * VOID_CELLS + wb;
        if (occ[row] & zmask) {
          const uint32_t* rs = start + row * VOID_CELLS;
          const int s0 = stretch_state(V, q, wa, wb, rs[z.a0], rs[z.b0]);
          const int s1 = s0 != 1 && z.b1 > z.a1
                           ? stretch_state(V, q, wa, wb, rs[z.a1], rs[z.b1]) : 0;
          st = (s0 == 1 || s1 == 1) ? 1 : (st | s0 | s1);
        }
      }
      if (__ballot(st == 1))
        return 1;
    }
  }
  return __ballot(st == 2) ? 2 : 0;
}

// keep[k] = 0 refused (masked node, pole probe, or an overlap at the rung
// radius with the catalogue as the rung began, decided as the reference
// decides it), 1 kept. A kept candidate that some void collides with under
// the periodic distance but outside the reference's scan (overlap state 2)
// gets bit 31 of its flat index set: the host re-tests it against the whole
// catalogue.
// The refusal counts of a workgroup, summed by wave and then by workgroup:
// one atomic per workgroup and reason (same-address atomics from every wave
// of a launch of millions of candidates serialise at the L2 and held the
// launch's end).
template <int R>
__device__ __forceinline__ void block_add_counts(const uint32_t (&c)[R],
  unsigned long long* __restrict__ out) {
  __shared__ uint32_t part[R][256 / 64];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < R; r++) {
    uint32_t v = c[r];
    for (int off = 32; off > 0; off >>= 1)
      v += __shfl_xor(v, off, 64);
    if (lane == 0)
      part[r][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    uint64_t t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++)
      t += part[threadIdx.x][w];
    if (t)
      atomicAdd(&out[threadIdx.x], (unsigned long long)t);
  }
}

// Workgroups of the candidate kernels below: enough to fill the device, each
// striding over its share of the candidates (consecutive threads, consecutive
// candidates).
constexpr uint32_t CAND_BLOCKS = 2048;

__global__ void __launch_bounds__(256) prefilter_kernel(uint32_t* __restrict__ flat,
  uint32_t n, uint32_t grid_n, float l, float box, const uint64_t* __restrict__ mask,
  int32_t proxy, float radius, float f, float scale, VoidIndex V,
  uint32_t* __restrict__ keep, unsigned long long* __restrict__ counts,
  uint8_t* __restrict__ why_out) {
  uint32_t refused[3] = {0u, 0u, 0u};  // masked, proxy, overlap
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    uint32_t why = 0;  // 0 kept, 1 masked, 2 proxy, 3 overlap
    const uint32_t node = flat[k];
    bool edge = false;
    const uint32_t ix = node / (grid_n * grid_n);
    const uint32_t iy = (node / grid_n) % grid_n;
    const uint32_t iz = node % grid_n;
    // the node's word and, with a pole distance, the poles' words loaded
    // together; the tests then decide in the reference's order
    const uint64_t own = mask[node >> 6];
    const bool poles = proxy > 0 && overlap_quick(mask, grid_n, ix, iy, iz, (uint32_t)proxy);
    if ((own >> (node & 63)) & 1ull) {
      why = 1;
    } else if (poles) {
      why = 2;
    } else if (V.n) {
      OverlapQuery q;
      overlap_query(q, (float)ix * l, (float)iy * l, (float)iz * l, radius, V.max_r, box,
        f, scale, V.inv_l);
      const int s = overlap_state(V, q);
      if (s == 1)
        why = 3;
      else
        edge = s == 2;
    }
    keep[k] = why == 0 ? 1u : 0u;
    if (why_out)
      why_out[k] = (uint8_t)why;
    if (edge)
      flat[k] = node | 0x80000000u;
    if (why)
      refused[why - 1]++;
  }
  block_add_counts<3>(refused, counts);
}

// The overlap state of each grown sphere (radius >= 0) against the catalogue
// as the rung began, for a reference max_r of the catalogue as the batch
// began (exodus_finder.c:2302-2311, the test at the rescaled radius). One
// wavefront per candidate: a grown sphere reaches many void cells, and the
// lanes split them (wave_overlap_state).
constexpr int GO_WAVES = 4;

__global__ void __launch_bounds__(64 * GO_WAVES) grown_overlap_kernel(
  const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz,
  const uint32_t* __restrict__ flat, const float* __restrict__ radius, uint32_t n,
  VoidIndex V, float max_r, float box, float f, float scale, const uint64_t* __restrict__ mask,
  uint32_t grid_n, float l, uint32_t* __restrict__ state) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t k = blockIdx.x * GO_WAVES + (threadIdx.x >> 6);
  if (k >= n)
    return;
  const float r = radius[k];
  int st = 0;
  // the pole probe at the rescaled radius against the mask as the rung
  // began (exodus_finder.c:2290-2300)
  bool pole_hit = false;
  if (r >= 0.0f && lane == 0) {
    const int32_t pole = (int32_t)(scale * r * (1.0f - f) / l) - 1;
    const uint32_t node = flat[k];
    if (pole > 0)
      pole_hit = overlap_quick(mask, grid_n, node / (grid_n * grid_n), (node / grid_n) % grid_n,
        node % grid_n, (uint32_t)pole);
  }
  if (r >= 0.0f && V.n) {  // r is the wave's own: uniform
    OverlapQuery q;
    overlap_query(q, cx[k], cy[k], cz[k], r, max_r, box, f, scale, V.inv_l);
    st = wave_overlap_state(V, q, lane);
  }
  if (lane == 0)
    state[k] = (uint32_t)st | (pole_hit ? GROWN_POLE_BIT : 0u);
}

// sif__mark_sphere (utils.c:965-1033) for a list of voids into the device
// mask: one workgroup per void walks the reference's cube of nodes, a row
// (dx, dy) per thread, and ORs each 64-bit word's bits in with one atomic.
// Setting bits is idempotent, so the mask does not depend on the order.
__global__ void __launch_bounds__(256) paint_voids_kernel(const float* __restrict__ vx,
  const float* __restrict__ vy, const float* __restrict__ vz, const float* __restrict__ vr,
  uint32_t nv, float scale, float l, uint32_t grid_n, unsigned long long* __restrict__ mask) {
  const uint32_t v = blockIdx.x;
  if (v >= nv)
    return;
  PaintSphere P;
  paint_sphere(P, vx[v], vy[v], vz[v], scale * vr[v], l, grid_n);
  const int32_t span = 2 * P.c + 1;
  const uint64_t N = grid_n;
  for (int32_t row = (int32_t)threadIdx.x; row < span * span; row += (int32_t)blockDim.x) {
    const int32_t gx = P.ox + row / span - P.c, gy = P.oy + row % span - P.c;
    const float ddx = paint_dist(P, gx, P.cx);
    const float dx2 = ddx * ddx;
    if (dx2 > P.r2)
      continue;
    const float ddy = paint_dist(P, gy, P.cy);
    const float dxy2 = dx2 + ddy * ddy;
    if (dxy2 > P.r2)
      continue;
    const uint64_t base = ((uint64_t)paint_wrap(gx, P.n) * N + paint_wrap(gy, P.n)) * N;
    uint64_t word = ~0ull, bits = 0;
    for (int32_t gz = P.oz - P.c; gz <= P.oz + P.c; gz++) {
      const float ddz = paint_dist(P, gz, P.cz);
      if (!(dxy2 + ddz * ddz <= P.r2))
        continue;
      const uint64_t idx = base + paint_wrap(gz, P.n);
      if ((idx >> 6) != word) {
        if (bits)
          atomicOr(&mask[word], (unsigned long long)bits);
        word = idx >> 6;
        bits = 0;
      }
      bits |= 1ull << (idx & 63);
    }
    if (bits)
      atomicOr(&mask[word], (unsigned long long)bits);
  }
}

// Candidate nodes -> centres (exodus_finder.c:386-405 for a centre on its
// node: (float)i * cell_length).
__global__ void node_centres_kernel(const uint32_t* __restrict__ flat, uint32_t n,
  uint32_t grid_n, float l, float* __restrict__ cx, float* __restrict__ cy,
  float* __restrict__ cz) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n)
    return;
  const uint64_t node = flat[k], nn = grid_n;
  cx[k] = __fmul_rn((float)(uint32_t)(node / (nn * nn)), l);
  cy[k] = __fmul_rn((float)(uint32_t)((node / nn) % nn), l);
  cz[k] = __fmul_rn((float)(uint32_t)(node % nn), l);
}

// --- host side ---

// exodus_rescale.c:175-283, on the host: the per-rung stencil of mesh-cell
// offsets with their squared-distance bounds and classes.
int template_build(uint32_t mesh_n, float l, float rmin, float r_search,
  std::vector<int4>& cells, std::vector<float2>& bounds) {
  cells.clear();
  bounds.clear();
  const int32_t c = (int32_t)(r_search / l) + 1;
  if (c >= (int32_t)mesh_n) {
    error_set(SIF_ERR_RANGE,
      "search sphere of %g (%d mesh cells): wider than the mesh (%u cells); use a "
      "finer mesh or smaller radii",
      (double)r_search, c, mesh_n);
    return SIF_ERR_RANGE;
  }
  const float rs2 = r_search * r_search;
  const float rc2 = rmin * rmin;
  for (int32_t dx = -c; dx <= c; dx++)
    for (int32_t dy = -c; dy <= c; dy++)
      for (int32_t dz = -c; dz <= c; dz++) {
        const float nx = dx == 0 ? 0.0f : (fabsf((float)dx) - 1.0f) * l;
        const float ny = dy == 0 ? 0.0f : (fabsf((float)dy) - 1.0f) * l;
        const float nz = dz == 0 ? 0.0f : (fabsf((float)dz) - 1.0f) * l;
        const float lo = nx * nx + ny * ny + nz * nz;
        if (lo > rs2)
          continue;
        const float fx = (fabsf((float)dx) + 1.0f) * l;
        const float fy = (fabsf((float)dy) + 1.0f) * l;
        const float fz = (fabsf((float)dz) + 1.0f) * l;
        const float hi = fx * fx + fy * fy + fz * fz;
        int type;
        if (hi <= rc2)
          type = CELL_CORE;
        else if (lo > rc2)
          type = hi <= rs2 ? CELL_SHELL : CELL_EDGE;
        else
          type = CELL_STRADDLES;
        cells.push_back(make_int4(dx, dy, dz, type));
        bounds.push_back(make_float2(lo, hi));
      }
  return SIF_OK;
}

namespace {
struct TemplateCache {
  uint32_t mesh_n = 0;
  float l = 0, rmin = -1, r_search = -1;
  uint32_t count = 0;
  DeviceBuffer cells, bounds;
};
// The templates on the device, one per (mesh, core, reach): a rung's own and,
// under SIF_FINDER_EXTEND_REACH, one per level of the extended reach, each
// built once per run. At most TEMPLATE_SLOTS are kept; the oldest goes first.
constexpr size_t TEMPLATE_SLOTS = 48;
thread_local std::vector<std::unique_ptr<TemplateCache>> g_tpls;
// the scratch of the extended-reach passes: the candidates still beyond their
// reach, gathered, their results, and the count
thread_local DeviceBuffer g_ext_idx, g_ext_cx, g_ext_cy, g_ext_cz, g_ext_rad, g_ext_reason, g_ext_nin,
  g_ext_count;
thread_local DeviceBuffer g_pf_counts;
// the split first pass's partial histograms and counts
thread_local DeviceBuffer g_part_bins, g_part_counts;

// A launch of fewer candidates than the chip holds waves (256 CUs, 16 waves
// each at this kernel's registers and LDS) gives each candidate's first pass
// to several waves, up to MAX_SLICES, so that one large candidate does not
// hold the launch for the time a single wave takes over its sphere. The wave
// count aimed at is a tuning (several runs side by side share the chip, and
// each should aim lower); the split changes no bit (integer partial sums).
constexpr uint32_t WAVE_SLOTS = 4096;
constexpr uint32_t MAX_SLICES = 16;
std::atomic<uint32_t> g_split_waves{WAVE_SLOTS};
}  // namespace

uint32_t grow_split_waves() { return g_split_waves.load(std::memory_order_relaxed); }
int set_grow_split_waves(uint32_t waves) {
  if (waves < 64 || waves > 65536) {
    error_set(SIF_ERR_INVALID, "growth split target %u waves (64 to 65536)", waves);
    return SIF_ERR_INVALID;
  }
  g_split_waves.store(waves, std::memory_order_relaxed);
  return SIF_OK;
}

static int template_device(uint32_t mesh_n, float l, float rmin, float r_search,
  const TemplateCache** out) {
  for (const auto& t : g_tpls)
    if (t->mesh_n == mesh_n && t->l == l && t->rmin == rmin && t->r_search == r_search) {
      *out = t.get();
      return SIF_OK;
    }
  std::vector<int4> cells;
  std::vector<float2> bounds;
  int st = template_build(mesh_n, l, rmin, r_search, cells, bounds);
  if (st != SIF_OK)
    return st;
  if (g_tpls.size() == TEMPLATE_SLOTS) {
    SIF_HIP(hipStreamSynchronize(stream()));  // no launch still reads it
    g_tpls.erase(g_tpls.begin());
  }
  auto t = std::make_unique<TemplateCache>();
  st = t->cells.reserve(cells.size() * sizeof(int4));
  if (st == SIF_OK)
    st = t->bounds.reserve(bounds.size() * sizeof(float2));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(t->cells.ptr, cells.data(), cells.size() * sizeof(int4),
    hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipMemcpyAsync(t->bounds.ptr, bounds.data(), bounds.size() * sizeof(float2),
    hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  t->mesh_n = mesh_n;
  t->l = l;
  t->rmin = rmin;
  t->r_search = r_search;
  t->count = (uint32_t)cells.size();
  *out = t.get();
  g_tpls.push_back(std::move(t));
  return SIF_OK;
}

int grow_batch(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, uint32_t mesh_n, float box, const float* cx, const float* cy,
  const float* cz, uint32_t n, float rmin, float r_search, float threshold,
  float vol_factor, float* radius, uint32_t* reason, uint64_t* n_in,
  uint64_t* algo_bytes, const MeshWeights* weights) {
  const float l = box / (float)mesh_n;
  const bool weighted = weights && weights->w;
  const TemplateCache* tpl = nullptr;
  int st = template_device(mesh_n, l, rmin, r_search, &tpl);
  if (st != SIF_OK)
    return st;
  if (n == 0)
    return SIF_OK;
  GrowParams P;
  P.offsets = offsets;
  P.mx = mx;
  P.my = my;
  P.mz = mz;
  P.mesh_n = (int32_t)mesh_n;
  P.box = box;
  P.inv_l = 1.0f / l;
  P.tcell = tpl->cells.as<int4>();
  P.tbound = tpl->bounds.as<float2>();
  P.tcount = tpl->count;
  P.rmin = rmin;
  P.r_search = r_search;
  P.threshold = threshold;
  P.vol_factor = vol_factor;
  P.cx = cx;
  P.cy = cy;
  P.cz = cz;
  P.n_cand = n;
  P.slices = 1;
  P.part_bins = nullptr;
  P.part_counts = nullptr;
  P.mw = weighted ? weights->w : nullptr;
  P.cw = weighted ? weights->cell : nullptr;
  P.mass_factor = weighted ? weights->mass_factor : vol_factor;
  P.radius = radius;
  P.reason = reason;
  P.n_in = n_in;
  P.algo_bytes = algo_bytes;
  const uint32_t nb = grow_bin_count(r_search, rmin, vol_factor);
  if (weighted)
    grow_lds_words_weighted(nb, &P.lds_bins, &P.lds_cache);
  else
    grow_lds_words(nb, 0, &P.lds_bins, &P.lds_cache);
  const size_t lds_bytes =
    (size_t)GROW_WAVES * grow_wave_words(P.lds_bins, P.lds_cache, weighted) * sizeof(uint32_t);
  const uint32_t split_waves = grow_split_waves();
  if (!weighted && n < split_waves / 2) {
    const uint32_t slices = std::min<uint32_t>(MAX_SLICES, split_waves / n);
    int st2 = g_part_bins.reserve((size_t)n * slices * nb * sizeof(uint32_t));
    if (st2 == SIF_OK)
      st2 = g_part_counts.reserve((size_t)n * slices * PART_COUNTS * sizeof(uint64_t));
    if (st2 != SIF_OK)
      return st2;
    P.slices = slices;
    P.part_bins = g_part_bins.as<uint32_t>();
    P.part_counts = g_part_counts.as<uint64_t>();
  }
  timing_begin(SIF_AMD_STAGE_GROW);
  if (P.slices > 1) {
    const size_t bin_lds = (size_t)(P.lds_bins + LDS_TAIL_WORDS) * sizeof(uint32_t);
    grow_bin_kernel<<<n * P.slices, 64, bin_lds, stream()>>>(P);
    SIF_HIP(hipGetLastError());
  }
  const uint64_t blocks = ((uint64_t)n + GROW_WAVES - 1) / GROW_WAVES;
  if (weighted)
    grow_kernel<true><<<(unsigned)blocks, 64 * GROW_WAVES, lds_bytes, stream()>>>(P);
  else
    grow_kernel<false><<<(unsigned)blocks, 64 * GROW_WAVES, lds_bytes, stream()>>>(P);
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_GROW);
  return SIF_OK;
}

// --- the survey rescaling (grow_survey.h) ---

__global__ void u32_max_kernel(const uint32_t* __restrict__ v, uint32_t n,
  uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = v[i] > m ? v[i] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m)
    atomicMax(out, m);
}

namespace {
thread_local DeviceBuffer g_survey_nb, g_survey_max;

GrowParams survey_params(const SurveyMesh& M, float box, const TemplateCache* t) {
  GrowParams P{};
  P.offsets = M.offsets;
  P.mx = M.x;
  P.my = M.y;
  P.mz = M.z;
  P.mesh_n = (int32_t)M.mesh_n;
  P.box = box;
  P.inv_l = 1.0f / (box / (float)M.mesh_n);
  P.tcell = t->cells.as<int4>();
  P.tbound = t->bounds.as<float2>();
  P.tcount = t->count;
  P.mw = M.w;
  P.cw = M.w ? M.cw : nullptr;
  P.slices = 1;
  return P;
}

template <bool DW, bool RW>
void survey_launch(const GrowParams& PD, const GrowParams& PR, const SurveyParams& Q,
  bool second, size_t lds) {
  if (!second)
    survey_totals_kernel<DW, RW><<<Q.n_cand, 64, 0, stream()>>>(PD, PR, Q);
  else
    survey_grow_kernel<DW, RW><<<Q.n_cand, 64, lds, stream()>>>(PD, PR, Q);
}
}  // namespace

int grow_survey(const SurveyMesh& D, const SurveyMesh& Rm, float box, const float* cx,
  const float* cy, const float* cz, uint32_t n, float rmin, float r_search, float K,
  float* radius, uint32_t* reason, uint64_t* n_in) {
  const TemplateCache *td = nullptr, *tr = nullptr;
  int st = template_device(D.mesh_n, box / (float)D.mesh_n, rmin, r_search, &td);
  if (st == SIF_OK)
    st = template_device(Rm.mesh_n, box / (float)Rm.mesh_n, rmin, r_search, &tr);
  // the second build may have evicted the first from the cache: found again
  // (or rebuilt, evicting an older one than the random mesh's)
  if (st == SIF_OK)
    st = template_device(D.mesh_n, box / (float)D.mesh_n, rmin, r_search, &td);
  if (st != SIF_OK)
    return st;
  if (n == 0)
    return SIF_OK;
  if ((st = g_survey_nb.reserve((size_t)n * 4)) != SIF_OK ||
      (st = g_survey_max.reserve(4)) != SIF_OK)
    return st;
  const GrowParams PD = survey_params(D, box, td);
  const GrowParams PR = survey_params(Rm, box, tr);
  SurveyParams Q{};
  Q.cx = cx;
  Q.cy = cy;
  Q.cz = cz;
  Q.n_cand = n;
  Q.rmin = rmin;
  Q.r_search = r_search;
  Q.K = K;
  Q.radius = radius;
  Q.reason = reason;
  Q.n_in = n_in;
  Q.nb = g_survey_nb.as<uint32_t>();
  Q.nb_max = 0;
  auto launch = [&](bool second, size_t lds) {
    if (D.w && Rm.w)
      survey_launch<true, true>(PD, PR, Q, second, lds);
    else if (D.w)
      survey_launch<true, false>(PD, PR, Q, second, lds);
    else if (Rm.w)
      survey_launch<false, true>(PD, PR, Q, second, lds);
    else
      survey_launch<false, false>(PD, PR, Q, second, lds);
  };
  timing_begin(SIF_AMD_STAGE_GROW);
  launch(false, 0);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemsetAsync(g_survey_max.ptr, 0, 4, stream()));
  u32_max_kernel<<<grid_for(n, 256) > 256 ? 256 : grid_for(n, 256), 256, 0, stream()>>>(
    Q.nb, n, g_survey_max.as<uint32_t>());
  SIF_HIP(hipGetLastError());
  uint32_t nb_max = 0;
  SIF_HIP(hipMemcpyAsync(&nb_max, g_survey_max.ptr, 4, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  if (nb_max > 0) {
    Q.nb_max = nb_max;
    launch(true, (size_t)survey_wave_words(nb_max) * sizeof(uint32_t));
    SIF_HIP(hipGetLastError());
  }
  timing_end(SIF_AMD_STAGE_GROW);
  return SIF_OK;
}

// --- SIF_FINDER_EXTEND_REACH: the candidates still beyond their reach, again
// at the next level (exodus_finder.c:2060-2072) ---

__global__ void ext_gather_kernel(const float* __restrict__ radius,
  const uint32_t* __restrict__ reason, uint32_t n, const float* __restrict__ cx,
  const float* __restrict__ cy, const float* __restrict__ cz, uint32_t* __restrict__ idx,
  float* __restrict__ gx, float* __restrict__ gy, float* __restrict__ gz,
  uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(radius[i] < 0.0f) || reason[i] != R_BEYOND)
    return;
  // the order of the gathered list is free: every candidate's result is its own
  const uint32_t k = atomicAdd(count, 1u);
  idx[k] = i;
  gx[k] = cx[i];
  gy[k] = cy[i];
  gz[k] = cz[i];
}

__global__ void ext_scatter_kernel(const uint32_t* __restrict__ idx, uint32_t m,
  const float* __restrict__ gr, const uint32_t* __restrict__ greason,
  const uint64_t* __restrict__ gnin, float* __restrict__ radius, uint32_t* __restrict__ reason,
  uint64_t* __restrict__ n_in) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m)
    return;
  const uint32_t i = idx[k];
  radius[i] = gr[k];
  reason[i] = greason[k];
  if (n_in)
    n_in[i] = gnin[k];
}

int grow_extend(const uint64_t* offsets, const float* mx, const float* my, const float* mz,
  uint32_t mesh_n, float box, const float* cx, const float* cy, const float* cz, uint32_t n,
  const float* ext_rmin, const float* ext_reach, uint32_t first, uint32_t levels,
  float threshold, float vol_factor, float* radius, uint32_t* reason, uint64_t* n_in,
  const MeshWeights* weights) {
  if (n == 0 || first >= levels)
    return SIF_OK;
  int st = g_ext_idx.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_cx.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_cy.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_cz.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_rad.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_reason.reserve((size_t)n * 4);
  if (st == SIF_OK) st = g_ext_nin.reserve((size_t)n * 8);
  if (st == SIF_OK) st = g_ext_count.reserve(4);
  if (st != SIF_OK)
    return st;
  for (uint32_t e = first; e < levels; e++) {
    SIF_HIP(hipMemsetAsync(g_ext_count.ptr, 0, 4, stream()));
    ext_gather_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(radius, reason, n, cx, cy, cz,
      g_ext_idx.as<uint32_t>(), g_ext_cx.as<float>(), g_ext_cy.as<float>(),
      g_ext_cz.as<float>(), g_ext_count.as<uint32_t>());
    SIF_HIP(hipGetLastError());
    uint32_t m = 0;
    SIF_HIP(hipMemcpyAsync(&m, g_ext_count.ptr, 4, hipMemcpyDeviceToHost, stream()));
    SIF_HIP(hipStreamSynchronize(stream()));
    if (m == 0)
      break;
    st = grow_batch(offsets, mx, my, mz, mesh_n, box, g_ext_cx.as<float>(),
      g_ext_cy.as<float>(), g_ext_cz.as<float>(), m, ext_rmin[e], ext_reach[e], threshold,
      vol_factor, g_ext_rad.as<float>(), g_ext_reason.as<uint32_t>(),
      g_ext_nin.as<uint64_t>(), nullptr, weights);
    if (st != SIF_OK)
      return st;
    ext_scatter_kernel<<<grid_for(m, 256), 256, 0, stream()>>>(g_ext_idx.as<uint32_t>(), m,
      g_ext_rad.as<float>(), g_ext_reason.as<uint32_t>(), g_ext_nin.as<uint64_t>(), radius,
      reason, n_in);
    SIF_HIP(hipGetLastError());
  }
  return SIF_OK;
}

int node_centres(const uint32_t* flat, uint32_t n, uint32_t grid_n, float l,
  float* cx, float* cy, float* cz) {
  if (n == 0)
    return SIF_OK;
  node_centres_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(flat, n, grid_n, l, cx, cy, cz);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int prefilter(uint32_t* flat, uint32_t n, uint32_t grid_n, float l, float box,
  const uint64_t* mask, int32_t proxy, float radius, float f, float scale,
  const VoidIndex& V, uint32_t* keep, uint64_t* d_counts, uint8_t* why) {
  if (n)
    prefilter_kernel<<<std::min<unsigned>(grid_for(n, 256), CAND_BLOCKS), 256, 0, stream()>>>(
      flat, n, grid_n, l, box,
      mask, proxy, radius, f, scale, V, keep, reinterpret_cast<unsigned long long*>(d_counts),
      why);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

__global__ void __launch_bounds__(256) masked_since_kernel(const uint32_t* __restrict__ flat,
  const uint8_t* __restrict__ why, uint64_t n, const uint64_t* __restrict__ mask,
  unsigned long long* __restrict__ out) {
  uint32_t c[4] = {0u, 0u, 0u, 0u};
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x)
    if (mask_bit(mask, flat[k] & 0x7fffffffu))
      c[why[k]]++;
  block_add_counts<4>(c, out);
}

int count_masked_since(const uint32_t* flat, const uint8_t* why, uint64_t n,
  const uint64_t* mask, uint64_t* d_out) {
  if (n)
    masked_since_kernel<<<std::min<unsigned>(grid_for(n, 256), CAND_BLOCKS), 256, 0,
      stream()>>>(flat, why, n, mask,
      reinterpret_cast<unsigned long long*>(d_out));
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

// --- the device void index: class-split CSR over the reference's 32^3 cells ---

__global__ void void_slot_kernel(const float* __restrict__ vx, const float* __restrict__ vy,
  const float* __restrict__ vz, const float* __restrict__ vr, uint32_t nv, float inv_l,
  float cell_w, uint32_t* __restrict__ slot, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv)
    return;
  const float r = vr[i];
  const uint32_t cls = r < 0.25f * cell_w ? 0u : r < 0.5f * cell_w ? 1u : r < cell_w ? 2u : 3u;
  const uint32_t cell =
    (void_axis_cell(vx[i], inv_l) * VOID_CELLS + void_axis_cell(vy[i], inv_l)) * VOID_CELLS +
    void_axis_cell(vz[i], inv_l);
  const uint32_t sl = cls * VOID_CELLS * VOID_CELLS * VOID_CELLS + cell;
  slot[i] = sl;
  atomicAdd(&counts[sl], 1u);
}

// Also the occupancy words: slot s's bit is bit (s mod 32) of word s / 32,
// set when the slot lists a void (slots is a multiple of 64, so each wave's
// two words lie within it or wholly past it).
__global__ void void_start_kernel(const uint64_t* __restrict__ pos, uint32_t slots,
  uint32_t* __restrict__ start, uint32_t* __restrict__ cursor,
  uint32_t* __restrict__ occupied) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= slots)
    start[s] = (uint32_t)pos[s];
  if (s < slots)
    cursor[s] = 0;
  const uint64_t m = __ballot(s < slots && pos[s + 1] > pos[s]);
  const uint32_t lane = threadIdx.x & 63;
  if ((lane & 31) == 0 && s < slots)
    occupied[s >> 5] = (uint32_t)(m >> lane);
}

// Within a cell the order of the voids is not fixed; nothing that reads the
// index depends on it (an overlap state is a yes/no over a set).
__global__ void void_scatter_kernel(const uint32_t* __restrict__ slot, uint32_t nv,
  const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
  const float* __restrict__ vr, const uint32_t* __restrict__ start,
  uint32_t* __restrict__ cursor, float4* __restrict__ voids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv)
    return;
  const uint32_t sl = slot[i];
  voids[start[sl] + atomicAdd(&cursor[sl], 1u)] = make_float4(vx[i], vy[i], vz[i], vr[i]);
}

namespace {
thread_local DeviceBuffer g_vi_slot, g_vi_counts, g_vi_pos, g_vi_cursor;
}

int build_void_index(const float* vx, const float* vy, const float* vz, const float* vr,
  uint32_t nv, float box, uint32_t* d_start, float4* d_voids, uint32_t* d_occupied) {
  const uint32_t slots = VOID_CLASSES * VOID_CELLS * VOID_CELLS * VOID_CELLS;
  static_assert((VOID_CLASSES * VOID_CELLS * VOID_CELLS * VOID_CELLS) % 64 == 0,
    "whole waves of slots");
  int st = g_vi_slot.reserve((size_t)(nv ? nv : 1) * 4);
  if (st == SIF_OK) st = g_vi_counts.reserve((size_t)(slots + 1) * 4);
  if (st == SIF_OK) st = g_vi_pos.reserve((size_t)(slots + 1) * 8);
  if (st == SIF_OK) st = g_vi_cursor.reserve((size_t)slots * 4);
  if (st != SIF_OK)
    return st;
  // counts[slots] stays 0 so that the exclusive scan's last entry is nv
  SIF_HIP(hipMemsetAsync(g_vi_counts.ptr, 0, (size_t)(slots + 1) * 4, stream()));
  if (nv)
    void_slot_kernel<<<grid_for(nv, 256), 256, 0, stream()>>>(vx, vy, vz, vr, nv,
      (float)VOID_CELLS / box, box / (float)VOID_CELLS, g_vi_slot.as<uint32_t>(),
      g_vi_counts.as<uint32_t>());
  SIF_HIP(hipGetLastError());
  st = exclusive_scan_u32(g_vi_counts.as<uint32_t>(), slots + 1, g_vi_pos.as<uint64_t>(),
    nullptr);
  if (st != SIF_OK)
    return st;
  void_start_kernel<<<grid_for(slots + 1, 256), 256, 0, stream()>>>(g_vi_pos.as<uint64_t>(),
    slots, d_start, g_vi_cursor.as<uint32_t>(), d_occupied);
  SIF_HIP(hipGetLastError());
  if (nv)
    void_scatter_kernel<<<grid_for(nv, 256), 256, 0, stream()>>>(g_vi_slot.as<uint32_t>(), nv,
      vx, vy, vz, vr, d_start, g_vi_cursor.as<uint32_t>(), d_voids);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int grown_overlap(const float* cx, const float* cy, const float* cz, const uint32_t* flat,
  const float* radius, uint32_t n, const VoidIndex& V, float max_r, float box, float f,
  float scale, const uint64_t* mask, uint32_t grid_n, float l, uint32_t* state) {
  if (n == 0)
    return SIF_OK;
  grown_overlap_kernel<<<(n + GO_WAVES - 1) / GO_WAVES, 64 * GO_WAVES, 0, stream()>>>(cx, cy,
    cz, flat, radius, n, V, max_r, box, f, scale, mask, grid_n, l, state);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int paint_voids(const float* vx, const float* vy, const float* vz, const float* vr,
  uint32_t nv, float scale, float l, uint32_t grid_n, uint64_t* mask) {
  if (nv == 0)
    return SIF_OK;
  paint_voids_kernel<<<nv, 256, 0, stream()>>>(vx, vy, vz, vr, nv, scale, l, grid_n,
    reinterpret_cast<unsigned long long*>(mask));
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_mark_spheres(uint64_t* d_mask_words, uint32_t n_cells, float box,
  const float* d_cx, const float* d_cy, const float* d_cz, const float* d_r,
  uint32_t n_voids, float scale) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!d_mask_words || n_cells == 0 || !(box > 0.0f) ||
      (n_voids && (!d_cx || !d_cy || !d_cz || !d_r))) {
    error_set(SIF_ERR_INVALID, "sif_amd_mark_spheres: invalid arguments");
    return SIF_ERR_INVALID;
  }
  st = paint_voids(d_cx, d_cy, d_cz, d_r, n_voids, scale, box / (float)n_cells, n_cells,
    d_mask_words);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_template(uint32_t mesh_cells, float box, float rmin, float r_search,
  int32_t* cells4, float* bounds2, uint32_t capacity, uint32_t* count) {
  std::vector<int4> cells;
  std::vector<float2> bounds;
  int st = template_build(mesh_cells, box / (float)mesh_cells, rmin, r_search, cells, bounds);
  if (st != SIF_OK)
    return st;
  *count = (uint32_t)cells.size();
  for (uint32_t i = 0; i < *count && i < capacity; i++) {
    cells4[4 * i] = cells[i].x;
    cells4[4 * i + 1] = cells[i].y;
    cells4[4 * i + 2] = cells[i].z;
    cells4[4 * i + 3] = cells[i].w;
    bounds2[2 * i] = bounds[i].x;
    bounds2[2 * i + 1] = bounds[i].y;
  }
  return SIF_OK;
}

int sif_amd_find_radius(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, uint32_t mesh_cells, float box, const float* cx,
  const float* cy, const float* cz, uint32_t n, float rmin, float r_search,
  float threshold, float vol_factor, float* radius, uint32_t* reason,
  uint64_t* n_in) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!offsets || !mx || !my || !mz || !cx || !cy || !cz || !radius || !reason ||
      mesh_cells == 0 || !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "sif_amd_find_radius: invalid arguments");
    return SIF_ERR_INVALID;
  }
  st = grow_batch(offsets, mx, my, mz, mesh_cells, box, cx, cy, cz, n, rmin,
    r_search, threshold, vol_factor, radius, reason, n_in);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_find_radius_survey(const sif_amd_mesh_t* data, const sif_amd_mesh_t* randoms,
  float box, const float* cx, const float* cy, const float* cz, uint32_t n, float rmin,
  float r_search, float K, float* radius, uint32_t* reason, uint64_t* n_in) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!data || !randoms || !data->offsets || !data->x || !data->y || !data->z ||
      !randoms->offsets || !randoms->x || !randoms->y || !randoms->z || data->n_cells == 0 ||
      randoms->n_cells == 0 || (data->weights && !data->cell_weights) ||
      (randoms->weights && !randoms->cell_weights) || (n && (!cx || !cy || !cz)) ||
      !radius || !reason || !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "sif_amd_find_radius_survey: invalid arguments");
    return SIF_ERR_INVALID;
  }
  SurveyMesh D, Rm;
  D.offsets = data->offsets;
  D.x = data->x;
  D.y = data->y;
  D.z = data->z;
  D.w = data->weights;
  D.cw = data->cell_weights;
  D.mesh_n = data->n_cells;
  Rm.offsets = randoms->offsets;
  Rm.x = randoms->x;
  Rm.y = randoms->y;
  Rm.z = randoms->z;
  Rm.w = randoms->weights;
  Rm.cw = randoms->cell_weights;
  Rm.mesh_n = randoms->n_cells;
  st = grow_survey(D, Rm, box, cx, cy, cz, n, rmin, r_search, K, radius, reason, n_in);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_find_radius_weighted(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, const float* mw, const float* cell_weights, uint32_t mesh_cells, float box,
  const float* cx, const float* cy, const float* cz, uint32_t n, float rmin, float r_search,
  float threshold, float vol_factor, float mass_factor, float* radius, uint32_t* reason,
  uint64_t* n_in) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!offsets || !mx || !my || !mz || !mw || !cell_weights || !cx || !cy || !cz ||
      !radius || !reason || mesh_cells == 0 || !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "sif_amd_find_radius_weighted: invalid arguments");
    return SIF_ERR_INVALID;
  }
  MeshWeights w;
  w.w = mw;
  w.cell = cell_weights;
  w.mass_factor = mass_factor;
  st = grow_batch(offsets, mx, my, mz, mesh_cells, box, cx, cy, cz, n, rmin, r_search,
    threshold, vol_factor, radius, reason, n_in, nullptr, &w);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
