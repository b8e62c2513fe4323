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
#include <vector>

#include "common.h"
#include "internal.h"
#include "sif_amd.h"
#include "grow_device.h"

namespace sif_amd {

// --- prefilter: the rung-radius tests of exodus_finder.c:871-944 ---

constexpr int VOID_CELLS = 32;

__device__ __forceinline__ uint32_t wrap_pbc(int32_t v, uint32_t n) {
  const int32_t w = v % (int32_t)n;
  return (uint32_t)(w < 0 ? w + (int32_t)n : w);
}

__device__ __forceinline__ bool mask_bit(const uint64_t* m, uint64_t i) {
  return (m[i >> 6] >> (i & 63)) & 1ull;
}

// utils.c:819-848
__device__ bool overlap_quick(const uint64_t* mask, uint32_t n, uint32_t ix,
  uint32_t iy, uint32_t iz, uint32_t r) {
  const uint64_t nn = n;
  if (mask_bit(mask, ((uint64_t)wrap_pbc((int32_t)(ix + r), n) * nn + iy) * nn + iz))
    return true;
  if (mask_bit(mask, ((uint64_t)wrap_pbc((int32_t)ix - (int32_t)r, n) * nn + iy) * nn + iz))
    return true;
  if (mask_bit(mask, ((uint64_t)ix * nn + wrap_pbc((int32_t)(iy + r), n)) * nn + iz))
    return true;
  if (mask_bit(mask, ((uint64_t)ix * nn + wrap_pbc((int32_t)iy - (int32_t)r, n)) * nn + iz))
    return true;
  if (mask_bit(mask, ((uint64_t)ix * nn + iy) * nn + wrap_pbc((int32_t)(iz + r), n)))
    return true;
  if (mask_bit(mask, ((uint64_t)ix * nn + iy) * nn + wrap_pbc((int32_t)iz - (int32_t)r, n)))
    return true;
  return false;
}

// utils.c:850-960, over a CSR void index instead of a linked list (the
// answer is a yes/no over a set, so the visiting order does not matter).
__device__ bool overlap_exact(const VoidIndex& V, float cx, float cy, float cz,
  float r, float box, float f, float scale) {
  const float half = __fmul_rn(box, 0.5f);
  const float mr0 = r < V.max_r ? r : V.max_r;
  const float sr = __fmul_rn(scale, __fsub_rn(__fadd_rn(r, V.max_r), __fmul_rn(f, mr0)));
  const int32_t lo0 = (int32_t)floorf(__fmul_rn(__fsub_rn(cx, sr), V.inv_l));
  const int32_t hi0 = (int32_t)floorf(__fmul_rn(__fadd_rn(cx, sr), V.inv_l));
  const int32_t lo1 = (int32_t)floorf(__fmul_rn(__fsub_rn(cy, sr), V.inv_l));
  const int32_t hi1 = (int32_t)floorf(__fmul_rn(__fadd_rn(cy, sr), V.inv_l));
  const int32_t lo2 = (int32_t)floorf(__fmul_rn(__fsub_rn(cz, sr), V.inv_l));
  const int32_t hi2 = (int32_t)floorf(__fmul_rn(__fadd_rn(cz, sr), V.inv_l));
  const bool interior = (__fsub_rn(cx, sr) >= 0.0f) && (__fadd_rn(cx, sr) < box) &&
                        (__fsub_rn(cy, sr) >= 0.0f) && (__fadd_rn(cy, sr) < box) &&
                        (__fsub_rn(cz, sr) >= 0.0f) && (__fadd_rn(cz, sr) < box);
  for (int32_t a = lo0; a <= hi0; a++) {
    const uint32_t wa = wrap_pbc(a, VOID_CELLS);
    for (int32_t b = lo1; b <= hi1; b++) {
      const uint32_t wb = wrap_pbc(b, VOID_CELLS);
      for (int32_t c = lo2; c <= hi2; c++) {
        const uint32_t cell = (wa * VOID_CELLS + wb) * VOID_CELLS + wrap_pbc(c, VOID_CELLS);
        for (uint32_t k = V.cell_start[cell]; k < V.cell_start[cell + 1]; k++) {
          const uint32_t i = V.items[k];
          const float ri = V.r[i];
          const float mr = r < ri ? r : ri;
          const float tot = __fmul_rn(scale, __fsub_rn(__fadd_rn(r, ri), __fmul_rn(f, mr)));
          const float t2 = __fmul_rn(tot, tot);
          if (interior) {
            const float dx = __fsub_rn(cx, V.cx[i]), dy = __fsub_rn(cy, V.cy[i]),
                        dz = __fsub_rn(cz, V.cz[i]);
            if (__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)) < t2)
              return true;
          } else {
            float dx = fabsf(__fsub_rn(cx, V.cx[i]));
            dx = dx > half ? __fsub_rn(box, dx) : dx;
            const float dx2 = __fmul_rn(dx, dx);
            if (dx2 >= t2)
              continue;
            float dy = fabsf(__fsub_rn(cy, V.cy[i]));
            dy = dy > half ? __fsub_rn(box, dy) : dy;
            const float dy2 = __fmul_rn(dy, dy);
            if (__fadd_rn(dx2, dy2) >= t2)
              continue;
            float dz = fabsf(__fsub_rn(cz, V.cz[i]));
            dz = dz > half ? __fsub_rn(box, dz) : dz;
            if (__fadd_rn(__fadd_rn(dx2, dy2), __fmul_rn(dz, dz)) < t2)
              return true;
          }
        }
      }
    }
  }
  return false;
}

// keep[k] = 0 refused (masked node, pole probe or exact overlap at the rung
// radius, against the catalogue as the rung began), 1 kept.
__global__ void prefilter_kernel(const uint32_t* __restrict__ flat, uint32_t n,
  uint32_t grid_n, float l, float box, const uint64_t* __restrict__ mask,
  int32_t proxy, float radius, float f, float scale, VoidIndex V, uint32_t n_voids,
  uint32_t* __restrict__ keep, unsigned long long* __restrict__ counts) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n)
    return;
  const uint64_t node = flat[k];
  uint32_t why = 0;  // 0 kept, 1 masked, 2 proxy, 3 overlap
  if (mask_bit(mask, node)) {
    why = 1;
  } else {
    const uint64_t nn = grid_n;
    const uint32_t ix = (uint32_t)(node / (nn * nn));
    const uint32_t iy = (uint32_t)((node / nn) % nn);
    const uint32_t iz = (uint32_t)(node % nn);
    if (proxy > 0 && overlap_quick(mask, grid_n, ix, iy, iz, (uint32_t)proxy))
      why = 2;
    else if (n_voids && overlap_exact(V, __fmul_rn((float)ix, l), __fmul_rn((float)iy, l),
                          __fmul_rn((float)iz, l), radius, box, f, scale))
      why = 3;
  }
  keep[k] = why == 0 ? 1u : 0u;
  if (why)
    atomicAdd(&counts[why - 1], 1ull);
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
TemplateCache g_tpl;
DeviceBuffer g_pf_counts;
}  // namespace

static int template_device(uint32_t mesh_n, float l, float rmin, float r_search) {
  if (g_tpl.mesh_n == mesh_n && g_tpl.l == l && g_tpl.rmin == rmin &&
      g_tpl.r_search == r_search)
    return SIF_OK;
  std::vector<int4> cells;
  std::vector<float2> bounds;
  int st = template_build(mesh_n, l, rmin, r_search, cells, bounds);
  if (st == SIF_OK)
    st = g_tpl.cells.reserve(cells.size() * sizeof(int4));
  if (st == SIF_OK)
    st = g_tpl.bounds.reserve(bounds.size() * sizeof(float2));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(g_tpl.cells.ptr, cells.data(), cells.size() * sizeof(int4),
    hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipMemcpyAsync(g_tpl.bounds.ptr, bounds.data(), bounds.size() * sizeof(float2),
    hipMemcpyHostToDevice, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  g_tpl.mesh_n = mesh_n;
  g_tpl.l = l;
  g_tpl.rmin = rmin;
  g_tpl.r_search = r_search;
  g_tpl.count = (uint32_t)cells.size();
  return SIF_OK;
}

int grow_batch(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, uint32_t mesh_n, float box, const float* cx, const float* cy,
  const float* cz, uint32_t n, float rmin, float r_search, float threshold,
  float vol_factor, float* radius, uint32_t* reason, uint64_t* n_in) {
  const float l = box / (float)mesh_n;
  int st = template_device(mesh_n, l, rmin, r_search);
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
  P.tcell = g_tpl.cells.as<int4>();
  P.tbound = g_tpl.bounds.as<float2>();
  P.tcount = g_tpl.count;
  P.rmin = rmin;
  P.r_search = r_search;
  P.threshold = threshold;
  P.vol_factor = vol_factor;
  P.cx = cx;
  P.cy = cy;
  P.cz = cz;
  P.n_cand = n;
  P.radius = radius;
  P.reason = reason;
  P.n_in = n_in;
  timing_begin(SIF_AMD_STAGE_GROW);
  const uint64_t blocks = ((uint64_t)n + GROW_WAVES - 1) / GROW_WAVES;
  grow_kernel<<<(unsigned)blocks, 64 * GROW_WAVES, 0, stream()>>>(P);
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_GROW);
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

int prefilter(const uint32_t* flat, uint32_t n, uint32_t grid_n, float l, float box,
  const uint64_t* mask, int32_t proxy, float radius, float f, float scale,
  const VoidIndex& V, uint32_t n_voids, uint32_t* keep, uint64_t counts_out[3]) {
  int st = g_pf_counts.reserve(3 * sizeof(unsigned long long));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_pf_counts.ptr, 0, 3 * sizeof(unsigned long long), stream()));
  if (n)
    prefilter_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(flat, n, grid_n, l, box,
      mask, proxy, radius, f, scale, V, n_voids, keep,
      g_pf_counts.as<unsigned long long>());
  SIF_HIP(hipGetLastError());
  unsigned long long c[3];
  SIF_HIP(hipMemcpyAsync(c, g_pf_counts.ptr, sizeof(c), hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  for (int i = 0; i < 3; i++)
    counts_out[i] = c[i];
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

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

}  // extern "C"
