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

namespace sif_amd {

constexpr int GROW_WAVES = 4;
constexpr int MAX_BINS = 2048;
constexpr int REFINE_CAP = 1024;
constexpr int SUB_BINS = 64;

enum { CELL_CORE = 1, CELL_STRADDLES = 2, CELL_EDGE = 3, CELL_SHELL = 4 };

enum {
  R_OK = 0,
  R_EMPTY = 1,
  R_BEYOND = 2,
  R_BELOW_RUNG = 3,
  R_DEGENERATE = 4,
  R_OVERFLOW = 5
};

struct GrowParams {
  const uint64_t* offsets;
  const float* mx;
  const float* my;
  const float* mz;
  int32_t mesh_n;
  float box;
  float inv_l;
  const int4* tcell;    // dx, dy, dz, type
  const float2* tbound; // min_d2, max_d2
  uint32_t tcount;
  float rmin, r_search, threshold, vol_factor;
  const float* cx;
  const float* cy;
  const float* cz;
  uint32_t n_cand;
  uint32_t* work_counter;
  float* radius;
  uint32_t* reason;
  uint64_t* n_in;
};

struct WaveLDS {
  uint32_t bins[MAX_BINS];
  float buf[REFINE_CAP];
  uint32_t t_start[64];
  uint32_t t_scan[64];
  float t_ex[64], t_ey[64], t_ez[64];
  uint32_t t_type[64];
  uint32_t sub[SUB_BINS];
  uint32_t counter;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_max_i32(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(v, off, 64);
    if (lane >= off)
      v += y;
  }
  return v;
}

// exodus_rescale.c:349-355, no contraction
__device__ __forceinline__ float dist2(float px, float py, float pz, float cx,
  float cy, float cz) {
  const float dx = __fsub_rn(px, cx), dy = __fsub_rn(py, cy), dz = __fsub_rn(pz, cz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// exodus_rescale.c:360-369
__device__ __forceinline__ uint32_t bin_of(float d2, float rc2, float inv_w, uint32_t nb) {
  const int32_t b = (int32_t)__fmul_rn(__fsub_rn(d2, rc2), inv_w);
  if (b < 0)
    return 0;
  if ((uint32_t)b >= nb)
    return nb - 1;
  return (uint32_t)b;
}

struct Query {
  float cx, cy, cz;
  int32_t ix, iy, iz;
};

// exodus_rescale.c:295-342: template entry -> particle range and the image
// of the centre it is measured against.
__device__ __forceinline__ void query_cell(const GrowParams& P, const Query& q,
  const int4 t, uint32_t* start, uint32_t* count, float* ex, float* ey, float* ez) {
  const int32_t N = P.mesh_n;
  const float L = P.box;
  int32_t ix = q.ix + t.x;
  float cx = q.cx;
  if (ix < 0) {
    ix += N;
    cx = __fadd_rn(cx, L);
  } else if (ix >= N) {
    ix -= N;
    cx = __fsub_rn(cx, L);
  }
  int32_t iy = q.iy + t.y;
  float cy = q.cy;
  if (iy < 0) {
    iy += N;
    cy = __fadd_rn(cy, L);
  } else if (iy >= N) {
    iy -= N;
    cy = __fsub_rn(cy, L);
  }
  int32_t iz = q.iz + t.z;
  float cz = q.cz;
  if (iz < 0) {
    iz += N;
    cz = __fadd_rn(cz, L);
  } else if (iz >= N) {
    iz -= N;
    cz = __fsub_rn(cz, L);
  }
  const uint64_t flat = ((uint64_t)ix * N + (uint64_t)iy) * N + (uint64_t)iz;
  const uint64_t s = P.offsets[flat], e = P.offsets[flat + 1];
  *start = (uint32_t)s;
  *count = (uint32_t)(e - s);
  *ex = cx;
  *ey = cy;
  *ez = cz;
}

// Walks the template 64 entries at a time. `want(entry, type, bound)` says
// whether the entry is visited; `whole(type, count)` consumes a whole cell
// without reading its particles (returns true when it did); `visit(d2, type)`
// sees every particle of the remaining cells.
template <class Want, class Whole, class Visit>
__device__ __forceinline__ void for_particles(const GrowParams& P, const Query& q,
  WaveLDS& S, int lane, Want want, Whole whole, Visit visit) {
  for (uint32_t base = 0; base < P.tcount; base += 64) {
    const uint32_t e = base + lane;
    uint32_t cnt = 0, start = 0, type = 0;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (e < P.tcount) {
      const int4 t = P.tcell[e];
      type = (uint32_t)t.w;
      if (want(type, P.tbound[e])) {
        query_cell(P, q, t, &start, &cnt, &ex, &ey, &ez);
        if (cnt && whole(type, cnt))
          cnt = 0;
      }
    }
    const uint32_t incl = wave_incl_scan_u32(cnt, lane);
    const uint32_t total = __shfl(incl, 63, 64);
    if (total == 0)
      continue;
    S.t_start[lane] = start;
    S.t_scan[lane] = incl;
    S.t_ex[lane] = ex;
    S.t_ey[lane] = ey;
    S.t_ez[lane] = ez;
    S.t_type[lane] = type;
    wave_sync();
    for (uint32_t j = lane; j < total; j += 64) {
      // first entry whose inclusive prefix exceeds j
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (S.t_scan[mid] > j)
          hi = mid;
        else
          lo = mid + 1;
      }
      const uint32_t before = lo ? S.t_scan[lo - 1] : 0u;
      const uint32_t p = S.t_start[lo] + (j - before);
      const float d2 = dist2(P.mx[p], P.my[p], P.mz[p], S.t_ex[lo], S.t_ey[lo], S.t_ez[lo]);
      visit(d2, S.t_type[lo]);
    }
    wave_sync();
  }
}

// Ascending bitonic sort of buf[0..n) (n <= REFINE_CAP) by one wave.
__device__ void wave_bitonic(float* buf, uint32_t n, int lane) {
  uint32_t P2 = 1;
  while (P2 < n)
    P2 <<= 1;
  for (uint32_t i = n + lane; i < P2; i += 64)
    buf[i] = __int_as_float(0x7f800000);  // +inf pads sort last
  wave_sync();
  for (uint32_t k = 2; k <= P2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = lane; t < P2 / 2; t += 64) {
        const uint32_t lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const uint32_t hi = lo + j;
        const bool asc = (lo & k) == 0;
        const float a = buf[lo], b = buf[hi];
        if ((a > b) == asc) {
          buf[lo] = b;
          buf[hi] = a;
        }
      }
      wave_sync();
    }
  }
}

__global__ void __launch_bounds__(64 * GROW_WAVES) grow_kernel(GrowParams P) {
  __shared__ WaveLDS lds[GROW_WAVES];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  WaveLDS& S = lds[w];

  const float rs2 = __fmul_rn(P.r_search, P.r_search);
  const float rc2 = __fmul_rn(P.rmin, P.rmin);
  const float span = __fsub_rn(rs2, rc2);
  const float rs3 = __fmul_rn(__fmul_rn(P.r_search, P.r_search), P.r_search);
  const float rm3 = __fmul_rn(__fmul_rn(P.rmin, P.rmin), P.rmin);
  uint32_t nb = MAX_BINS;
  {
    const float exp_shell = __fmul_rn(P.vol_factor, __fsub_rn(rs3, rm3));
    const float want = __fdiv_rn(exp_shell, 16.0f);
    if (!(want >= 64.0f))
      nb = 64;
    else if (want < (float)MAX_BINS)
      nb = (uint32_t)want;
  }
  const float expected = __fmul_rn(__fmul_rn(__fmul_rn(P.vol_factor, P.r_search), P.r_search), P.r_search);
  const float K = __fmul_rn(__fadd_rn(P.threshold, 1.0f), P.vol_factor);
  const float K2 = __fmul_rn(K, K);
  const float inv_w = __fdiv_rn((float)nb, span);
  const float bin_w = __fdiv_rn(span, (float)nb);

  for (;;) {
    uint32_t c = 0;
    if (lane == 0)
      c = atomicAdd(P.work_counter, 1u);
    c = __shfl(c, 0, 64);
    if (c >= P.n_cand)
      break;

    Query q;
    q.cx = P.cx[c];
    q.cy = P.cy[c];
    q.cz = P.cz[c];
    q.ix = (int32_t)__fmul_rn(q.cx, P.inv_l);
    q.iy = (int32_t)__fmul_rn(q.cy, P.inv_l);
    q.iz = (int32_t)__fmul_rn(q.cz, P.inv_l);

    float radius = -1.0f;
    uint32_t reason = R_OK;
    uint64_t n_in_out = 0;

    if (!(span > 0.0f)) {
      reason = R_DEGENERATE;
    } else {
      // 1. early refusal, O(surface): exodus_rescale.c:711-782
      uint64_t within = 0;
      for_particles(P, q, S, lane,
        [](uint32_t, float2) { return true; },
        [&](uint32_t type, uint32_t cnt) {
          if (type == CELL_CORE || type == CELL_SHELL) {
            within += cnt;
            return true;
          }
          return false;
        },
        [&](float d2, uint32_t) { within += d2 <= rs2 ? 1u : 0u; });
      within = wave_sum_u64(within);
      if (within == 0) {
        reason = R_EMPTY;
      } else if (__fsub_rn(__fdiv_rn((float)within, expected), 1.0f) <= P.threshold) {
        reason = R_BEYOND;
      }
    }

    if (reason == R_OK) {
      // 2. histogram of the annulus, uniform in d^2: exodus_rescale.c:793-865
      for (uint32_t b = lane; b < nb; b += 64)
        S.bins[b] = 0;
      wave_sync();
      uint64_t n_core = 0, n_shell = 0;
      for_particles(P, q, S, lane,
        [](uint32_t type, float2) { return true; },
        [&](uint32_t type, uint32_t cnt) {
          if (type == CELL_CORE) {
            n_core += cnt;
            return true;
          }
          return false;
        },
        [&](float d2, uint32_t type) {
          if (type == CELL_SHELL) {
            atomicAdd(&S.bins[bin_of(d2, rc2, inv_w, nb)], 1u);
            n_shell++;
          } else {
            n_core += d2 <= rc2 ? 1u : 0u;
            if (d2 > rc2 && d2 <= rs2) {
              atomicAdd(&S.bins[bin_of(d2, rc2, inv_w, nb)], 1u);
              n_shell++;
            }
          }
        });
      wave_sync();
      const uint64_t total = wave_sum_u64(n_core) + wave_sum_u64(n_shell);
      if (total == 0) {
        reason = R_EMPTY;
      } else if (__fsub_rn(__fdiv_rn((float)total, expected), 1.0f) <= P.threshold) {
        reason = R_BEYOND;
      } else {
        // 3. skip flags: lane L owns bins [L*per, (L+1)*per), top down.
        const uint32_t per = (nb + 63) / 64;
        const uint32_t b0 = lane * per;
        const uint32_t b1 = b0 + per < nb ? b0 + per : nb;
        uint64_t mine = 0;
        for (uint32_t b = b0; b < b1; b++)
          mine += S.bins[b];
        // bins above this lane's range: suffix over higher lanes
        uint64_t suffix = mine;
        for (int off = 1; off < 64; off <<= 1) {
          const uint64_t y = __shfl_down(suffix, off, 64);
          if (lane + off < 64)
            suffix += y;
        }
        uint64_t above = suffix - mine;
        for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--) {
          const uint32_t cnt = S.bins[b];
          if (cnt == 0)
            continue;
          float hi = __fadd_rn(rc2, __fmul_rn((float)(b + 1), bin_w));
          if (b == (int32_t)nb - 1)
            hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
          const float n_min = (float)(total - above - cnt);
          const float lhs = __fmul_rn(n_min, n_min);
          const float rhs = __fmul_rn(__fmul_rn(__fmul_rn(K2, hi), hi), hi);
          if (!(lhs > rhs))
            S.bins[b] = cnt | 0x80000000u;  // a bin the walk must open
          above += cnt;
        }
        wave_sync();

        // 4. walk: next flagged bin below `limit`, resolve a window, repeat.
        int32_t limit = (int32_t)nb;
        while (reason == R_OK && radius < 0.0f) {
          int best = -1;
          for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--)
            if (b < limit && (S.bins[b] & 0x80000000u)) {
              best = b;
              break;
            }
          best = wave_max_i32(best);
          if (best < 0) {
            reason = R_BELOW_RUNG;
            break;
          }
          // particles at or below the top of bin `best`
          uint64_t abv = 0;
          for (uint32_t b = (uint32_t)best + 1 + lane; b < nb; b += 64)
            abv += S.bins[b] & 0x7fffffffu;
          abv = wave_sum_u64(abv);
          const uint64_t cur_n = total - abv;
          // window [best - W + 1, best]: as many bins as fit the buffer
          uint32_t wcnt = 0;
          if (best - lane >= 0)
            wcnt = S.bins[best - lane] & 0x7fffffffu;
          const uint32_t wscan = wave_incl_scan_u32(wcnt, lane);
          const uint64_t fits = __ballot(best - lane >= 0 && wscan <= REFINE_CAP);
          int W = fits ? __popcll(fits) : 0;
          const bool split = W == 0;  // the top bin alone overflows
          if (split)
            W = 1;
          const int32_t blo = best - W + 1;
          float hi = __fadd_rn(rc2, __fmul_rn((float)(best + 1), bin_w));
          if (best == (int32_t)nb - 1)
            hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
          const float lo = __fadd_rn(rc2, __fmul_rn((float)blo, bin_w));
          const uint32_t win_total = __shfl(wscan, W - 1, 64);

          // sub-ranges of squared distance: 1 normally, SUB_BINS when split
          const int n_sub = split ? SUB_BINS : 1;
          const float inv_sub = n_sub > 1 ? __fdiv_rn((float)n_sub, __fsub_rn(hi, lo)) : 0.0f;
          uint32_t sub_count_mine = 0;  // lane s holds the count of sub-range s
          if (split) {
            S.sub[lane] = 0;
            wave_sync();
            for_particles(P, q, S, lane,
              [&](uint32_t type, float2 bd) {
                return type != CELL_CORE && !(bd.y < lo) && !(bd.x > hi);
              },
              [](uint32_t, uint32_t) { return false; },
              [&](float d2, uint32_t type) {
                if (type != CELL_SHELL && !(d2 > rc2 && d2 <= rs2))
                  return;
                const uint32_t b = bin_of(d2, rc2, inv_w, nb);
                if ((int32_t)b < blo || (int32_t)b > best)
                  return;
                int32_t sb = (int32_t)__fmul_rn(__fsub_rn(d2, lo), inv_sub);
                sb = sb < 0 ? 0 : (sb >= SUB_BINS ? SUB_BINS - 1 : sb);
                atomicAdd(&S.sub[sb], 1u);
              });
            wave_sync();
            sub_count_mine = S.sub[lane];
          }

          uint64_t top_n = cur_n;  // particles at or below the top of the range
          bool exhausted = false;
          for (int s = n_sub - 1; s >= 0 && radius < 0.0f && !exhausted; s--) {
            uint32_t expect_cnt = split ? __shfl(sub_count_mine, s, 64) : win_total;
            if (expect_cnt == 0)
              continue;
            if (expect_cnt > (uint32_t)REFINE_CAP) {
              reason = R_OVERFLOW;
              break;
            }
            if (lane == 0)
              S.counter = 0;
            wave_sync();
            for_particles(P, q, S, lane,
              [&](uint32_t type, float2 bd) {
                return type != CELL_CORE && !(bd.y < lo) && !(bd.x > hi);
              },
              [](uint32_t, uint32_t) { return false; },
              [&](float d2, uint32_t type) {
                if (type != CELL_SHELL && !(d2 > rc2 && d2 <= rs2))
                  return;
                const uint32_t b = bin_of(d2, rc2, inv_w, nb);
                if ((int32_t)b < blo || (int32_t)b > best)
                  return;
                if (split) {
                  int32_t sb = (int32_t)__fmul_rn(__fsub_rn(d2, lo), inv_sub);
                  sb = sb < 0 ? 0 : (sb >= SUB_BINS ? SUB_BINS - 1 : sb);
                  if (sb != s)
                    return;
                }
                const uint32_t slot = atomicAdd(&S.counter, 1u);
                if (slot < (uint32_t)REFINE_CAP)
                  S.buf[slot] = d2;
              });
            wave_sync();
            const uint32_t got = S.counter;
            if (got != expect_cnt) {
              reason = R_OVERFLOW;  // counts disagree: never expected
              break;
            }
            wave_bitonic(S.buf, got, lane);
            // outside in: element k encloses top_n - got + k particles
            int found = -1;
            for (uint32_t k = lane; k < got; k += 64) {
              const float d2 = S.buf[k];
              if (d2 == 0.0f)
                continue;
              const uint64_t n_in = top_n - got + k;
              const float rn = (float)n_in;
              const float lhs = __fmul_rn(rn, rn);
              const float rhs = __fmul_rn(K2, __fmul_rn(__fmul_rn(d2, d2), d2));
              if (lhs <= rhs)
                found = (int)k;
            }
            found = wave_max_i32(found);
            if (found >= 0) {
              const float d2 = S.buf[found];
              radius = __fsqrt_rn(d2);
              n_in_out = top_n - got + (uint64_t)found;
            }
            top_n -= got;
            wave_sync();
          }
          if (radius < 0.0f && reason == R_OK)
            limit = blo;
        }
      }
    }
    if (lane == 0) {
      P.radius[c] = radius;
      P.reason[c] = reason;
      if (P.n_in)
        P.n_in[c] = n_in_out;
    }
  }
}


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
DeviceBuffer g_counter, g_pf_counts;
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
  if (st == SIF_OK)
    st = g_counter.reserve(sizeof(uint32_t));
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
  P.work_counter = g_counter.as<uint32_t>();
  P.radius = radius;
  P.reason = reason;
  P.n_in = n_in;
  timing_begin(SIF_AMD_STAGE_GROW);
  SIF_HIP(hipMemsetAsync(g_counter.ptr, 0, sizeof(uint32_t), stream()));
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  uint64_t blocks = ((uint64_t)n + GROW_WAVES - 1) / GROW_WAVES;
  const uint64_t cap = (uint64_t)cus * 2;
  if (blocks > cap)
    blocks = cap;
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
