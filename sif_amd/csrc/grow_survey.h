// Device code of the survey rescaling (exodus_rescale.c:1909-2317): one
// wavefront per candidate, on two meshes -- the data and the randoms, each
// with its own cells and template -- through the same template walk as the
// box's growth (grow_device.h).
//
// The test at a data distance d is D(<d) <= K R(<d), K = (1 + t) alpha: the
// data weight enclosed against the random weight, with R interpolated
// linearly in d^2 inside each histogram bin. Two launches:
//
//   survey_totals_kernel  the weight of both meshes inside r_search (whole
//                         CORE and SHELL cells from the cell table or the
//                         count), the early refusals (no randoms; beyond the
//                         search sphere), and the candidate's bin count, which
//                         the data tracers inside set (exodus_rescale.c:
//                         2214-2225);
//   survey_grow_kernel    the two histograms of the annulus on those bins,
//                         the skip test per bin, and the exact pass: the data
//                         tracers of a window of bins sorted in LDS and walked
//                         outside in against the interpolated random weight.
//
// The host sizes the second launch's LDS for the largest bin count of the
// batch. A window holds as many bins as fit REFINE_CAP data tracers; a single
// bin of more is refused with R_OVERFLOW.
#pragma once

#include "grow_device.h"

namespace sif_amd {

enum { R_NO_RANDOMS = 6 };

struct SurveyParams {
  const float* cx;
  const float* cy;
  const float* cz;
  uint32_t n_cand;
  float rmin, r_search, K;
  float* radius;
  uint32_t* reason;
  uint64_t* n_in;
  uint32_t* nb;      // per candidate: its bin count, 0 once refused
  uint32_t nb_max;   // the LDS of survey_grow_kernel holds this many bins
};

// The weight of a whole cell: the cell table's, or its tracer count.
__device__ __forceinline__ double survey_cell_w(const GrowParams& P, uint64_t cell,
  uint32_t cnt) {
  return P.cw ? (double)P.cw[cell] : (double)cnt;
}

__device__ __forceinline__ bool survey_centre(const GrowParams& P, float cx, float cy,
  float cz, Query* q) {
  q->cx = cx;
  q->cy = cy;
  q->cz = cz;
  q->ix = (int32_t)__fmul_rn(cx, P.inv_l);
  q->iy = (int32_t)__fmul_rn(cy, P.inv_l);
  q->iz = (int32_t)__fmul_rn(cz, P.inv_l);
  return cx >= 0.0f && cx <= P.box && cy >= 0.0f && cy <= P.box && cz >= 0.0f &&
         cz <= P.box && q->ix <= P.mesh_n && q->iy <= P.mesh_n && q->iz <= P.mesh_n;
}

// exodus_rescale.c:1961-2001: the weight and tracers inside r_search.
template <bool WT>
__device__ __forceinline__ void survey_sphere_total(const GrowParams& P, const Query& q,
  WaveLDS& S, int lane, float rs2, double* w_out, uint64_t* n_out) {
  double w = 0.0;
  uint64_t n = 0;
  for_particles<WT>(P, q, S, lane,
    [](uint32_t, float2) { return true; },
    [&](uint32_t type, uint32_t cnt, uint64_t cell) {
      if (type == CELL_CORE || type == CELL_SHELL) {
        w += survey_cell_w(P, cell, cnt);
        n += cnt;
        return true;
      }
      return false;
    },
    [&](float d2, float wt, uint32_t, bool valid) {
      if (valid && d2 <= rs2) {
        w += (double)wt;
        n++;
      }
    });
  *w_out = unid(wave_sum_f64(w));
  *n_out = uni(wave_sum_u64(n));
}

// exodus_rescale.c:2008-2066: one mesh's histogram of the annulus in weight
// (wbins, nb doubles; counts in cnt when given), with its core and annulus
// weights and the tracers of the core. Whole CORE cells from the table.
template <bool WT>
__device__ __forceinline__ void survey_histogram(const GrowParams& P, const Query& q,
  WaveLDS& S, int lane, double* wbins, uint32_t* cnt, uint32_t nb, float rc2, float rs2,
  float inv_w, double* core_out, double* shell_out, uint64_t* n_core_out) {
  for (uint32_t b = lane; b < nb; b += 64) {
    wbins[b] = 0.0;
    if (cnt)
      cnt[b] = 0;
  }
  wave_sync();
  double core = 0.0, shell = 0.0;
  uint64_t n_core = 0;
  for_particles<WT>(P, q, S, lane,
    [](uint32_t, float2) { return true; },
    [&](uint32_t type, uint32_t c, uint64_t cell) {
      if (type == CELL_CORE) {
        core += survey_cell_w(P, cell, c);
        n_core += c;
        return true;
      }
      return false;
    },
    [&](float d2, float wt, uint32_t type, bool valid) {
      if (!valid)
        return;
      bool annulus;
      if (type == CELL_SHELL) {
        annulus = true;
      } else if (d2 <= rc2) {
        core += (double)wt;
        n_core++;
        annulus = false;
      } else {
        annulus = d2 <= rs2;
      }
      if (annulus) {
        const uint32_t b = bin_of(d2, rc2, inv_w, nb);
        // one wave owns these bins: the additions follow its program and
        // lane order
        atomicAdd(&wbins[b], (double)wt);
        if (cnt)
          atomicAdd(&cnt[b], 1u);
        shell += (double)wt;
      }
    });
  wave_sync();
  *core_out = unid(wave_sum_f64(core));
  *shell_out = unid(wave_sum_f64(shell));
  *n_core_out = uni(wave_sum_u64(n_core));
}

// exodus_rescale.c:2121-2160: the window's pairs key/val[0..cnt) in ascending
// order, walked from the top. Element k encloses d_top less the weight of
// every pair from k up, itself included, against the random weight below its
// bin (rpre) plus its bin's weight times where it sits in the bin, in d^2;
// it passes when R > 0 and D <= K R. Returns the highest passing index
// (uniform), or -1.
__device__ __forceinline__ int survey_walk(const float* key, const float* val, uint32_t cnt,
  double d_top, float K, float rc2, float inv_w, float bin_w, uint32_t nb,
  const double* rbins, const double* rpre, int lane) {
  double carry = 0.0;  // uniform: weight of the blocks above
  for (int32_t top = (int32_t)cnt; top > 0; top -= 64) {
    const int32_t k = top - 64 + lane;
    float d2 = 0.0f, w = 0.0f;
    if (k >= 0) {
      d2 = key[k];
      w = val[k];
    }
    double s = (double)w;  // weight from k to the top of the block
    for (int off = 1; off < 64; off <<= 1) {
      const double y = __shfl_down(s, off, 64);
      if (lane + off < 64)
        s += y;
    }
    int f = -1;
    if (k >= 0 && d2 != 0.0f) {
      const uint32_t b = bin_of(d2, rc2, inv_w, nb);
      float frac =
        __fmul_rn(__fsub_rn(d2, __fadd_rn(rc2, __fmul_rn((float)b, bin_w))), inv_w);
      frac = frac < 0.0f ? 0.0f : (frac > 1.0f ? 1.0f : frac);
      const float r_in = (float)__dadd_rn(rpre[b], __dmul_rn(rbins[b], (double)frac));
      const float d_in = (float)(d_top - (carry + s));
      if (r_in > 0.0f && d_in <= __fmul_rn(K, r_in))
        f = (int)k;
    }
    f = uni(wave_max_i32(f));
    if (f >= 0)
      return f;
    carry += unid(__shfl(s, 0, 64));
  }
  return -1;
}

// The first launch: sphere totals and refusals, the bin count.
template <bool DW, bool RW>
__global__ void __launch_bounds__(64, 4) survey_totals_kernel(GrowParams PD, GrowParams PR,
  SurveyParams Q) {
  const int lane = threadIdx.x & 63;
  const uint32_t c = blockIdx.x;
  if (c >= Q.n_cand)
    return;
  WaveLDS S{};
  const float rs2 = __fmul_rn(Q.r_search, Q.r_search);
  const float rc2 = __fmul_rn(Q.rmin, Q.rmin);
  const float span = __fsub_rn(rs2, rc2);
  uint32_t reason = R_OK, nb = 0;
  Query qd, qr;
  const float cx = Q.cx[c], cy = Q.cy[c], cz = Q.cz[c];
  const bool in_d = survey_centre(PD, cx, cy, cz, &qd);
  const bool in_r = survey_centre(PR, cx, cy, cz, &qr);
  if (!in_d || !in_r || !(span > 0.0f)) {
    reason = R_DEGENERATE;
  } else {
    double d_within, r_within;
    uint64_t nd_within, nr_within;
    survey_sphere_total<DW>(PD, qd, S, lane, rs2, &d_within, &nd_within);
    survey_sphere_total<RW>(PR, qr, S, lane, rs2, &r_within, &nr_within);
    if (!(r_within > 0.0)) {
      reason = R_NO_RANDOMS;
    } else if ((float)d_within <= __fmul_rn(Q.K, (float)r_within)) {
      reason = R_BEYOND;
    } else {
      // exodus_rescale.c:2214-2225, float steps as written
      const float rm3 = __fmul_rn(__fmul_rn(Q.rmin, Q.rmin), Q.rmin);
      const float rs3 = __fmul_rn(__fmul_rn(Q.r_search, Q.r_search), Q.r_search);
      const float core_fraction = __fdiv_rn(rm3, rs3);
      const float want =
        __fdiv_rn(__fmul_rn((float)nd_within, __fsub_rn(1.0f, core_fraction)), 16.0f);
      nb = (uint32_t)MAX_BINS;
      if (!(want >= 64.0f))
        nb = 64;
      else if (want < (float)MAX_BINS)
        nb = (uint32_t)want;
    }
  }
  if (lane == 0) {
    Q.nb[c] = reason == R_OK ? nb : 0u;
    Q.reason[c] = reason;
    Q.radius[c] = -1.0f;
    if (Q.n_in)
      Q.n_in[c] = 0;
  }
}

// LDS words a wave of survey_grow_kernel holds for nb bins: three double
// arrays (data and random weights, the randoms' running total), the data
// counts, REFINE_CAP (distance, weight) pairs.
SIF_GROW_HD uint32_t survey_wave_words(uint32_t nb) {
  const uint32_t nb4 = (nb + 3u) & ~3u;
  return 3u * 2u * nb4 + nb4 + 2u * (uint32_t)REFINE_CAP;
}

// The second launch: histograms, skip flags, windows.
template <bool DW, bool RW>
__global__ void __launch_bounds__(64, 4) survey_grow_kernel(GrowParams PD, GrowParams PR,
  SurveyParams Q) {
  GROW_DYNAMIC_LDS(lds_words);
  const int lane = threadIdx.x & 63;
  const uint32_t c = blockIdx.x;
  if (c >= Q.n_cand)
    return;
  const uint32_t nb = uni(Q.nb[c]);
  if (nb == 0 || nb > Q.nb_max)
    return;  // refused by the first launch (its reason stands)
  const uint32_t nb4 = (Q.nb_max + 3u) & ~3u;
  double* dbins = reinterpret_cast<double*>(lds_words);
  double* rbins = dbins + nb4;
  double* rpre = rbins + nb4;
  uint32_t* dcnt = reinterpret_cast<uint32_t*>(rpre + nb4);
  float* buf = reinterpret_cast<float*>(dcnt + nb4);
  float* buf_w = buf + REFINE_CAP;
  WaveLDS S{};

  const float rs2 = __fmul_rn(Q.r_search, Q.r_search);
  const float rc2 = __fmul_rn(Q.rmin, Q.rmin);
  const float span = __fsub_rn(rs2, rc2);
  const float inv_w = __fdiv_rn((float)nb, span);
  const float bin_w = __fdiv_rn(span, (float)nb);
  const float K = Q.K;
  Query qd, qr;
  const float cx = Q.cx[c], cy = Q.cy[c], cz = Q.cz[c];
  survey_centre(PD, cx, cy, cz, &qd);
  survey_centre(PR, cx, cy, cz, &qr);

  double d_core, d_shell, r_core, r_shell;
  uint64_t nd_core, unused;
  survey_histogram<DW>(PD, qd, S, lane, dbins, dcnt, nb, rc2, rs2, inv_w, &d_core, &d_shell,
    &nd_core);
  survey_histogram<RW>(PR, qr, S, lane, rbins, nullptr, nb, rc2, rs2, inv_w, &r_core,
    &r_shell, &unused);
  const double d_total = d_core + d_shell;
  const double r_total = r_core + r_shell;
  float radius = -1.0f;
  uint32_t reason = R_OK;
  uint64_t n_in_out = 0;
  // exodus_rescale.c:2244-2254: the backstops
  if (!(r_total > 0.0)) {
    reason = R_NO_RANDOMS;
  } else if ((float)d_total <= __fmul_rn(K, (float)r_total)) {
    reason = R_BEYOND;
  } else {
    // lane L owns bins [L per, (L + 1) per): running random weight below
    // each bin (from the core up), and the skip flags, top down
    const uint32_t per = (nb + 63) / 64;
    const uint32_t b0 = lane * per;
    const uint32_t b1 = b0 + per < nb ? b0 + per : nb;
    double mine_d = 0.0, mine_r = 0.0;
    uint64_t mine_n = 0;
    for (uint32_t b = b0; b < b1; b++) {
      mine_d += dbins[b];
      mine_r += rbins[b];
      mine_n += dcnt[b];
    }
    double below_r = mine_r;  // inclusive prefix over lanes, then exclusive
    for (int off = 1; off < 64; off <<= 1) {
      const double y = __shfl_up(below_r, off, 64);
      if (lane >= off)
        below_r += y;
    }
    below_r -= mine_r;
    double run = r_core + below_r;
    for (uint32_t b = b0; b < b1; b++) {
      rpre[b] = run;
      run += rbins[b];
    }
    double above_d = mine_d, above_r = mine_r;  // inclusive suffix, then exclusive
    for (int off = 1; off < 64; off <<= 1) {
      const double yd = __shfl_down(above_d, off, 64);
      const double yr = __shfl_down(above_r, off, 64);
      if (lane + off < 64) {
        above_d += yd;
        above_r += yr;
      }
    }
    above_d -= mine_d;
    above_r -= mine_r;
    for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--) {
      const double bd = dbins[b], br = rbins[b];
      if (bd > 0.0) {
        // exodus_rescale.c:2277-2290: the least data a tracer of this bin
        // can have inside it against the most randoms it can
        const float d_in_min = (float)(d_total - above_d - bd);
        const float r_in_max = (float)(r_total - above_r);
        if (!(!(r_in_max > 0.0f) || d_in_min > __fmul_rn(K, r_in_max)))
          dcnt[b] |= 0x80000000u;  // a bin the walk must open
      }
      above_d += bd;
      above_r += br;
    }
    const uint64_t n_all = nd_core + uni(wave_sum_u64(mine_n));
    wave_sync();

    // the walk: next flagged bin below `limit`, a window resolved, repeat
    int32_t limit = (int32_t)nb;
    while (reason == R_OK && radius < 0.0f) {
      int best = -1;
      for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--)
        if (b < limit && (dcnt[b] & 0x80000000u) != 0u) {
          best = b;
          break;
        }
      best = uni(wave_max_i32(best));
      if (best < 0) {
        reason = R_BELOW_RUNG;
        break;
      }
      // data weight and tracers at or below the top of bin `best`
      double abv_d = 0.0;
      uint64_t abv_n = 0;
      for (uint32_t b = (uint32_t)best + 1 + lane; b < nb; b += 64) {
        abv_d += dbins[b];
        abv_n += dcnt[b] & 0x7fffffffu;
      }
      const double d_top = d_total - unid(wave_sum_f64(abv_d));
      const uint64_t n_top = n_all - uni(wave_sum_u64(abv_n));
      // window [best - W + 1, best]: as many bins as fit the pair buffer
      uint32_t wcnt = 0;
      if (best - lane >= 0)
        wcnt = dcnt[best - lane] & 0x7fffffffu;
      const uint32_t wscan = wave_incl_scan_u32(wcnt, lane);
      const uint64_t fits = __ballot(best - lane >= 0 && wscan <= (uint32_t)REFINE_CAP);
      const int W = uni(fits ? __popcll(fits) : 0);
      if (W == 0) {
        reason = R_OVERFLOW;  // one bin holds more than the buffer
        break;
      }
      const int32_t blo = best - W + 1;
      float hi = __fadd_rn(rc2, __fmul_rn((float)(best + 1), bin_w));
      if (best == (int32_t)nb - 1)
        hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
      const float lo = __fadd_rn(rc2, __fmul_rn((float)blo, bin_w));
      const uint32_t win_total = uni(__shfl(wscan, W - 1, 64));
      uint32_t got = 0;  // uniform
      for_particles<DW>(PD, qd, S, lane,
        [&](uint32_t type, float2 bd) {
          return type != CELL_CORE && !(bd.y < lo) && !(bd.x > hi);
        },
        [](uint32_t, uint32_t, uint64_t) { return false; },
        [&](float d2, float wt, uint32_t type, bool valid) {
          bool take = valid && (type == CELL_SHELL || (d2 > rc2 && d2 <= rs2));
          if (take) {
            const uint32_t b = bin_of(d2, rc2, inv_w, nb);
            take = (int32_t)b >= blo && (int32_t)b <= best;
          }
          wave_append2(buf, buf_w, (uint32_t)REFINE_CAP, got, take, d2, wt, lane);
        });
      wave_sync();
      got = uni(got);
      if (got != win_total) {
        reason = R_OVERFLOW;  // counts disagree: never expected
        break;
      }
      wave_sort_pairs(buf, buf_w, got, lane);
      const int found =
        survey_walk(buf, buf_w, got, d_top, K, rc2, inv_w, bin_w, nb, rbins, rpre, lane);
      if (found >= 0) {
        radius = __builtin_sqrtf(unif(buf[found]));
        n_in_out = n_top - win_total + (uint64_t)found;
      } else {
        limit = blo;
      }
      wave_sync();
    }
  }
  if (lane == 0) {
    Q.radius[c] = radius;
    Q.reason[c] = reason;
    if (Q.n_in)
      Q.n_in[c] = n_in_out;
  }
}

}  // namespace sif_amd
