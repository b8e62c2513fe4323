/*
 * TEST INFRASTRUCTURE ONLY -- oracle restatement of the exodus finder of
 * sif v0.5.0: candidate scan and sort, mesh template, radius rescaling,
 * crossing witness, overlap tests, mask painting and the commit. See
 * oracle.h for who may use it. Built with -ffp-contract=off.
 */
#include "oracle.h"
#include "oracle_rescale.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#  include <omp.h>
#endif

#define PI_F ((float)3.14159265358979323846)

static inline int bit_get(const uint64_t* m, uint64_t i) {
  return (int)((m[i >> 6] >> (i & 63)) & 1u);
}

/* --- candidates: src/finder/utils.c:123-213 --- */

uint64_t orc_scan_candidates(const float* delta, const uint64_t* mask,
  uint64_t total, float t, orc_candidate_t* out) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < total; i++) {
    if (delta[i] <= t && !(mask && bit_get(mask, i))) {
      if (out) {
        out[k].delta = delta[i];
        out[k].refined = 0;
        out[k].flat_idx = i;
      }
      k++;
    }
  }
  return k;
}

/* --- minima: src/finder/utils.c:228-262, 406-552 (minima_only, no
 * refinement) ---
 *
 * The candidates of SIF_FINDER_SEEDS_MINIMA: of the cells above, those whose
 * value is below each of the 26 periodic neighbours', a neighbour of lower
 * flat index winning a tie. On their nodes, at their own depth. */
static int orc_is_minimum(const float* f, uint32_t n, uint64_t i) {
  const uint64_t nn = (uint64_t)n * n;
  const int64_t c[3] = {(int64_t)(i / nn), (int64_t)((i / n) % n), (int64_t)(i % n)};
  const float v = f[i];
  for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
      for (int d = -1; d <= 1; d++) {
        if (!a && !b && !d)
          continue;
        const uint64_t x = (uint64_t)((c[0] + a + n) % n);
        const uint64_t y = (uint64_t)((c[1] + b + n) % n);
        const uint64_t z = (uint64_t)((c[2] + d + n) % n);
        const uint64_t j = x * nn + y * n + z;
        const float u = f[j];
        if (j < i ? !(v < u) : !(v <= u))
          return 0;
      }
  return 1;
}

uint64_t orc_scan_minima(const float* delta, uint32_t n, const uint64_t* mask,
  float t, orc_candidate_t* out) {
  const uint64_t total = (uint64_t)n * n * n;
  uint64_t k = 0;
  for (uint64_t i = 0; i < total; i++) {
    if (delta[i] <= t && !(mask && bit_get(mask, i)) && orc_is_minimum(delta, n, i)) {
      if (out) {
        out[k].delta = delta[i];
        out[k].refined = 0;
        out[k].flat_idx = i;
      }
      k++;
    }
  }
  return k;
}

/* include/sif/utils/sort.h:56-99 (three-way partition, middle pivot, the
 * serial permutation that the task-parallel form also produces, :101-145),
 * with the predicate a.delta < b.delta of utils.c:24. */
static void ref_partition(orc_candidate_t* a, int64_t lo, int64_t hi,
  int64_t* lt_out, int64_t* gt_out) {
  int64_t lt = lo, gt = hi, i = lo;
  const orc_candidate_t pivot = a[lo + (hi - lo) / 2];
  while (i <= gt) {
    const int is_less = a[i].delta < pivot.delta;
    const int is_greater = pivot.delta < a[i].delta;
    if (is_less) {
      orc_candidate_t t = a[lt];
      a[lt] = a[i];
      a[i] = t;
      lt++;
      i++;
    } else if (is_greater) {
      orc_candidate_t t = a[i];
      a[i] = a[gt];
      a[gt] = t;
      gt--;
    } else {
      i++;
    }
  }
  *lt_out = lt;
  *gt_out = gt;
}

static void ref_quicksort(orc_candidate_t* a, int64_t lo, int64_t hi) {
  while (lo < hi) {
    int64_t lt, gt;
    ref_partition(a, lo, hi, &lt, &gt);
    /* recurse on the smaller side; the permutation is the same as the
     * reference's recursion order since the two sides are disjoint */
    if (lt - 1 - lo < hi - (gt + 1)) {
      ref_quicksort(a, lo, lt - 1);
      lo = gt + 1;
    } else {
      ref_quicksort(a, gt + 1, hi);
      hi = lt - 1;
    }
  }
}

void orc_sort_candidates_ref(orc_candidate_t* c, uint64_t n) {
  if (n > 1)
    ref_quicksort(c, 0, (int64_t)n - 1);
}

static int cmp_flat(const void* pa, const void* pb) {
  const orc_candidate_t* a = pa;
  const orc_candidate_t* b = pb;
  if (a->delta < b->delta)
    return -1;
  if (b->delta < a->delta)
    return 1;
  return (a->flat_idx > b->flat_idx) - (a->flat_idx < b->flat_idx);
}

void orc_sort_candidates_flat(orc_candidate_t* c, uint64_t n) {
  qsort(c, n, sizeof(*c), cmp_flat);
}

/* --- template: src/finder/exodus_rescale.c:175-283 --- */

int orc_template_build(orc_template_t* t, uint32_t mesh_cells, float l,
  float rmin, float r_search) {
  memset(t, 0, sizeof(*t));
  const int32_t c = (int32_t)(r_search / l) + 1;
  if (c >= (int32_t)mesh_cells)
    return -3;
  const uint64_t cap = (uint64_t)(2 * c + 1) * (2 * c + 1) * (2 * c + 1);
  t->dx = malloc(cap * sizeof(int32_t));
  t->dy = malloc(cap * sizeof(int32_t));
  t->dz = malloc(cap * sizeof(int32_t));
  t->type = malloc(cap);
  t->min_d2 = malloc(cap * sizeof(float));
  t->max_d2 = malloc(cap * sizeof(float));
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
        const uint32_t k = t->count++;
        t->dx[k] = dx;
        t->dy[k] = dy;
        t->dz[k] = dz;
        t->min_d2[k] = lo;
        t->max_d2[k] = hi;
        if (hi <= rc2)
          t->type[k] = ORC_CELL_CORE;
        else if (lo > rc2)
          t->type[k] = hi <= rs2 ? ORC_CELL_SHELL : ORC_CELL_EDGE;
        else
          t->type[k] = ORC_CELL_STRADDLES;
      }
  return 0;
}

void orc_template_free(orc_template_t* t) {
  free(t->dx);
  free(t->dy);
  free(t->dz);
  free(t->type);
  free(t->min_d2);
  free(t->max_d2);
  memset(t, 0, sizeof(*t));
}

/* --- rescaling: exodus_rescale.c:295-369, 391-507, 642-1012 --- */

/* The mesh cell of template entry i, as query_cell() finds it. */
static uint64_t orc_query_flat(const query_t* q, uint32_t i, int32_t N) {
  int32_t ix = q->ix + q->t->dx[i], iy = q->iy + q->t->dy[i],
          iz = q->iz + q->t->dz[i];
  ix = ix < 0 ? ix + N : ix >= N ? ix - N : ix;
  iy = iy < 0 ? iy + N : iy >= N ? iy - N : iy;
  iz = iz < 0 ? iz + N : iz >= N ? iz - N : iz;
  return (uint64_t)ix * N * N + (uint64_t)iy * N + iz;
}

static int cmp_float(const void* pa, const void* pb) {
  const float a = *(const float*)pa, b = *(const float*)pb;
  return (a > b) - (a < b);
}

/* exodus_rescale.c:391-507. Returns 1 found, 0 exhausted. cur_w is read
 * when weighted, cur_n always (the enclosed count reported). */
static int resolve(const query_t* q, uint32_t blo, uint32_t bhi, float lo,
  float hi, float rc2, float rs2, float inv_w, uint32_t nb, uint64_t cur_n,
  double cur_w, float K2, float* out, uint64_t* n_in_out, int weighted) {
  const orc_template_t* t = q->t;
  uint32_t cap = 1024, cnt = 0;
  pair_t* buf = malloc(sizeof(pair_t) * cap);
  for (uint32_t i = 0; i < t->count; i++) {
    const uint8_t ty = t->type[i];
    if (ty == ORC_CELL_CORE || t->max_d2[i] < lo || t->min_d2[i] > hi)
      continue;
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(q, i, &s, &e, &ex, &ey, &ez))
      continue;
    for (uint64_t p = s; p < e; p++) {
      const float d2 = dist2(q->m->x[p], q->m->y[p], q->m->z[p], ex, ey, ez);
      if (ty != ORC_CELL_SHELL && !(d2 > rc2 && d2 <= rs2))
        continue;
      const uint32_t b = bin_of(d2, rc2, inv_w, nb);
      if (b < blo || b > bhi)
        continue;
      if (cnt == cap) {
        cap *= 2;
        buf = realloc(buf, sizeof(pair_t) * cap);
      }
      buf[cnt].d2 = d2;
      buf[cnt].w = weighted ? q->m->w[p] : 1.0f;
      cnt++;
    }
  }
  if (cnt > 1) {
    if (weighted) {
      qsort(buf, cnt, sizeof(pair_t), cmp_pair);
    } else {
      float* d = malloc(sizeof(float) * cnt);
      for (uint32_t k = 0; k < cnt; k++)
        d[k] = buf[k].d2;
      qsort(d, cnt, sizeof(float), cmp_float);
      for (uint32_t k = 0; k < cnt; k++)
        buf[k].d2 = d[k];
      free(d);
    }
  }
  int found = 0;
  for (int64_t k = (int64_t)cnt - 1; k >= 0; k--) {
    const float d2 = buf[k].d2;
    const double w = (double)buf[k].w;
    if (d2 == 0.0f) {
      cur_n--;
      cur_w -= w;
      continue;
    }
    const uint64_t n_in = cur_n - 1;
    const float d2c = d2 * d2 * d2;
    const float rn = weighted ? (float)(cur_w - w) : (float)n_in;
    if (rn * rn <= K2 * d2c) {
      *out = sqrtf(d2);
      if (n_in_out)
        *n_in_out = n_in;
      found = 1;
      break;
    }
    cur_n--;
    cur_w -= w;
  }
  free(buf);
  return found;
}

/* exodus_rescale.c:642-1012, find_exact_radius_impl, both forms. */
static float find_radius_impl(const orc_mesh_t* m, float cx, float cy,
  float cz, float r_search, float threshold, float vol_factor,
  float mass_factor, float rmin, const orc_template_t* tpl, uint32_t window,
  uint8_t* reason, uint64_t* n_in_out, int weighted) {
  *reason = ORC_RESCALE_OK;
  const float inv_l = 1.0f / m->cell_length;
  const query_t q = {m, tpl, cx, cy, cz, (int32_t)(cx * inv_l),
    (int32_t)(cy * inv_l), (int32_t)(cz * inv_l)};
  const float* mw = m->w;
  const float* cw = m->cell_w;
  const float rs2 = r_search * r_search;
  const float rc2 = rmin * rmin;
  const float span = rs2 - rc2;
  uint32_t nb = 2048;
  {
    const float exp_shell =
      vol_factor * (r_search * r_search * r_search - rmin * rmin * rmin);
    const float want = exp_shell / 16.0f;
    if (!(want >= 64.0f))
      nb = 64;
    else if (want < 2048.0f)
      nb = (uint32_t)want;
  }
  if (!(span > 0.0f)) {
    *reason = ORC_RESCALE_DEGENERATE;
    return -1.0f;
  }
  /* early refusal, :711-782; weighted, whole cells from the cell table */
  {
    const float expected = mass_factor * r_search * r_search * r_search;
    uint64_t within = 0;
    double w_within = 0.0;
    for (uint32_t i = 0; i < tpl->count; i++) {
      const uint8_t ty = tpl->type[i];
      uint64_t s, e;
      float ex, ey, ez;
      if (!query_cell(&q, i, &s, &e, &ex, &ey, &ez))
        continue;
      if (ty == ORC_CELL_CORE || ty == ORC_CELL_SHELL) {
        if (weighted) {
          const int32_t N = (int32_t)m->n_cells;
          const uint64_t flat = orc_query_flat(&q, i, N);
          w_within += (double)cw[flat];
        } else {
          within += e - s;
        }
        continue;
      }
      for (uint64_t p = s; p < e; p++) {
        const int in = dist2(m->x[p], m->y[p], m->z[p], ex, ey, ez) <= rs2;
        if (weighted) {
          if (in)
            w_within += (double)mw[p];
        } else {
          within += in;
        }
      }
    }
    if (weighted ? !(w_within > 0.0) : within == 0) {
      *reason = ORC_RESCALE_EMPTY;
      return -1.0f;
    }
    const float inside = weighted ? (float)w_within : (float)within;
    if ((inside / expected) - 1.0f <= threshold) {
      *reason = ORC_RESCALE_BEYOND;
      return -1.0f;
    }
  }
  const float inv_w = (float)nb / span;
  const float bin_w = span / (float)nb;
  uint32_t* bins = calloc(nb, sizeof(uint32_t));
  double* wbins = calloc(nb, sizeof(double));
  uint64_t n_core = 0, n_shell = 0;
  double w_core = 0.0, w_shell = 0.0;
  for (uint32_t i = 0; i < tpl->count; i++) {
    const uint8_t ty = tpl->type[i];
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(&q, i, &s, &e, &ex, &ey, &ez))
      continue;
    if (ty == ORC_CELL_CORE) {
      n_core += e - s;
      if (weighted)
        w_core += (double)cw[orc_query_flat(&q, i, (int32_t)m->n_cells)];
    } else if (ty == ORC_CELL_SHELL) {
      for (uint64_t p = s; p < e; p++) {
        const uint32_t b =
          bin_of(dist2(m->x[p], m->y[p], m->z[p], ex, ey, ez), rc2, inv_w, nb);
        bins[b]++;
        if (weighted) {
          wbins[b] += (double)mw[p];
          w_shell += (double)mw[p];
        }
      }
      n_shell += e - s;
    } else {
      for (uint64_t p = s; p < e; p++) {
        const float d2 = dist2(m->x[p], m->y[p], m->z[p], ex, ey, ez);
        const double w = weighted ? (double)mw[p] : 0.0;
        if (d2 <= rc2) {
          n_core++;
          w_core += w;
        } else if (d2 <= rs2) {
          const uint32_t b = bin_of(d2, rc2, inv_w, nb);
          bins[b]++;
          n_shell++;
          wbins[b] += w;
          w_shell += w;
        }
      }
    }
  }
  const uint64_t total = n_core + n_shell;
  const double total_w = w_core + w_shell;
  const float expected = mass_factor * r_search * r_search * r_search;
  if (weighted ? !(total_w > 0.0) : total == 0) {
    free(bins);
    free(wbins);
    *reason = ORC_RESCALE_EMPTY;
    return -1.0f;
  }
  if (((weighted ? (float)total_w : (float)total) / expected) - 1.0f <= threshold) {
    free(bins);
    free(wbins);
    *reason = ORC_RESCALE_BEYOND;
    return -1.0f;
  }
  const float K = (threshold + 1.0f) * mass_factor;
  const float K2 = K * K;
  if (window < 1)
    window = 1;
  uint64_t above = 0;
  double above_w = 0.0;
  for (int32_t b = (int32_t)nb - 1; b >= 0; b--) {
    const uint32_t cnt = bins[b];
    const double cnt_w = wbins[b];
    if (weighted ? !(cnt_w > 0.0) : cnt == 0) {
      above += cnt; /* weightless tracers: skipped, still counted above */
      continue;
    }
    float hi = rc2 + (float)(b + 1) * bin_w;
    if (b == (int32_t)nb - 1)
      hi = rs2 * (1.0f + 1e-6f);
    const float n_min = weighted ? (float)(total_w - above_w - cnt_w)
                                 : (float)(total - above - cnt);
    if (n_min * n_min > K2 * hi * hi * hi) {
      above += cnt;
      above_w += cnt_w;
      continue;
    }
    int32_t blo = b - (int32_t)(window - 1);
    if (blo < 0)
      blo = 0;
    uint64_t wcount = 0;
    double wweight = 0.0;
    for (int32_t bb = blo; bb <= b; bb++) {
      wcount += bins[bb];
      wweight += wbins[bb];
    }
    const float wlo = rc2 + (float)blo * bin_w;
    float r = -1.0f;
    if (resolve(&q, (uint32_t)blo, (uint32_t)b, wlo, hi, rc2, rs2, inv_w, nb,
          total - above, total_w - above_w, K2, &r, n_in_out, weighted)) {
      free(bins);
      free(wbins);
      return r;
    }
    above += wcount;
    above_w += wweight;
    b = blo;
  }
  free(bins);
  free(wbins);
  *reason = ORC_RESCALE_BELOW_RUNG;
  return -1.0f;
}

float orc_find_radius(const orc_mesh_t* m, float cx, float cy, float cz,
  float r_search, float threshold, float vol_factor, float rmin,
  const orc_template_t* tpl, uint32_t window, uint8_t* reason,
  uint64_t* n_in_out) {
  return find_radius_impl(m, cx, cy, cz, r_search, threshold, vol_factor,
    vol_factor, rmin, tpl, window, reason, n_in_out, 0);
}

float orc_find_radius_weighted(const orc_mesh_t* m, float cx, float cy,
  float cz, float r_search, float threshold, float vol_factor,
  float mass_factor, float rmin, const orc_template_t* tpl, uint32_t window,
  uint8_t* reason, uint64_t* n_in_out) {
  return find_radius_impl(m, cx, cy, cz, r_search, threshold, vol_factor,
    mass_factor, rmin, tpl, window, reason, n_in_out, 1);
}

/* exodus_rescale.c:1074-1167 */
int orc_witness(const orc_mesh_t* m, float cx, float cy, float cz, float rho2,
  float r_search, float threshold, float vol_factor, const orc_template_t* tpl,
  int closed) {
  const float inv_l = 1.0f / m->cell_length;
  const query_t q = {m, tpl, cx, cy, cz, (int32_t)(cx * inv_l),
    (int32_t)(cy * inv_l), (int32_t)(cz * inv_l)};
  const float rs2 = r_search * r_search;
  if (!(rho2 < rs2))
    return 0;
  const float K = (threshold + 1.0f) * vol_factor;
  const float K2 = K * K;
  const float slack = 1e-5f;
  const float bound = K2 * (rho2 * rho2 * rho2) * (1.0f - slack);
  const double n_cap = sqrt((double)bound) + 2.0;
  const float whole_in = rho2 * (1.0f - slack);
  const float whole_out = rho2 * (1.0f + slack);
  uint64_t n = 0, within = 0;
  int beyond = 0;
  for (uint32_t i = 0; i < tpl->count; i++) {
    const uint8_t ty = tpl->type[i];
    if (!closed && beyond && tpl->min_d2[i] > whole_out)
      continue;
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(&q, i, &s, &e, &ex, &ey, &ez))
      continue;
    const uint64_t cnt = e - s;
    if (ty == ORC_CELL_CORE || tpl->max_d2[i] <= whole_in) {
      n += cnt;
      within += cnt;
    } else if (ty == ORC_CELL_SHELL && tpl->min_d2[i] > whole_out) {
      within += cnt;
      beyond = 1;
    } else {
      for (uint64_t p = s; p < e; p++) {
        const float d2 = dist2(m->x[p], m->y[p], m->z[p], ex, ey, ez);
        if (d2 <= rho2) {
          n++;
          within++;
        } else if (ty == ORC_CELL_SHELL || d2 <= rs2) {
          within++;
          beyond = 1;
        }
      }
    }
    if ((double)n > n_cap)
      return 0;
  }
  const float rn = (float)n;
  if (!beyond || !(rn * rn <= bound))
    return 0;
  if (closed) {
    const double expected =
      (double)vol_factor * (double)r_search * (double)r_search * (double)r_search;
    if (!((double)within > (1.0 + (double)threshold) * expected * (1.0 + 1e-5)))
      return 0;
  }
  return 1;
}

static int cmp_double(const void* pa, const void* pb) {
  const double a = *(const double*)pa, b = *(const double*)pb;
  return (a > b) - (a < b);
}

/* tests/test_finders.c:433-465 */
double orc_brute_force_radius(const float* x, const float* y, const float* z,
  uint64_t n, float box, double cx, double cy, double cz, double rung,
  double r_search, double threshold) {
  const double L = (double)box, half = 0.5 * L;
  double* shell = malloc(sizeof(double) * (n ? n : 1));
  uint64_t n_core = 0, n_shell = 0;
  for (uint64_t p = 0; p < n; p++) {
    double dx = fabs((double)x[p] - cx);
    double dy = fabs((double)y[p] - cy);
    double dz = fabs((double)z[p] - cz);
    dx = dx > half ? L - dx : dx;
    dy = dy > half ? L - dy : dy;
    dz = dz > half ? L - dz : dz;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d <= rung)
      n_core++;
    else if (d <= r_search)
      shell[n_shell++] = d;
  }
  qsort(shell, n_shell, sizeof(double), cmp_double);
  const double pi = 3.14159265358979323846;
  const double k = (1.0 + threshold) * (double)n / (L * L * L) * 4.0 / 3.0 * pi;
  double r = -1.0;
  for (uint64_t i = n_shell; i-- > 0;) {
    const double d = shell[i];
    if ((double)(n_core + i) <= k * d * d * d) {
      r = d;
      break;
    }
  }
  free(shell);
  return r;
}

/* --- overlap and mask: src/finder/utils.c:819-1033 --- */

static inline uint32_t wrap_pbc(int32_t v, uint32_t n) {
  const int32_t w = v % (int32_t)n;
  return (uint32_t)(w < 0 ? w + (int32_t)n : w);
}

int orc_overlap_quick(const uint64_t* mask, uint32_t n, uint32_t ix,
  uint32_t iy, uint32_t iz, uint32_t r) {
  const uint64_t nn = n;
#define FLAT(a, b, c) (((uint64_t)(a) * nn + (b)) * nn + (c))
  if (bit_get(mask, FLAT(wrap_pbc((int32_t)(ix + r), n), iy, iz)))
    return 1;
  if (bit_get(mask, FLAT(wrap_pbc((int32_t)ix - (int32_t)r, n), iy, iz)))
    return 1;
  if (bit_get(mask, FLAT(ix, wrap_pbc((int32_t)(iy + r), n), iz)))
    return 1;
  if (bit_get(mask, FLAT(ix, wrap_pbc((int32_t)iy - (int32_t)r, n), iz)))
    return 1;
  if (bit_get(mask, FLAT(ix, iy, wrap_pbc((int32_t)(iz + r), n))))
    return 1;
  if (bit_get(mask, FLAT(ix, iy, wrap_pbc((int32_t)iz - (int32_t)r, n))))
    return 1;
#undef FLAT
  return 0;
}

/* utils.c:965-1039: the nodes of a sphere set (sif__mark_sphere) or cleared
 * (sif__clear_sphere) */
static void paint_nodes(uint64_t* mask, float cx, float cy, float cz, float r,
  uint32_t n, float l, int set) {
  const float r2 = r * r;
  const float inv_l = 1.0f / l;
  const float L = (float)n * l;
  const float half = L * 0.5f;
  const int32_t c = (int32_t)(r * inv_l) + 1;
  const int32_t ox = (int32_t)(cx * inv_l), oy = (int32_t)(cy * inv_l),
                oz = (int32_t)(cz * inv_l);
  for (int32_t dx = -c; dx <= c; dx++) {
    const int32_t gx = ox + dx;
    float ddx = fabsf((float)gx * l - cx);
    ddx = ddx > half ? L - ddx : ddx;
    const float dx2 = ddx * ddx;
    if (dx2 > r2)
      continue;
    const uint32_t wx = wrap_pbc(gx, n);
    for (int32_t dy = -c; dy <= c; dy++) {
      const int32_t gy = oy + dy;
      float ddy = fabsf((float)gy * l - cy);
      ddy = ddy > half ? L - ddy : ddy;
      const float dxy2 = dx2 + ddy * ddy;
      if (dxy2 > r2)
        continue;
      const uint32_t wy = wrap_pbc(gy, n);
      for (int32_t dz = -c; dz <= c; dz++) {
        const int32_t gz = oz + dz;
        float ddz = fabsf((float)gz * l - cz);
        ddz = ddz > half ? L - ddz : ddz;
        if (dxy2 + ddz * ddz <= r2) {
          const uint64_t cell =
            ((uint64_t)wx * n + wy) * (uint64_t)n + wrap_pbc(gz, n);
          if (set)
            mask[cell >> 6] |= 1ull << (cell & 63);
          else
            mask[cell >> 6] &= ~(1ull << (cell & 63));
        }
      }
    }
  }
}

void orc_mark_sphere(uint64_t* mask, float cx, float cy, float cz, float r,
  uint32_t n, float l) {
  paint_nodes(mask, cx, cy, cz, r, n, l, 1);
}

/* --- void index: a cell linked list over a 32^3 grid,
 * src/structures/cell_linked_list.c:181-210 and exodus_finder.c:215 --- */

#define VOID_CELLS 32

typedef struct {
  int32_t head[VOID_CELLS * VOID_CELLS * VOID_CELLS];
  int32_t* next;
  uint64_t cap;
  float inv_l;
} void_index_t;

/* cell_linked_list.c:41-59, periodic */
static int32_t vcell(float v, float inv_l) {
  int32_t c = (int32_t)floorf(v * inv_l);
  c %= VOID_CELLS;
  return c < 0 ? c + VOID_CELLS : c;
}

static void vindex_insert(void_index_t* vi, uint64_t k, float x, float y,
  float z) {
  if (k >= vi->cap) {
    vi->cap = vi->cap ? 2 * vi->cap : 1024;
    while (vi->cap <= k)
      vi->cap *= 2;
    vi->next = realloc(vi->next, sizeof(int32_t) * vi->cap);
  }
  const int32_t c =
    (vcell(x, vi->inv_l) * VOID_CELLS + vcell(y, vi->inv_l)) * VOID_CELLS +
    vcell(z, vi->inv_l);
  vi->next[k] = vi->head[c];
  vi->head[c] = (int32_t)k;
}

/* cell_linked_list.c:212-232: unlinked from its cell's chain */
static void vindex_remove(void_index_t* vi, uint64_t k, float x, float y, float z) {
  const int32_t c =
    (vcell(x, vi->inv_l) * VOID_CELLS + vcell(y, vi->inv_l)) * VOID_CELLS +
    vcell(z, vi->inv_l);
  int32_t* link = &vi->head[c];
  while (*link != -1 && *link != (int32_t)k)
    link = &vi->next[*link];
  if (*link != -1)
    *link = vi->next[k];
}

/* The window of 32^3 cells within `reach` of x, whole along an axis it spans
 * (exodus_finder.c:706-717) */
static void vwindow(const void_index_t* vi, const float x[3], float reach,
  int32_t lo[3], int32_t hi[3]) {
  for (int a = 0; a < 3; a++) {
    lo[a] = (int32_t)floorf((x[a] - reach) * vi->inv_l);
    hi[a] = (int32_t)floorf((x[a] + reach) * vi->inv_l);
    if (hi[a] - lo[a] >= VOID_CELLS) {
      lo[a] = 0;
      hi[a] = VOID_CELLS - 1;
    }
  }
}

/* utils.c:850-960 */
static int overlap_exact(const orc_catalogue_t* cat, const void_index_t* vi,
  float cx, float cy, float cz, float r, float max_r, float box, float f,
  float scale) {
  const float half = box * 0.5f;
  const float sr = scale * (r + max_r - f * (r < max_r ? r : max_r));
  const int32_t lo[3] = {(int32_t)floorf((cx - sr) * vi->inv_l),
    (int32_t)floorf((cy - sr) * vi->inv_l),
    (int32_t)floorf((cz - sr) * vi->inv_l)};
  const int32_t hi[3] = {(int32_t)floorf((cx + sr) * vi->inv_l),
    (int32_t)floorf((cy + sr) * vi->inv_l),
    (int32_t)floorf((cz + sr) * vi->inv_l)};
  const int interior = (cx - sr >= 0.0f) && (cx + sr < box) &&
                       (cy - sr >= 0.0f) && (cy + sr < box) &&
                       (cz - sr >= 0.0f) && (cz + sr < box);
  for (int32_t a = lo[0]; a <= hi[0]; a++) {
    const uint32_t wa = wrap_pbc(a, VOID_CELLS);
    for (int32_t b = lo[1]; b <= hi[1]; b++) {
      const uint32_t wb = wrap_pbc(b, VOID_CELLS);
      for (int32_t c = lo[2]; c <= hi[2]; c++) {
        const uint32_t wc = wrap_pbc(c, VOID_CELLS);
        for (int32_t i = vi->head[(wa * VOID_CELLS + wb) * VOID_CELLS + wc];
             i != -1; i = vi->next[i]) {
          const float ri = cat->r[i];
          const float mr = r < ri ? r : ri;
          const float tot = scale * (r + ri - f * mr);
          const float t2 = tot * tot;
          if (interior) {
            const float dx = cx - cat->cx[i], dy = cy - cat->cy[i],
                        dz = cz - cat->cz[i];
            if (dx * dx + dy * dy + dz * dz < t2)
              return 1;
          } else {
            float dx = fabsf(cx - cat->cx[i]);
            dx = dx > half ? box - dx : dx;
            const float dx2 = dx * dx;
            if (dx2 >= t2)
              continue;
            float dy = fabsf(cy - cat->cy[i]);
            dy = dy > half ? box - dy : dy;
            const float dy2 = dy * dy;
            if (dx2 + dy2 >= t2)
              continue;
            float dz = fabsf(cz - cat->cz[i]);
            dz = dz > half ? box - dz : dz;
            if (dx2 + dy2 + dz * dz < t2)
              return 1;
          }
        }
      }
    }
  }
  return 0;
}

/* utils.c:850-960 for many queries against one catalogue, its 32^3 list
 * built as the reference builds it (insertion in catalogue order). */
void orc_overlap_exact_many(const float* vcx, const float* vcy, const float* vcz,
  const float* vr, uint64_t nv, float max_r, const float* qx, const float* qy,
  const float* qz, const float* qr, uint64_t nq, float box, float f, float scale,
  uint8_t* out) {
  void_index_t* vi = calloc(1, sizeof(void_index_t));
  for (int32_t c = 0; c < VOID_CELLS * VOID_CELLS * VOID_CELLS; c++)
    vi->head[c] = -1;
  vi->inv_l = (float)VOID_CELLS / box;
  for (uint64_t k = 0; k < nv; k++)
    vindex_insert(vi, k, vcx[k], vcy[k], vcz[k]);
  orc_catalogue_t cat;
  memset(&cat, 0, sizeof(cat));
  cat.cx = (float*)vcx;
  cat.cy = (float*)vcy;
  cat.cz = (float*)vcz;
  cat.r = (float*)vr;
  cat.n = nv;
  for (uint64_t q = 0; q < nq; q++)
    out[q] = (uint8_t)overlap_exact(&cat, vi, qx[q], qy[q], qz[q], qr[q], max_r, box, f,
      scale);
  free(vi->next);
  free(vi);
}

/* exodus_finder.c:779-845, with min_rj = 0 (no displacement) */
static double collision_radius(const orc_catalogue_t* cat,
  const void_index_t* vi, float cx, float cy, float cz, float r_hi,
  float max_r, float f, float box, float scale) {
  const double L = (double)box, half = 0.5 * L, ff = (double)f;
  const float reach = scale * (r_hi + max_r);
  const int32_t n = VOID_CELLS;
  const float x[3] = {cx, cy, cz};
  int32_t lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (int32_t)floorf((x[a] - reach) * vi->inv_l);
    hi[a] = (int32_t)floorf((x[a] + reach) * vi->inv_l);
    if (hi[a] - lo[a] >= n) {
      lo[a] = 0;
      hi[a] = n - 1;
    }
  }
  double best = INFINITY;
  for (int32_t a = lo[0]; a <= hi[0]; a++) {
    const int32_t wa = ((a % n) + n) % n;
    for (int32_t b = lo[1]; b <= hi[1]; b++) {
      const int32_t wb = ((b % n) + n) % n;
      for (int32_t c = lo[2]; c <= hi[2]; c++) {
        const int32_t wc = ((c % n) + n) % n;
        for (int32_t i = vi->head[(wa * n + wb) * n + wc]; i != -1;
             i = vi->next[i]) {
          const double rj = (double)cat->r[i];
          double dx = fabs((double)cx - (double)cat->cx[i]);
          double dy = fabs((double)cy - (double)cat->cy[i]);
          double dz = fabs((double)cz - (double)cat->cz[i]);
          dx = dx > half ? L - dx : dx;
          dy = dy > half ? L - dy : dy;
          dz = dz > half ? L - dz : dz;
          const double D =
            (sqrt(dx * dx + dy * dy + dz * dz) + 1e-5 * (L + (double)r_hi + rj)) /
            (double)scale;
          double rs;
          if (D <= rj)
            rs = 0.0;
          else if (D <= (2.0 - ff) * rj)
            rs = (D - rj) / (1.0 - ff);
          else
            rs = D - (1.0 - ff) * rj;
          if (rs < best)
            best = rs;
        }
      }
    }
  }
  return best;
}

static int cat_append(orc_catalogue_t* c, float x, float y, float z, float r,
  uint64_t n_in, uint32_t rung) {
  if (c->n == c->capacity) {
    c->capacity = c->capacity ? 2 * c->capacity : 1024;
    c->cx = realloc(c->cx, sizeof(float) * c->capacity);
    c->cy = realloc(c->cy, sizeof(float) * c->capacity);
    c->cz = realloc(c->cz, sizeof(float) * c->capacity);
    c->r = realloc(c->r, sizeof(float) * c->capacity);
    c->n_in = realloc(c->n_in, sizeof(uint64_t) * c->capacity);
    c->rung = realloc(c->rung, sizeof(uint32_t) * c->capacity);
  }
  c->cx[c->n] = x;
  c->cy[c->n] = y;
  c->cz[c->n] = z;
  c->r[c->n] = r;
  c->n_in[c->n] = n_in;
  c->rung[c->n] = rung;
  c->n++;
  return 0;
}

void orc_catalogue_free(orc_catalogue_t* c) {
  free(c->cx);
  free(c->cy);
  free(c->cz);
  free(c->r);
  free(c->n_in);
  free(c->rung);
  memset(c, 0, sizeof(*c));
}

/* --- displacement: exodus_finder.c:684-1037 --- */

/* find_conflicts (:696-750): the voids a sphere collides with, in the
 * window's cell order and each chain's order, and the largest radius among
 * them (-1 for none). */
static uint64_t find_conflicts(const orc_catalogue_t* cat, const void_index_t* vi,
  float cx, float cy, float cz, float r, float f, float box, float max_r,
  float scale, uint64_t** list, uint64_t* cap, float* largest) {
  const float half = 0.5f * box;
  const float reach = scale * (r + max_r - f * (r < max_r ? r : max_r));
  const float x[3] = {cx, cy, cz};
  int32_t lo[3], hi[3];
  vwindow(vi, x, reach, lo, hi);
  uint64_t m = 0;
  *largest = -1.0f;
  for (int32_t a = lo[0]; a <= hi[0]; a++)
    for (int32_t b = lo[1]; b <= hi[1]; b++)
      for (int32_t c = lo[2]; c <= hi[2]; c++) {
        const uint32_t cell = (wrap_pbc(a, VOID_CELLS) * VOID_CELLS +
                                wrap_pbc(b, VOID_CELLS)) * VOID_CELLS +
                              wrap_pbc(c, VOID_CELLS);
        for (int32_t i = vi->head[cell]; i != -1; i = vi->next[i]) {
          float dx = fabsf(cx - cat->cx[i]);
          float dy = fabsf(cy - cat->cy[i]);
          float dz = fabsf(cz - cat->cz[i]);
          dx = dx > half ? box - dx : dx;
          dy = dy > half ? box - dy : dy;
          dz = dz > half ? box - dz : dz;
          const float ri = cat->r[i];
          const float tot = scale * (r + ri - f * (r < ri ? r : ri));
          if (dx * dx + dy * dy + dz * dz >= tot * tot)
            continue;
          if (m == *cap) {
            *cap = *cap ? 2 * *cap : 64;
            *list = realloc(*list, sizeof(uint64_t) * *cap);
          }
          (*list)[m++] = (uint64_t)i;
          if (ri > *largest)
            *largest = ri;
        }
      }
  return m;
}

/* displace_void (:946-1002): out of the index, its radius negated, its nodes
 * cleared, then the nodes of every standing void it intersects set again. */
static void displace_void(orc_catalogue_t* cat, void_index_t* vi, uint64_t* mask,
  uint64_t i, float max_r, float scale, uint32_t n, float l, float box) {
  const float cx = cat->cx[i], cy = cat->cy[i], cz = cat->cz[i], r = cat->r[i];
  vindex_remove(vi, i, cx, cy, cz);
  cat->r[i] = -r;
  paint_nodes(mask, cx, cy, cz, scale * r, n, l, 0);
  const float half = 0.5f * box;
  const float x[3] = {cx, cy, cz};
  int32_t lo[3], hi[3];
  vwindow(vi, x, scale * (r + max_r), lo, hi);
  for (int32_t a = lo[0]; a <= hi[0]; a++)
    for (int32_t b = lo[1]; b <= hi[1]; b++)
      for (int32_t c = lo[2]; c <= hi[2]; c++) {
        const uint32_t cell = (wrap_pbc(a, VOID_CELLS) * VOID_CELLS +
                                wrap_pbc(b, VOID_CELLS)) * VOID_CELLS +
                              wrap_pbc(c, VOID_CELLS);
        for (int32_t j = vi->head[cell]; j != -1; j = vi->next[j]) {
          float dx = fabsf(cx - cat->cx[j]);
          float dy = fabsf(cy - cat->cy[j]);
          float dz = fabsf(cz - cat->cz[j]);
          dx = dx > half ? box - dx : dx;
          dy = dy > half ? box - dy : dy;
          dz = dz > half ? box - dz : dz;
          const float t = scale * (r + cat->r[j]);
          if (dx * dx + dy * dy + dz * dz < t * t)
            paint_nodes(mask, cat->cx[j], cat->cy[j], cat->cz[j], scale * cat->r[j],
              n, l, 1);
        }
      }
}

/* compact_catalogue (:1004-1032): the displaced voids dropped, the rest
 * re-indexed in catalogue order. */
static void compact_catalogue(orc_catalogue_t* cat, void_index_t* vi) {
  uint64_t j = 0;
  for (uint64_t i = 0; i < cat->n; i++) {
    if (!(cat->r[i] > 0.0f))
      continue;
    cat->cx[j] = cat->cx[i];
    cat->cy[j] = cat->cy[i];
    cat->cz[j] = cat->cz[i];
    cat->r[j] = cat->r[i];
    cat->n_in[j] = cat->n_in[i];
    cat->rung[j] = cat->rung[i];
    j++;
  }
  cat->n = j;
  for (int32_t c = 0; c < VOID_CELLS * VOID_CELLS * VOID_CELLS; c++)
    vi->head[c] = -1;
  for (uint64_t i = 0; i < cat->n; i++)
    vindex_insert(vi, i, cat->cx[i], cat->cy[i], cat->cz[i]);
}

/* A candidate's speculative evaluation against the frozen catalogue. */
typedef struct {
  uint8_t status; /* 1 refused (final), 2 rescale failed, 3 radius found */
  uint8_t why;    /* for status 1: 0 proxy, 1 overlap, 2 witness */
  float r;
  uint64_t n_in;
} eval_t;

#define BATCH 4096u

/* The exodus run of exodus_finder.c:1585-2492 for a periodic box, unweighted,
 * SIF_FINDER_SEEDS_CELLS on their nodes, no extended reach or displacement.
 * Candidates are evaluated in batches against the catalogue as the batch
 * began (exodus_finder.c:1969-2084) and committed in order against the
 * catalogue as it stands (2181-2329), with every rung-radius test repeated
 * at commit; the catalogue is the one a strictly sequential pass gives. */
/* rmesh set: a survey run (exodus_finder.c:1592, 2045-2049), each radius
 * grown against the randoms by orc_find_radius_survey with K, no witness. */
static int exodus_run_impl(float* const* fields, orc_field_fn field_fn, void* field_ctx,
  uint32_t n, float box,
  const orc_mesh_t* mesh, const orc_mesh_t* rmesh, float survey_K, uint64_t n_particles,
  const float* radii, uint32_t n_radii, float t, float f, int lagrangian,
  uint32_t search_opt, int tie_rule, int use_witness, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats) {
  const uint64_t total = (uint64_t)n * n * n;
  const float l = box / (float)n;
  uint64_t* mask = calloc((total + 63) / 64, sizeof(uint64_t));
  orc_candidate_t* cand = malloc(sizeof(orc_candidate_t) * total);
  void_index_t* vi = calloc(1, sizeof(void_index_t));
  memset(vi->head, 0xff, sizeof(vi->head));
  vi->inv_l = (float)VOID_CELLS / box; /* cell_linked_list.c:92 */
  eval_t* ev = malloc(sizeof(eval_t) * BATCH);
  uint64_t* bidx = malloc(sizeof(uint64_t) * BATCH);

  const float box_volume = box * box * box;
  const float mean_density = (float)n_particles / box_volume;
  const float vol = (4.0f / 3.0f) * PI_F * mean_density;
  /* exodus_finder.c:1639-1642: a weighted mesh's density tests use the mean
   * weight density; the witness is never run on one (:1868) */
  const int weighted = mesh->w != NULL;
  const float mass =
    weighted ? (4.0f / 3.0f) * PI_F * ((float)mesh->total_weight / box_volume)
             : vol;
  if (weighted || rmesh)
    use_witness = 0;
  const float scale = lagrangian ? (float)cbrt(1.0 + (double)t) : 1.0f;
  const float factors[4] = {1.5f, 1.25f, 1.75f, 2.0f};
  const float sf = factors[(search_opt >> 12) & 3u];
  float max_r = 0.0f;
  int status = 0;
  /* the witness is a shortcut of the commit; under displacement the commit
   * runs in full, without it */
  const int displace = (search_opt & ORC_DISPLACE) != 0;
  if (displace)
    use_witness = 0;
  uint64_t* conflicts = NULL;
  uint64_t conflicts_cap = 0;

  /* exodus_finder.c:234-244, 1670-1702, 1342-1353: the extended reach's
   * levels, from the run's largest search radius down past the bottom rung
   * in steps of 1.5; level k's core 1.2 below level k - 1's reach */
  uint32_t n_ext = 0;
  float ext_reach[32], ext_rmin[32];
  orc_template_t etpl[32];
  if (search_opt & ORC_EXTEND_REACH) {
    float r_max = radii[0];
    for (uint32_t i = 1; i < n_radii; i++)
      if (radii[i] > r_max)
        r_max = radii[i];
    const float r_ext = r_max + fmaxf((sf - 1.0f) * r_max, 3.0f * l);
    const float r_low = radii[n_radii - 1];
    while (n_ext < 32u && (n_ext == 0 || r_ext / powf(1.5f, (float)(n_ext - 1)) > r_low))
      n_ext++;
    for (uint32_t k = 0; k < n_ext; k++)
      ext_reach[k] = r_ext / powf(1.5f, (float)(n_ext - 1 - k));
    for (uint32_t k = 0; k < n_ext; k++) {
      ext_rmin[k] = (k ? ext_reach[k - 1] : ext_reach[0] / 1.5f) / 1.2f;
      if (orc_template_build(&etpl[k], mesh->n_cells, mesh->cell_length, ext_rmin[k],
            ext_reach[k]) != 0) {
        for (uint32_t j = 0; j < k; j++)
          orc_template_free(&etpl[j]);
        free(mask);
        free(cand);
        free(vi);
        free(ev);
        free(bidx);
        return -3;
      }
    }
  }

  for (uint32_t i = 0; i < n_radii && status == 0; i++) {
    const float radius = radii[i];
    orc_rung_stats_t st;
    memset(&st, 0, sizeof(st));
    const float* field = fields ? fields[i] : field_fn(field_ctx, i);
    if (!field) {
      status = -5;
      break;
    }
    const uint64_t nc = (search_opt & ORC_SEEDS_MINIMA)
                          ? orc_scan_minima(field, n, mask, t, cand)
                          : orc_scan_candidates(field, mask, total, t, cand);
    st.candidates = nc;
    if (tie_rule == 0)
      orc_sort_candidates_ref(cand, nc);
    else
      orc_sort_candidates_flat(cand, nc);
    const float rmin = 1.00f * radius;
    float r_search = radius + fmaxf((sf - 1.0f) * radius, 3.0f * l);
    if (i > 0 && radii[i - 1] > r_search)
      r_search = radii[i - 1];
    orc_template_t tpl, rtpl;
    memset(&rtpl, 0, sizeof(rtpl));
    if (orc_template_build(&tpl, mesh->n_cells, mesh->cell_length, rmin,
          r_search) != 0) {
      status = -3;
      break;
    }
    /* exodus_finder.c:1834-1840: the random mesh's own template */
    if (rmesh && orc_template_build(&rtpl, rmesh->n_cells, rmesh->cell_length, rmin,
                   r_search) != 0) {
      orc_template_free(&tpl);
      status = -3;
      break;
    }
    const int32_t proxy =
      (int32_t)(scale * radius * (1.0f - f) / l) - 1;
    const float rc2 = rmin * rmin;
    /* exodus_finder.c:1824-1832: the first level this rung enters */
    uint32_t ext0 = n_ext;
    for (uint32_t k = 0; k < n_ext && ext0 == n_ext; k++)
      if (ext_reach[k] >= 1.2f * r_search)
        ext0 = k;
    if (ext0 == n_ext && n_ext > 0 && ext_reach[n_ext - 1] > r_search)
      ext0 = n_ext - 1;

    uint64_t k = displace ? nc : 0;
    /* SIF_FINDER_DISPLACE (exodus_finder.c:1871-1880, 2210-2275): no
     * prefilter; each candidate, in order, tested at the rung against the
     * catalogue as it stands, grown, and then refused only by a void at least
     * as large, displacing the smaller ones it collides with. The reference's
     * batches end at each displacement, so its commit is this sequence. */
    for (uint64_t q = 0; displace && q < nc; q++) {
      const uint64_t flat = cand[q].flat_idx;
      if (bit_get(mask, flat)) {
        st.masked++;
        continue;
      }
      const uint32_t ix = (uint32_t)(flat / ((uint64_t)n * n));
      const uint32_t iy = (uint32_t)((flat / n) % n);
      const uint32_t iz = (uint32_t)(flat % n);
      if (proxy > 0 && orc_overlap_quick(mask, n, ix, iy, iz, (uint32_t)proxy)) {
        st.proxy++;
        continue;
      }
      const float cx = (float)ix * l, cy = (float)iy * l, cz = (float)iz * l;
      if (overlap_exact(cat, vi, cx, cy, cz, radius, max_r, box, f, scale)) {
        st.overlap++;
        continue;
      }
      uint8_t reason;
      uint64_t n_in = 0;
      float r = weighted
        ? orc_find_radius_weighted(mesh, cx, cy, cz, r_search, t, vol, mass, rmin,
            &tpl, window, &reason, &n_in)
        : orc_find_radius(mesh, cx, cy, cz, r_search, t, vol, rmin, &tpl, window,
            &reason, &n_in);
      for (uint32_t x = ext0; x < n_ext && r < 0.0f && reason == ORC_RESCALE_BEYOND;
           x++) {
        n_in = 0;
        r = weighted
          ? orc_find_radius_weighted(mesh, cx, cy, cz, ext_reach[x], t, vol, mass,
              ext_rmin[x], &etpl[x], window, &reason, &n_in)
          : orc_find_radius(mesh, cx, cy, cz, ext_reach[x], t, vol, ext_rmin[x],
              &etpl[x], window, &reason, &n_in);
      }
      if (r < 0.0f) {
        st.rescale_fail++;
        continue;
      }
      float largest;
      const uint64_t m = find_conflicts(cat, vi, cx, cy, cz, r, f, box, max_r, scale,
        &conflicts, &conflicts_cap, &largest);
      if (largest >= r) {
        st.overlap++;
        continue;
      }
      for (uint64_t c = 0; c < m; c++)
        displace_void(cat, vi, mask, conflicts[c], max_r, scale, n, l, box);
      st.displaced += m;
      cat_append(cat, cx, cy, cz, r, n_in, i);
      vindex_insert(vi, cat->n - 1, cx, cy, cz);
      orc_mark_sphere(mask, cx, cy, cz, scale * r, n, l);
      if (r > max_r)
        max_r = r;
      st.accepted++;
    }
    if (displace)
      compact_catalogue(cat, vi);
    while (k < nc) {
      /* phase 0: masked nodes, sequentially */
      uint32_t bc = 0;
      while (k < nc && bc < BATCH) {
        if (bit_get(mask, cand[k].flat_idx))
          st.masked++;
        else
          bidx[bc++] = k;
        k++;
      }
      if (bc == 0)
        break;
      /* phase 1: speculative, against the catalogue frozen here */
#pragma omp parallel for schedule(dynamic, 16)
      for (uint32_t b = 0; b < bc; b++) {
        const uint64_t flat = cand[bidx[b]].flat_idx;
        const uint32_t ix = (uint32_t)(flat / ((uint64_t)n * n));
        const uint32_t iy = (uint32_t)((flat / n) % n);
        const uint32_t iz = (uint32_t)(flat % n);
        eval_t* e = &ev[b];
        e->status = 0;
        if (proxy > 0 && orc_overlap_quick(mask, n, ix, iy, iz, (uint32_t)proxy)) {
          e->status = 1;
          e->why = 0;
          continue;
        }
        const float cx = (float)ix * l, cy = (float)iy * l, cz = (float)iz * l;
        if (overlap_exact(cat, vi, cx, cy, cz, radius, max_r, box, f, scale)) {
          e->status = 1;
          e->why = 1;
          continue;
        }
        if (use_witness) {
          const double rho =
            collision_radius(cat, vi, cx, cy, cz, r_search, max_r, f, box, scale);
          if (rho < (double)r_search) {
            float rho2 = (float)(rho * rho * (1.0 + 1e-5));
            if (rho2 < rc2)
              rho2 = rc2;
            if (orc_witness(mesh, cx, cy, cz, rho2, r_search, t, vol, &tpl, 0)) {
              e->status = 1;
              e->why = 2;
              continue;
            }
          }
        }
        uint8_t reason;
        uint64_t n_in = 0;
        float r = rmesh
          ? orc_find_radius_survey(mesh, rmesh, &tpl, &rtpl, cx, cy, cz, r_search,
              survey_K, rmin, window, &reason, &n_in)
          : weighted
          ? orc_find_radius_weighted(mesh, cx, cy, cz, r_search, t, vol, mass,
              rmin, &tpl, window, &reason, &n_in)
          : orc_find_radius(mesh, cx, cy, cz, r_search, t, vol, rmin, &tpl,
              window, &reason, &n_in);
        /* exodus_finder.c:2060-2072: beyond the reach, the next level */
        for (uint32_t x = ext0; x < n_ext && r < 0.0f && reason == ORC_RESCALE_BEYOND;
             x++) {
          n_in = 0;
          r = weighted
            ? orc_find_radius_weighted(mesh, cx, cy, cz, ext_reach[x], t, vol,
                mass, ext_rmin[x], &etpl[x], window, &reason, &n_in)
            : orc_find_radius(mesh, cx, cy, cz, ext_reach[x], t, vol,
                ext_rmin[x], &etpl[x], window, &reason, &n_in);
        }
        e->status = r < 0.0f ? 2 : 3;
        e->r = r;
        e->n_in = n_in;
      }
      /* phase 2: commit in order */
      for (uint32_t b = 0; b < bc; b++) {
        const eval_t* e = &ev[b];
        const uint64_t flat = cand[bidx[b]].flat_idx;
        if (e->status == 1) {
          if (e->why == 0)
            st.proxy++;
          else if (e->why == 1)
            st.overlap++;
          else
            st.witness++;
          continue;
        }
        const uint32_t ix = (uint32_t)(flat / ((uint64_t)n * n));
        const uint32_t iy = (uint32_t)((flat / n) % n);
        const uint32_t iz = (uint32_t)(flat % n);
        const float cx = (float)ix * l, cy = (float)iy * l, cz = (float)iz * l;
        /* the rung-radius tests against the catalogue as it now stands */
        if (bit_get(mask, flat)) {
          st.masked++;
          continue;
        }
        if (proxy > 0 && orc_overlap_quick(mask, n, ix, iy, iz, (uint32_t)proxy)) {
          st.proxy++;
          continue;
        }
        if (overlap_exact(cat, vi, cx, cy, cz, radius, max_r, box, f, scale)) {
          st.overlap++;
          continue;
        }
        if (e->status == 2) {
          st.rescale_fail++;
          continue;
        }
        const float r = e->r;
        const int32_t pole = (int32_t)(scale * r * (1.0f - f) / l) - 1;
        if (pole > 0 && orc_overlap_quick(mask, n, ix, iy, iz, (uint32_t)pole)) {
          st.proxy++;
          continue;
        }
        if (overlap_exact(cat, vi, cx, cy, cz, r, max_r, box, f, scale)) {
          st.overlap++;
          continue;
        }
        cat_append(cat, cx, cy, cz, r, e->n_in, i);
        vindex_insert(vi, cat->n - 1, cx, cy, cz);
        orc_mark_sphere(mask, cx, cy, cz, scale * r, n, l);
        if (r > max_r)
          max_r = r;
        st.accepted++;
      }
    }
    orc_template_free(&tpl);
    if (rmesh)
      orc_template_free(&rtpl);
    if (stats)
      stats[i] = st;
  }
  for (uint32_t k = 0; k < n_ext; k++)
    orc_template_free(&etpl[k]);
  free(conflicts);

  free(mask);
  free(cand);
  free(vi->next);
  free(vi);
  free(ev);
  free(bidx);
  return status;
}

int orc_exodus_run(float* const* fields, uint32_t n, float box,
  const orc_mesh_t* mesh, uint64_t n_particles, const float* radii,
  uint32_t n_radii, float t, float f, int lagrangian, uint32_t search_opt,
  int tie_rule, int use_witness, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats) {
  return exodus_run_impl(fields, NULL, NULL, n, box, mesh, NULL, 0.0f, n_particles, radii,
    n_radii, t, f, lagrangian, search_opt, tie_rule, use_witness, window, cat, stats);
}

int orc_exodus_run_lazy(orc_field_fn field, void* ctx, uint32_t n, float box,
  const orc_mesh_t* mesh, uint64_t n_particles, const float* radii,
  uint32_t n_radii, float t, float f, int lagrangian, uint32_t search_opt,
  int tie_rule, int use_witness, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats) {
  return exodus_run_impl(NULL, field, ctx, n, box, mesh, NULL, 0.0f, n_particles, radii,
    n_radii, t, f, lagrangian, search_opt, tie_rule, use_witness, window, cat, stats);
}

/* sif__exodus_run for a survey (exodus_finder.c:1585-2492 with a random
 * mesh): fields[i] holds the survey contrast of rung i (orc_survey_contrast
 * of the smoothed data and randoms); K = (1 + t) alpha, alpha the data's
 * total weight over the randoms' (exodus_finder.c:1649-1653). Cells seeds on
 * their nodes, no extended reach or displacement (exodus_finder.c:1114-1137).
 * The footprint columns are orc_survey_void_footprint's, per void. */
int orc_exodus_run_survey(float* const* fields, uint32_t n, float box,
  const orc_mesh_t* data_mesh, const orc_mesh_t* random_mesh, float survey_K,
  const float* radii_desc, uint32_t n_radii, float threshold, float overlap_fraction,
  int lagrangian, uint32_t search_opt, int tie_rule, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats) {
  if (search_opt & (ORC_SEEDS_MINIMA | ORC_EXTEND_REACH | ORC_DISPLACE))
    return -2;
  return exodus_run_impl(fields, NULL, NULL, n, box, data_mesh, random_mesh, survey_K, 0,
    radii_desc,
    n_radii, threshold, overlap_fraction, lagrangian, search_opt, tie_rule, 0, window, cat,
    stats);
}
