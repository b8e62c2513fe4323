/*
 * TEST INFRASTRUCTURE ONLY -- oracle restatement of the survey half of the
 * exodus finder of sif v0.5.0 (src/finder/exodus_survey.c and the survey
 * rescaling, src/finder/exodus_rescale.c:1909-2317): the footprint the
 * randoms mark out and its closing, the survey contrast, the survey radius
 * and the observed fraction of each void. See oracle.h for who may use it.
 * Built with -ffp-contract=off.
 *
 * Occupancy masks here are one byte a grid node (0 or 1), flat index
 * (ix n + iy) n + iz, in place of the reference's bitmask.
 */
#include "oracle.h"
#include "oracle_rescale.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* exodus_survey.c:40-44: the node nearest p on an axis, clamped. */
uint32_t orc_survey_node(float p, float inv_l, uint32_t n) {
  const uint32_t i = (uint32_t)(p * inv_l + 0.5f);
  return i < n ? i : n - 1;
}

/* exodus_survey.c:62-76: lambda from m = lambda / (1 - exp(-lambda)), by 64
 * bisections of [0, m]. */
double orc_survey_randoms_per_node(double m) {
  if (!(m > 1.0))
    return 0.0;
  double lo = 0.0, hi = m;
  for (int it = 0; it < 64; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid < m * (1.0 - exp(-mid)))
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

/* exodus_survey.c:85-122: one axis of a dilation (erode 0: a node is set
 * when any node within k along the axis is) or an erosion (erode 1: when all
 * 2k + 1 are; nodes past a face count as unset) of src into dst. */
static void close_pass(const uint8_t* src, uint8_t* dst, uint32_t n, int axis, int32_t k,
  int erode) {
  const uint64_t n2 = (uint64_t)n * n;
  const uint64_t stride = axis == 0 ? n2 : axis == 1 ? n : 1;
  const uint64_t step_a = axis == 0 ? n : n2;
  const uint64_t step_b = axis == 2 ? n : 1;
  const int32_t len = (int32_t)n;
  const int32_t full = 2 * k + 1;
  memset(dst, 0, (size_t)(n2 * n));
#pragma omp parallel for schedule(static)
  for (uint64_t line = 0; line < n2; line++) {
    const uint64_t base = (line / n) * step_a + (line % n) * step_b;
    int32_t count = 0;
    for (int32_t j = 0; j < k && j < len; j++)
      count += src[base + (uint64_t)j * stride];
    for (int32_t i = 0; i < len; i++) {
      if (i + k < len)
        count += src[base + (uint64_t)(i + k) * stride];
      if (i - k - 1 >= 0)
        count -= src[base + (uint64_t)(i - k - 1) * stride];
      if (erode ? count == full : count > 0)
        dst[base + (uint64_t)i * stride] = 1;
    }
  }
}

/* exodus_survey.c:130-143: dilate x, y, z, then erode x, y, z, the two masks
 * taking turns, the result back in occ. */
void orc_survey_close(uint8_t* occ, uint32_t n, int32_t k) {
  uint8_t* tmp = malloc((size_t)n * n * n);
  close_pass(occ, tmp, n, 0, k, 0);
  close_pass(tmp, occ, n, 1, k, 0);
  close_pass(occ, tmp, n, 2, k, 0);
  close_pass(tmp, occ, n, 0, k, 1);
  close_pass(occ, tmp, n, 1, k, 1);
  close_pass(tmp, occ, n, 2, k, 1);
  free(tmp);
}

/* exodus_survey.c:148-207 (sif__exodus_survey_occupancy): the nodes the
 * randoms reach, closed over a (2k + 1)^3 box, k the smallest in [1, 4] for
 * which lambda (2k + 1)^3 reaches 128. Returns k (0: no random marked). */
int orc_survey_occupancy(const float* rx, const float* ry, const float* rz, uint64_t nr,
  uint32_t n, float cell_length, uint8_t* occ, double* lambda_out, uint64_t* marked_out,
  uint64_t* observed_out) {
  const uint64_t total = (uint64_t)n * n * n;
  const float inv_l = 1.0f / cell_length;
  memset(occ, 0, (size_t)total);
  for (uint64_t p = 0; p < nr; p++) {
    const uint32_t ix = orc_survey_node(rx[p], inv_l, n);
    const uint32_t iy = orc_survey_node(ry[p], inv_l, n);
    const uint32_t iz = orc_survey_node(rz[p], inv_l, n);
    occ[((uint64_t)ix * n + iy) * n + iz] = 1;
  }
  uint64_t marked = 0;
  for (uint64_t c = 0; c < total; c++)
    marked += occ[c];
  if (marked_out)
    *marked_out = marked;
  if (lambda_out)
    *lambda_out = 0.0;
  if (observed_out)
    *observed_out = 0;
  if (marked == 0)
    return 0;
  const double lambda = orc_survey_randoms_per_node((double)nr / (double)marked);
  int32_t k = 1;
  while (k < 4 && lambda * (2 * k + 1) * (2 * k + 1) * (2 * k + 1) < 128.0)
    k++;
  orc_survey_close(occ, n, k);
  uint64_t observed = 0;
  for (uint64_t c = 0; c < total; c++)
    observed += occ[c];
  if (lambda_out)
    *lambda_out = lambda;
  if (observed_out)
    *observed_out = observed;
  return k;
}

/* exodus_survey.c:223-234: rho_d / (alpha rho_r) - 1 on observed nodes with
 * randoms expected, 1e30 elsewhere. */
void orc_survey_contrast(float* data, const float* randoms, const uint8_t* occ, float alpha,
  uint64_t total) {
  for (uint64_t c = 0; c < total; c++) {
    const float expected = alpha * randoms[c];
    data[c] = (occ[c] && expected > 0.0f) ? data[c] / expected - 1.0f : 1e30f;
  }
}

/* exodus_survey.c:247-303: the observed fraction of a void's sphere and of
 * its shell out to 2r, on a lattice of spacing max(r / 8, l / 2) through the
 * cube around it; a point outside [0, box)^3 is unobserved. */
void orc_survey_void_footprint(const uint8_t* occ, uint32_t n, float l, float box, float cx,
  float cy, float cz, float r, float* inside, float* shell) {
  const float inv_l = 1.0f / l;
  float h = 0.125f * r;
  if (h < 0.5f * l)
    h = 0.5f * l;
  const int32_t half = (int32_t)(2.0f * r / h) + 1;
  const float r2 = r * r;
  const float r2_out = 4.0f * r2;
  uint64_t n_in = 0, occ_in = 0, n_sh = 0, occ_sh = 0;
  for (int32_t a = -half; a < half; a++) {
    const float dx = ((float)a + 0.5f) * h;
    for (int32_t b = -half; b < half; b++) {
      const float dy = ((float)b + 0.5f) * h;
      for (int32_t c = -half; c < half; c++) {
        const float dz = ((float)c + 0.5f) * h;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > r2_out)
          continue;
        const float px = cx + dx, py = cy + dy, pz = cz + dz;
        int o = 0;
        if (px >= 0.0f && px < box && py >= 0.0f && py < box && pz >= 0.0f && pz < box) {
          const uint64_t f = ((uint64_t)orc_survey_node(px, inv_l, n) * n +
                               orc_survey_node(py, inv_l, n)) * n +
                             orc_survey_node(pz, inv_l, n);
          o = occ[f];
        }
        if (d2 <= r2) {
          n_in++;
          occ_in += (uint64_t)o;
        } else {
          n_sh++;
          occ_sh += (uint64_t)o;
        }
      }
    }
  }
  *inside = n_in ? (float)((double)occ_in / (double)n_in) : 0.0f;
  *shell = n_sh ? (float)((double)occ_sh / (double)n_sh) : 0.0f;
}

/* --- the survey rescaling: exodus_rescale.c:1909-2317 --- */

/* exodus_rescale.c:1945-1954 */
static inline double tracer_w(const orc_mesh_t* m, uint64_t p) {
  return m->w ? (double)m->w[p] : 1.0;
}
static inline double cell_w(const orc_mesh_t* m, uint64_t flat, uint64_t s, uint64_t e) {
  return m->cell_w ? (double)m->cell_w[flat] : (double)(e - s);
}

/* exodus_rescale.c:1961-2001: the weight (and tracers) inside r_search,
 * whole CORE and SHELL cells from the cell table. */
static void sphere_total(const query_t* q, float rs2, double* w_out, uint64_t* n_out) {
  const orc_template_t* t = q->t;
  const int32_t N = (int32_t)q->m->n_cells;
  double w = 0.0;
  uint64_t cnt = 0;
  for (uint32_t i = 0; i < t->count; i++) {
    const uint8_t ty = t->type[i];
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(q, i, &s, &e, &ex, &ey, &ez))
      continue;
    if (ty == ORC_CELL_CORE || ty == ORC_CELL_SHELL) {
      int32_t ix = q->ix + t->dx[i], iy = q->iy + t->dy[i], iz = q->iz + t->dz[i];
      ix = ix < 0 ? ix + N : ix >= N ? ix - N : ix;
      iy = iy < 0 ? iy + N : iy >= N ? iy - N : iy;
      iz = iz < 0 ? iz + N : iz >= N ? iz - N : iz;
      w += cell_w(q->m, ((uint64_t)ix * N + iy) * N + iz, s, e);
      cnt += e - s;
      continue;
    }
    for (uint64_t p = s; p < e; p++)
      if (dist2(q->m->x[p], q->m->y[p], q->m->z[p], ex, ey, ez) <= rs2) {
        w += tracer_w(q->m, p);
        cnt++;
      }
  }
  *w_out = w;
  *n_out = cnt;
}

/* exodus_rescale.c:2008-2066: one mesh's histogram of the annulus in weight,
 * with its core and annulus weights; whole CORE cells from the table. */
static void histogram(const query_t* q, double* wb, uint32_t nb, float rc2, float rs2,
  float inv_w, double* core_out, double* shell_out) {
  const orc_template_t* t = q->t;
  const int32_t N = (int32_t)q->m->n_cells;
  double core = 0.0, sh = 0.0;
  memset(wb, 0, sizeof(double) * nb);
  for (uint32_t i = 0; i < t->count; i++) {
    const uint8_t ty = t->type[i];
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(q, i, &s, &e, &ex, &ey, &ez))
      continue;
    if (ty == ORC_CELL_CORE) {
      int32_t ix = q->ix + t->dx[i], iy = q->iy + t->dy[i], iz = q->iz + t->dz[i];
      ix = ix < 0 ? ix + N : ix >= N ? ix - N : ix;
      iy = iy < 0 ? iy + N : iy >= N ? iy - N : iy;
      iz = iz < 0 ? iz + N : iz >= N ? iz - N : iz;
      core += cell_w(q->m, ((uint64_t)ix * N + iy) * N + iz, s, e);
    } else if (ty == ORC_CELL_SHELL) {
      for (uint64_t p = s; p < e; p++) {
        const float d2 = dist2(q->m->x[p], q->m->y[p], q->m->z[p], ex, ey, ez);
        const double w = tracer_w(q->m, p);
        wb[bin_of(d2, rc2, inv_w, nb)] += w;
        sh += w;
      }
    } else {
      for (uint64_t p = s; p < e; p++) {
        const float d2 = dist2(q->m->x[p], q->m->y[p], q->m->z[p], ex, ey, ez);
        const double w = tracer_w(q->m, p);
        if (d2 <= rc2) {
          core += w;
        } else if (d2 <= rs2) {
          wb[bin_of(d2, rc2, inv_w, nb)] += w;
          sh += w;
        }
      }
    }
  }
  *core_out = core;
  *shell_out = sh;
}

/* exodus_rescale.c:2076-2163: the data tracers of bins [blo, bhi], sorted and
 * walked outside in against the random weight interpolated in d^2 inside
 * each bin. d_top, r_top: data and random weight at or below the top of bhi;
 * n_top the data tracers there. Returns 1 found (radius, enclosed tracers),
 * 0 exhausted. */
static int survey_resolve(const query_t* dq, uint32_t blo, uint32_t bhi, float lo, float hi,
  float rc2, float rs2, float inv_w, float bin_w, uint32_t nb, double d_top, double r_top,
  uint64_t n_top, const double* rb, float K, float* out, uint64_t* n_in_out) {
  const orc_template_t* t = dq->t;
  uint32_t cap = 1024, cnt = 0;
  pair_t* buf = malloc(sizeof(pair_t) * cap);
  for (uint32_t i = 0; i < t->count; i++) {
    const uint8_t ty = t->type[i];
    if (ty == ORC_CELL_CORE || t->max_d2[i] < lo || t->min_d2[i] > hi)
      continue;
    uint64_t s, e;
    float ex, ey, ez;
    if (!query_cell(dq, i, &s, &e, &ex, &ey, &ez))
      continue;
    for (uint64_t p = s; p < e; p++) {
      const float d2 = dist2(dq->m->x[p], dq->m->y[p], dq->m->z[p], ex, ey, ez);
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
      buf[cnt].w = (float)tracer_w(dq->m, p);
      cnt++;
    }
  }
  if (cnt > 1)
    qsort(buf, cnt, sizeof(pair_t), cmp_pair);
  uint32_t cur = bhi;
  double r_below = r_top - rb[bhi];
  double d_cur = d_top;
  uint64_t n_cur = n_top;
  int found = 0;
  for (int64_t k = (int64_t)cnt - 1; k >= 0; k--) {
    const float d2 = buf[k].d2;
    const double w = (double)buf[k].w;
    if (d2 == 0.0f) {
      d_cur -= w;
      n_cur--;
      continue;
    }
    const uint32_t b = bin_of(d2, rc2, inv_w, nb);
    while (cur > b) {
      cur--;
      r_below -= rb[cur];
    }
    float frac = (d2 - (rc2 + (float)cur * bin_w)) * inv_w;
    frac = frac < 0.0f ? 0.0f : (frac > 1.0f ? 1.0f : frac);
    const float r_in = (float)(r_below + rb[cur] * (double)frac);
    const float d_in = (float)(d_cur - w);
    if (r_in > 0.0f && d_in <= K * r_in) {
      *out = sqrtf(d2);
      if (n_in_out)
        *n_in_out = n_cur - 1;
      found = 1;
      break;
    }
    d_cur -= w;
    n_cur--;
  }
  free(buf);
  return found;
}

/* exodus_rescale.c:2172-2317 (sif__exodus_find_radius_survey): the largest
 * data distance d inside r_search with D(<d) <= K R(<d). */
float orc_find_radius_survey(const orc_mesh_t* dm, const orc_mesh_t* rm,
  const orc_template_t* dtpl, const orc_template_t* rtpl, float cx, float cy, float cz,
  float r_search, float K, float rmin, uint32_t window, uint8_t* reason, uint64_t* n_in) {
  *reason = ORC_RESCALE_OK;
  const float rs2 = r_search * r_search;
  const float rc2 = rmin * rmin;
  const float span = rs2 - rc2;
  if (!(span > 0.0f)) {
    *reason = ORC_RESCALE_DEGENERATE;
    return -1.0f;
  }
  query_t dq, rq;
  dq.m = dm;
  dq.t = dtpl;
  rq.m = rm;
  rq.t = rtpl;
  dq.cx = rq.cx = cx;
  dq.cy = rq.cy = cy;
  dq.cz = rq.cz = cz;
  const float dinv = 1.0f / dm->cell_length, rinv = 1.0f / rm->cell_length;
  dq.ix = (int32_t)(cx * dinv);
  dq.iy = (int32_t)(cy * dinv);
  dq.iz = (int32_t)(cz * dinv);
  rq.ix = (int32_t)(cx * rinv);
  rq.iy = (int32_t)(cy * rinv);
  rq.iz = (int32_t)(cz * rinv);

  double d_within, r_within;
  uint64_t nd_within, nr_within;
  sphere_total(&dq, rs2, &d_within, &nd_within);
  sphere_total(&rq, rs2, &r_within, &nr_within);
  if (!(r_within > 0.0)) {
    *reason = ORC_RESCALE_NO_RANDOMS;
    return -1.0f;
  }
  if ((float)d_within <= K * (float)r_within) {
    *reason = ORC_RESCALE_BEYOND;
    return -1.0f;
  }

  uint32_t nb = 2048;
  {
    const float core_fraction = (rmin * rmin * rmin) / (r_search * r_search * r_search);
    const float want = (float)nd_within * (1.0f - core_fraction) / 16.0f;
    if (!(want >= 64.0f))
      nb = 64;
    else if (want < 2048.0f)
      nb = (uint32_t)want;
  }
  const float inv_w = (float)nb / span;
  const float bin_w = span / (float)nb;
  double* db = malloc(sizeof(double) * nb);
  double* rb = malloc(sizeof(double) * nb);
  double d_core, d_shell, r_core, r_shell;
  histogram(&dq, db, nb, rc2, rs2, inv_w, &d_core, &d_shell);
  histogram(&rq, rb, nb, rc2, rs2, inv_w, &r_core, &r_shell);
  const double d_total = d_core + d_shell;
  const double r_total = r_core + r_shell;
  float radius = -1.0f;
  if (!(r_total > 0.0)) {
    *reason = ORC_RESCALE_NO_RANDOMS;
  } else if ((float)d_total <= K * (float)r_total) {
    *reason = ORC_RESCALE_BEYOND;
  } else {
    /* the data tracers at or below the top of a bin: those enclosed, the
     * core's and the annulus' (every tracer the data histogram binned) */
    uint64_t* nbin = calloc(nb, sizeof(uint64_t));
    uint64_t n_all = 0;
    {
      const orc_template_t* t = dq.t;
      for (uint32_t i = 0; i < t->count; i++) {
        uint64_t s, e;
        float ex, ey, ez;
        if (!query_cell(&dq, i, &s, &e, &ex, &ey, &ez))
          continue;
        if (t->type[i] == ORC_CELL_CORE) {
          n_all += e - s;
          continue;
        }
        for (uint64_t p = s; p < e; p++) {
          const float d2 = dist2(dm->x[p], dm->y[p], dm->z[p], ex, ey, ez);
          if (t->type[i] == ORC_CELL_SHELL || d2 <= rs2) {
            n_all++;
            if (t->type[i] == ORC_CELL_SHELL || d2 > rc2)
              nbin[bin_of(d2, rc2, inv_w, nb)]++;
          }
        }
      }
    }
    uint64_t n_above = 0;
    double d_above = 0.0, r_above = 0.0;
    int found = 0;
    for (int32_t b = (int32_t)nb - 1; b >= 0 && !found; b--) {
      const double bd = db[b], br = rb[b];
      if (!(bd > 0.0)) {
        r_above += br;
        n_above += nbin[b];
        continue;
      }
      float hi = rc2 + (float)(b + 1) * bin_w;
      if (b == (int32_t)nb - 1)
        hi = rs2 * (1.0f + 1e-6f);
      const float d_in_min = (float)(d_total - d_above - bd);
      const float r_in_max = (float)(r_total - r_above);
      if (!(r_in_max > 0.0f) || d_in_min > K * r_in_max) {
        d_above += bd;
        r_above += br;
        n_above += nbin[b];
        continue;
      }
      int32_t blo = b - (int32_t)(window - 1);
      if (blo < 0)
        blo = 0;
      double wd = 0.0, wr = 0.0;
      uint64_t wn = 0;
      for (int32_t bb = blo; bb <= b; bb++) {
        wd += db[bb];
        wr += rb[bb];
        wn += nbin[bb];
      }
      const float lo = rc2 + (float)blo * bin_w;
      found = survey_resolve(&dq, (uint32_t)blo, (uint32_t)b, lo, hi, rc2, rs2, inv_w, bin_w,
        nb, d_total - d_above, r_total - r_above, n_all - n_above, rb, K, &radius, n_in);
      if (!found) {
        d_above += wd;
        r_above += wr;
        n_above += wn;
        b = blo;
      }
    }
    free(nbin);
    if (!found)
      *reason = ORC_RESCALE_BELOW_RUNG;
  }
  free(db);
  free(rb);
  return radius;
}
