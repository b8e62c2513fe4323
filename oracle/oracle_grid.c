/*
 * TEST INFRASTRUCTURE ONLY -- oracle restatement of the field, CIC, density
 * contrast, chain mesh, filter/window and FFT stages of the sif v0.5.0 exodus
 * path. See oracle.h for who may use it.
 *
 * Built with -ffp-contract=off: every float expression is evaluated as
 * written, one rounding per operation.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#  include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* src/structures/field_internal.h:58-64 */
float orc_box_fold(float v, float box) {
  if (v >= box)
    v -= box;
  else if (v < 0)
    v += box;
  return v >= box ? 0.0f : v;
}

/* src/structures/field.c:889-935 with field_internal.h:41-52: a coordinate
 * is inside [0, box), or accepted within 4 float ulps of the box length. */
uint64_t orc_check_in_box(const float* x, const float* y, const float* z,
  uint64_t n, float box, uint64_t* n_edge, uint64_t* first_bad) {
  const double L = (double)box;
  const double tol = L * 4.0 * (double)FLT_EPSILON;
  uint64_t bad = 0, edge = 0, first = UINT64_MAX;
  for (uint64_t i = 0; i < n; i++) {
    const float p[3] = {x[i], y[i], z[i]};
    int inside = 1, accepted = 1;
    for (int a = 0; a < 3; a++) {
      if (!(p[a] >= 0.0f && p[a] < box))
        inside = 0;
      const double v = (double)p[a];
      if (!(v >= -tol && v <= L + tol))
        accepted = 0;
    }
    if (inside)
      continue;
    if (accepted) {
      edge++;
      continue;
    }
    bad++;
    if (i < first)
      first = i;
  }
  if (n_edge)
    *n_edge = edge;
  if (first_bad)
    *first_bad = first;
  return bad;
}

/* --- CIC --- */

/* One particle's cell and eight corner weights, as grid.c:585-650 computes
 * them: u = fold(x) * (N / L), base = (uint32)u with N wrapping to 0, t = u -
 * base, weights (x,y) products times (w * z). */
typedef struct {
  uint32_t ix, iy, iz, ix1, iy1, iz1;
  float v[8]; /* order of the eight writes at grid.c:642-649 */
} cic_corner_t;

static void cic_corners(float px, float py, float pz, float pw, uint32_t N,
  float idx, float box, cic_corner_t* c) {
  const float cx = orc_box_fold(px, box) * idx;
  const float cy = orc_box_fold(py, box) * idx;
  const float cz = orc_box_fold(pz, box) * idx;
  const uint32_t rx = (uint32_t)cx, ry = (uint32_t)cy, rz = (uint32_t)cz;
  c->ix = rx >= N ? 0 : rx;
  c->iy = ry >= N ? 0 : ry;
  c->iz = rz >= N ? 0 : rz;
  c->ix1 = c->ix + 1 == N ? 0 : c->ix + 1;
  c->iy1 = c->iy + 1 == N ? 0 : c->iy + 1;
  c->iz1 = c->iz + 1 == N ? 0 : c->iz + 1;
  const float tx = cx - (float)rx, ty = cy - (float)ry, tz = cz - (float)rz;
  const float wx0 = 1.0f - tx, wx1 = tx;
  const float wy0 = 1.0f - ty, wy1 = ty;
  const float wz0 = 1.0f - tz, wz1 = tz;
  const float w00 = wx0 * wy0, w01 = wx0 * wy1;
  const float w10 = wx1 * wy0, w11 = wx1 * wy1;
  const float mz0 = pw * wz0, mz1 = pw * wz1;
  c->v[0] = w00 * mz0;
  c->v[1] = w00 * mz1;
  c->v[2] = w01 * mz0;
  c->v[3] = w01 * mz1;
  c->v[4] = w10 * mz0;
  c->v[5] = w10 * mz1;
  c->v[6] = w11 * mz0;
  c->v[7] = w11 * mz1;
}

/* grid.c:1042-1175 -> grid_compute_cic (679-746) -> grid_cic_deposit_tile
 * (455-652) for CIC: 16-plane x slabs (:59), 256 chunks per tile (:38), the
 * slab-major counting sort of passes 1-3, the per-slab deposit of pass 4, the
 * serial ghost fold (270-309) and the merge x 1/l^3 (323-342). */
void orc_cic_float(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t N, float box, uint64_t tile,
  float* out) {
  const float idx = (float)N / box;
  const float cell = box / (float)N;
  int32_t n_slabs = (int32_t)(N / 16u);
  if (n_slabs < 1)
    n_slabs = 1;
  const uint32_t base_width = N / (uint32_t)n_slabs;
  const uint64_t plane = (uint64_t)N * N;

  uint32_t* x0 = malloc(sizeof(uint32_t) * n_slabs);
  uint32_t* width = malloc(sizeof(uint32_t) * n_slabs);
  float** vals = malloc(sizeof(float*) * n_slabs);
  for (int32_t t = 0; t < n_slabs; t++) {
    x0[t] = (uint32_t)t * (N / (uint32_t)n_slabs);
    const uint32_t end =
      t == n_slabs - 1 ? N : (uint32_t)(t + 1) * (N / (uint32_t)n_slabs);
    width[t] = end - x0[t];
    vals[t] = calloc((size_t)(width[t] + 1) * plane, sizeof(float));
  }

  if (tile == 0)
    tile = 67108864ull;
  if (tile > n)
    tile = n;
  if (tile == 0)
    tile = 1;
  const int n_chunks = 256;
  uint64_t* counts = malloc(sizeof(uint64_t) * n_chunks * n_slabs);
  uint64_t* sorted = malloc(sizeof(uint64_t) * (tile ? tile : 1));

  for (uint64_t lo = 0; lo < n; lo += tile) {
    const uint64_t hi = lo + tile > n ? n : lo + tile;
    const uint64_t tn = hi - lo;
    const uint64_t chunk = (tn + n_chunks - 1) / n_chunks;
    memset(counts, 0, sizeof(uint64_t) * n_chunks * n_slabs);
    for (int c = 0; c < n_chunks; c++) {
      const uint64_t s = lo + (uint64_t)c * chunk;
      const uint64_t e = s + chunk > hi ? hi : s + chunk;
      for (uint64_t p = s; p < e; p++) {
        uint32_t ix = (uint32_t)(orc_box_fold(x[p], box) * idx);
        ix = ix >= N ? ix - N : ix;
        uint32_t sl = ix / base_width;
        if (sl > (uint32_t)(n_slabs - 1))
          sl = (uint32_t)(n_slabs - 1);
        counts[(uint64_t)c * n_slabs + sl]++;
      }
    }
    uint64_t off = 0;
    for (int32_t sl = 0; sl < n_slabs; sl++)
      for (int c = 0; c < n_chunks; c++) {
        const uint64_t k = (uint64_t)c * n_slabs + sl;
        const uint64_t cnt = counts[k];
        counts[k] = off;
        off += cnt;
      }
    for (int c = 0; c < n_chunks; c++) {
      const uint64_t s = lo + (uint64_t)c * chunk;
      const uint64_t e = s + chunk > hi ? hi : s + chunk;
      for (uint64_t p = s; p < e; p++) {
        uint32_t ix = (uint32_t)(orc_box_fold(x[p], box) * idx);
        ix = ix >= N ? ix - N : ix;
        uint32_t sl = ix / base_width;
        if (sl > (uint32_t)(n_slabs - 1))
          sl = (uint32_t)(n_slabs - 1);
        sorted[counts[(uint64_t)c * n_slabs + sl]++] = p;
      }
    }
#pragma omp parallel for schedule(static, 1)
    for (int32_t sl = 0; sl < n_slabs; sl++) {
      const uint64_t s =
        sl == 0 ? 0 : counts[(uint64_t)(n_chunks - 1) * n_slabs + (sl - 1)];
      const uint64_t e = counts[(uint64_t)(n_chunks - 1) * n_slabs + sl];
      float* v = vals[sl];
      for (uint64_t q = s; q < e; q++) {
        const uint64_t p = sorted[q];
        cic_corner_t c;
        cic_corners(x[p], y[p], z[p], w ? w[p] : 1.0f, N, idx, box, &c);
        const uint64_t l0 = (uint64_t)(c.ix - x0[sl]) * plane;
        const uint64_t l1 = l0 + plane; /* ghost plane when ix is the last */
        v[l0 + (uint64_t)c.iy * N + c.iz] += c.v[0];
        v[l0 + (uint64_t)c.iy * N + c.iz1] += c.v[1];
        v[l0 + (uint64_t)c.iy1 * N + c.iz] += c.v[2];
        v[l0 + (uint64_t)c.iy1 * N + c.iz1] += c.v[3];
        v[l1 + (uint64_t)c.iy * N + c.iz] += c.v[4];
        v[l1 + (uint64_t)c.iy * N + c.iz1] += c.v[5];
        v[l1 + (uint64_t)c.iy1 * N + c.iz] += c.v[6];
        v[l1 + (uint64_t)c.iy1 * N + c.iz1] += c.v[7];
      }
    }
  }

  /* grid.c:270-309: upper ghosts into the next slab, serially. */
  for (int32_t t = 0; t < n_slabs; t++) {
    float* ghost = vals[t] + (uint64_t)width[t] * plane;
    float* next = vals[(t + 1) % n_slabs];
    for (uint64_t i = 0; i < plane; i++)
      next[i] += ghost[i];
  }
  /* grid.c:323-342 and :735-738 */
  const float inv_vol = 1.0f / (float)(cell * cell * cell);
  for (int32_t t = 0; t < n_slabs; t++)
    for (uint32_t l = 0; l < width[t]; l++)
      for (uint64_t i = 0; i < plane; i++)
        out[(uint64_t)(x0[t] + l) * plane + i] =
          vals[t][(uint64_t)l * plane + i] * inv_vol;

  for (int32_t t = 0; t < n_slabs; t++)
    free(vals[t]);
  free(vals);
  free(width);
  free(x0);
  free(counts);
  free(sorted);
}

/* The GPU deposit's arithmetic: the same eight float corner values as
 * grid.c:585-650, each rounded to a multiple of 2^-32 and added in int64. */
/* The fixed-point unit of a deposit, 2^-s: n particles whose largest |w| is
 * wmax put at most n wmax in a cell; with n wmax <= 2^e, s = 62 - e keeps
 * every corner and every cell sum below 2^62. 32 when nothing is deposited. */
int orc_cic_unit_exponent(uint64_t n, float wmax) {
  const double bound = (double)n * (double)wmax;
  if (!(bound > 0.0))
    return 32;
  int e;
  (void)frexp(bound, &e);
  return 62 - e;
}

/* The deposit in fixed point: each corner value rounded once to 2^-s units
 * (s by orc_cic_unit_exponent over the weights' largest magnitude) and summed
 * exactly in int64. Returns s. */
int orc_cic_fixed(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t N, float box, int64_t* acc) {
  const float idx = (float)N / box;
  const uint64_t plane = (uint64_t)N * N;
  memset(acc, 0, sizeof(int64_t) * plane * N);
  float wmax = 1.0f;
  if (w) {
    wmax = 0.0f;
    for (uint64_t p = 0; p < n; p++)
      if (fabsf(w[p]) > wmax)
        wmax = fabsf(w[p]);
  }
  const int s = orc_cic_unit_exponent(n, wmax);
  const double unit_inv = ldexp(1.0, s);
  for (uint64_t p = 0; p < n; p++) {
    cic_corner_t c;
    cic_corners(x[p], y[p], z[p], w ? w[p] : 1.0f, N, idx, box, &c);
    const uint64_t a0 = (uint64_t)c.ix * plane, a1 = (uint64_t)c.ix1 * plane;
    const uint64_t cells[8] = {a0 + (uint64_t)c.iy * N + c.iz,
      a0 + (uint64_t)c.iy * N + c.iz1, a0 + (uint64_t)c.iy1 * N + c.iz,
      a0 + (uint64_t)c.iy1 * N + c.iz1, a1 + (uint64_t)c.iy * N + c.iz,
      a1 + (uint64_t)c.iy * N + c.iz1, a1 + (uint64_t)c.iy1 * N + c.iz,
      a1 + (uint64_t)c.iy1 * N + c.iz1};
    for (int k = 0; k < 8; k++)
      acc[cells[k]] += (int64_t)llrint((double)c.v[k] * unit_inv);
  }
  return s;
}

/* Density per unit volume from the fixed-point cells, and the mean of the
 * fixed-point mass per unit volume (exact integer sum, one rounding). */
void orc_cic_fixed_density(const int64_t* acc, uint32_t N, float box, int s,
  float* out, double* mean_out) {
  const double unit = ldexp(1.0, -s);
  const float cell = box / (float)N;
  const float inv_vol = 1.0f / (float)(cell * cell * cell);
  const uint64_t total = (uint64_t)N * N * N;
  __int128 sum = 0;
  for (uint64_t i = 0; i < total; i++) {
    out[i] = (float)((double)acc[i] * unit) * inv_vol;
    sum += acc[i];
  }
  if (mean_out)
    *mean_out = (double)sum * unit * (double)inv_vol / (double)total;
}

/* Textbook CIC in double: each particle spreads unit mass (or its weight)
 * over the 8 nodes around it with trilinear weights, periodic. */
void orc_cic_brute_double(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t N, float box, double* out) {
  const double L = (double)box, h = L / (double)N;
  const uint64_t total = (uint64_t)N * N * N;
  memset(out, 0, sizeof(double) * total);
  for (uint64_t p = 0; p < n; p++) {
    const double u[3] = {
      (double)orc_box_fold(x[p], box) / h, (double)orc_box_fold(y[p], box) / h,
      (double)orc_box_fold(z[p], box) / h};
    int64_t b[3];
    double f[3];
    for (int a = 0; a < 3; a++) {
      b[a] = (int64_t)floor(u[a]);
      f[a] = u[a] - (double)b[a];
    }
    const double m = w ? (double)w[p] : 1.0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++) {
          const uint64_t gx = (uint64_t)(((b[0] + i) % N + N) % N);
          const uint64_t gy = (uint64_t)(((b[1] + j) % N + N) % N);
          const uint64_t gz = (uint64_t)(((b[2] + k) % N + N) % N);
          const double wt = (i ? f[0] : 1.0 - f[0]) * (j ? f[1] : 1.0 - f[1]) *
                            (k ? f[2] : 1.0 - f[2]);
          out[(gx * N + gy) * N + gz] += m * wt;
        }
  }
  const double inv = 1.0 / (h * h * h);
  for (uint64_t i = 0; i < total; i++)
    out[i] *= inv;
}

/* grid.c:1177-1244: mean over 64 fixed blocks summed serially in double,
 * delta = v * (float)(1 / mean) - 1. */
int orc_density_contrast(float* v, uint64_t total, double* mean_out) {
  const uint64_t block = (total + 63) / 64;
  double mass = 0.0;
  for (int b = 0; b < 64; b++) {
    const uint64_t s = (uint64_t)b * block;
    const uint64_t e = s + block < total ? s + block : total;
    double sum = 0.0;
    for (uint64_t i = s; i < e; i++)
      sum += (double)v[i];
    mass += sum;
  }
  const double mean = mass / (double)total;
  if (mean_out)
    *mean_out = mean;
  if (!(mean > 0.0))
    return -1;
  orc_density_contrast_with_mean(v, total, mean);
  return 0;
}

void orc_density_contrast_with_mean(float* v, uint64_t total, double mean) {
  const float inv = (float)(1.0 / mean);
  for (uint64_t i = 0; i < total; i++)
    v[i] = v[i] * inv - 1.0f;
}

/* --- chain mesh --- */

/* chain_mesh.c:92, :687, :316-331 */
uint64_t orc_chain_mesh_cell(float x, float y, float z, uint32_t nc, float box) {
  const float cell = box / (float)nc;
  const float inv_l = 1.0f / cell;
  int32_t ix = (int32_t)(orc_box_fold(x, box) * inv_l);
  int32_t iy = (int32_t)(orc_box_fold(y, box) * inv_l);
  int32_t iz = (int32_t)(orc_box_fold(z, box) * inv_l);
  const int32_t n = (int32_t)nc;
  ix = ix < 0 ? 0 : (ix >= n ? n - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= n ? n - 1 : iy);
  iz = iz < 0 ? 0 : (iz >= n ? n - 1 : iz);
  return (uint64_t)ix * nc * nc + (uint64_t)iy * nc + (uint64_t)iz;
}

/* chain_mesh.c:680-832 followed by the canonical order of :364-542: counts,
 * exclusive prefix, and every cell's particles in ascending source index --
 * which is exactly a stable counting sort by cell. Cell weights and the
 * total: :624-672 (64 ordered blocks of cells, double). */
void orc_chain_mesh(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t nc, float box, uint64_t* offsets,
  float* mx, float* my, float* mz, float* mw, uint64_t* orig, float* cw,
  double* total_weight) {
  const uint64_t cells = (uint64_t)nc * nc * nc;
  uint64_t* cell_of = malloc(sizeof(uint64_t) * (n ? n : 1));
  memset(offsets, 0, sizeof(uint64_t) * (cells + 1));
  for (uint64_t i = 0; i < n; i++) {
    cell_of[i] = orc_chain_mesh_cell(x[i], y[i], z[i], nc, box);
    offsets[cell_of[i]]++;
  }
  uint64_t run = 0;
  for (uint64_t c = 0; c < cells; c++) {
    const uint64_t k = offsets[c];
    offsets[c] = run;
    run += k;
  }
  offsets[cells] = run;
  uint64_t* cursor = malloc(sizeof(uint64_t) * cells);
  memcpy(cursor, offsets, sizeof(uint64_t) * cells);
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t pos = cursor[cell_of[i]]++;
    if (orig)
      orig[pos] = i;
    mx[pos] = orc_box_fold(x[i], box);
    my[pos] = orc_box_fold(y[i], box);
    mz[pos] = orc_box_fold(z[i], box);
    if (w && mw)
      mw[pos] = w[i];
  }
  free(cursor);
  free(cell_of);
  if (!w) {
    if (total_weight)
      *total_weight = (double)n;
    return;
  }
  const uint64_t block = (cells + 63) / 64;
  double total = 0.0;
  for (int b = 0; b < 64; b++) {
    const uint64_t s = (uint64_t)b * block;
    const uint64_t e = s + block < cells ? s + block : cells;
    double sum = 0.0;
    for (uint64_t c = s; c < e; c++) {
      double cell = 0.0;
      for (uint64_t p = offsets[c]; p < offsets[c + 1]; p++)
        cell += (double)mw[p];
      if (cw)
        cw[c] = (float)cell;
      sum += cell;
    }
    total += sum;
  }
  if (total_weight)
    *total_weight = total;
}

/* --- filter and window --- */

uint32_t orc_filter_lut_size(uint32_t n) {
  const uint32_t h = n >> 1;
  return 3 * h * h + 1;
}

/* src/math/fft.c:521-568: indexed by the integer |k|^2, evaluated in float. */
void orc_filter_lut(int filter, float r, float sigma, float box, uint32_t n,
  float* lut) {
  const uint32_t max_k2 = orc_filter_lut_size(n) - 1;
  const float pi = (float)3.14159265358979323846;
  const float factor = (float)(r * 2.0 * pi / box);
  const float factor2 = factor * factor;
  const float gfactor = (float)(sigma * 2.0 * pi / box);
  const float gfactor2 = gfactor * gfactor;
  for (uint32_t k2 = 0; k2 <= max_k2; k2++) {
    if (k2 == 0) {
      lut[k2] = 1.0f;
    } else if (filter == ORC_FILTER_TOP_HAT) {
      const float kr = factor * sqrtf((float)k2);
      if (kr > 1e-4f)
        lut[k2] = 3.0f * (sinf(kr) - kr * cosf(kr)) / (kr * kr * kr);
      else
        lut[k2] = 1.0f;
    } else if (filter == ORC_FILTER_GAUSSIAN) {
      lut[k2] = expf(-0.5f * factor2 * (float)k2);
    } else {
      lut[k2] = 1.0f;
    }
    if (sigma > 0.0f && k2 > 0)
      lut[k2] *= expf(-0.5f * gfactor2 * (float)k2);
  }
}

/* fft.c:890-975: x = pi k / n with the float pi, W = sinc^order by repeated
 * multiplication, S the image sum of the aliasing-aware mode. */
static double win_x(uint32_t i, uint32_t n) {
  const uint32_t h = n >> 1;
  const int32_t k = i > h ? (int32_t)i - (int32_t)n : (int32_t)i;
  return (double)(float)3.14159265358979323846 * (double)k / (double)n;
}
static double win_w(double x, int order) {
  const double s = (x < 1e-9 && x > -1e-9) ? 1.0 : sin(x) / x;
  double w = 1.0;
  for (int j = 0; j < order; j++)
    w *= s;
  return w;
}
static double win_s(double x, int order) {
  const double s2 = sin(x) * sin(x);
  switch (order) {
  case 1:
    return 1.0;
  case 2:
    return 1.0 - 2.0 * s2 / 3.0;
  case 3:
    return 1.0 - s2 + 2.0 * s2 * s2 / 15.0;
  default:
    return 1.0 - 4.0 * s2 / 3.0 + 2.0 * s2 * s2 / 5.0 -
           4.0 * s2 * s2 * s2 / 315.0;
  }
}
static double win_a(double x, int order) {
  const double c = cos(x), c2 = c * c;
  switch (order) {
  case 1:
    return c;
  case 2:
    return c * (c2 + 5.0) / 6.0;
  case 3:
    return c * (c2 * c2 + 58.0 * c2 + 61.0) / 120.0;
  default:
    return c * (c2 * c2 * c2 + 543.0 * c2 * c2 + 3111.0 * c2 + 1385.0) /
           5040.0;
  }
}

void orc_window_axis(uint32_t n, int order, int mode, double* axis) {
  for (uint32_t i = 0; i < n; i++) {
    if (order == 0) {
      axis[i] = 1.0;
      continue;
    }
    const double x = win_x(i, n), w = win_w(x, order);
    axis[i] = mode ? w / win_s(x, order) : 1.0 / w;
  }
}

/* fft.c:1001-1010 */
double orc_window_corner_gain(int order, int mode, int interlaced) {
  const double x = 0.5 * (double)(float)3.14159265358979323846;
  const double w = win_w(x, order);
  if (mode == 0)
    return 1.0 / (w * w * w);
  const double s = win_s(x, order), a = win_a(x, order);
  const double images = interlaced ? 0.5 * (s * s * s + a * a * a) : s * s * s;
  return w * w * w / images;
}

/* fft.c:753-780 */
float orc_mode_factor(const float* lut, const double* axis, uint32_t n,
  uint32_t ix, uint32_t iy, uint32_t iz) {
  const uint32_t h = n >> 1;
  const int32_t kx = ix > h ? (int32_t)ix - (int32_t)n : (int32_t)ix;
  const int32_t ky = iy > h ? (int32_t)iy - (int32_t)n : (int32_t)iy;
  const uint32_t k2 = (uint32_t)(kx * kx) + (uint32_t)(ky * ky) + iz * iz;
  float s = lut[k2];
  if (axis) {
    const double cxy = axis[ix] * axis[iy];
    s = (float)((double)s * cxy * axis[iz]);
  }
  return s;
}

/* --- FFT in double (power-of-two lengths), FFTW sign convention --- */

static void fft1d(double* re, double* im, uint32_t n, int sign) {
  for (uint32_t i = 1, j = 0; i < n; i++) {
    uint32_t bit = n >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i];
      re[i] = re[j];
      re[j] = t;
      t = im[i];
      im[i] = im[j];
      im[j] = t;
    }
  }
  for (uint32_t len = 2; len <= n; len <<= 1) {
    const double ang = sign * 2.0 * 3.14159265358979323846 / (double)len;
    for (uint32_t i = 0; i < n; i += len)
      for (uint32_t k = 0; k < len / 2; k++) {
        const double wr = cos(ang * k), wi = sin(ang * k);
        const uint32_t a = i + k, b = i + k + len / 2;
        const double xr = re[b] * wr - im[b] * wi;
        const double xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
  }
}

static int pow2(uint32_t n) { return n >= 2 && (n & (n - 1)) == 0; }

/* 1D c2c along axis `ax` (0 = x, 1 = y) of the half spectrum. */
static void spec_axis(double* spec, uint32_t n, int ax, int sign) {
  const uint32_t nz = n / 2 + 1;
#pragma omp parallel
  {
    double* re = malloc(sizeof(double) * n);
    double* im = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (uint64_t q = 0; q < (uint64_t)n * nz; q++) {
      const uint32_t a = (uint32_t)(q / nz), kz = (uint32_t)(q % nz);
      for (uint32_t i = 0; i < n; i++) {
        const uint64_t m = ax == 0 ? ((uint64_t)i * n + a) * nz + kz
                                   : ((uint64_t)a * n + i) * nz + kz;
        re[i] = spec[2 * m];
        im[i] = spec[2 * m + 1];
      }
      fft1d(re, im, n, sign);
      for (uint32_t i = 0; i < n; i++) {
        const uint64_t m = ax == 0 ? ((uint64_t)i * n + a) * nz + kz
                                   : ((uint64_t)a * n + i) * nz + kz;
        spec[2 * m] = re[i];
        spec[2 * m + 1] = im[i];
      }
    }
    free(re);
    free(im);
  }
}

int orc_fft_r2c(const double* in, uint32_t n, double* spec) {
  if (!pow2(n))
    return -1;
  const uint32_t nz = n / 2 + 1;
#pragma omp parallel
  {
    double* re = malloc(sizeof(double) * n);
    double* im = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (uint64_t row = 0; row < (uint64_t)n * n; row++) {
      for (uint32_t i = 0; i < n; i++) {
        re[i] = in[row * n + i];
        im[i] = 0.0;
      }
      fft1d(re, im, n, -1);
      for (uint32_t k = 0; k < nz; k++) {
        spec[2 * (row * nz + k)] = re[k];
        spec[2 * (row * nz + k) + 1] = im[k];
      }
    }
    free(re);
    free(im);
  }
  spec_axis(spec, n, 1, -1);
  spec_axis(spec, n, 0, -1);
  return 0;
}

/* The inverse of orc_fft_r2c, unnormalised: c2c along x and y, then each z
 * row extended by Hermitian symmetry and transformed, keeping the real part
 * (the imaginary parts of the kz = 0 and n/2 modes drop out, as in FFTW's
 * c2r). The input is not modified. */
int orc_fft_c2r(const double* spec_in, uint32_t n, double* out) {
  if (!pow2(n))
    return -1;
  const uint32_t nz = n / 2 + 1;
  const uint64_t modes = (uint64_t)n * n * nz;
  double* spec = malloc(sizeof(double) * 2 * modes);
  memcpy(spec, spec_in, sizeof(double) * 2 * modes);
  spec_axis(spec, n, 0, +1);
  spec_axis(spec, n, 1, +1);
#pragma omp parallel
  {
    double* re = malloc(sizeof(double) * n);
    double* im = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (uint64_t row = 0; row < (uint64_t)n * n; row++) {
      for (uint32_t k = 0; k < nz; k++) {
        re[k] = spec[2 * (row * nz + k)];
        im[k] = spec[2 * (row * nz + k) + 1];
      }
      for (uint32_t k = nz; k < n; k++) {
        re[k] = re[n - k];
        im[k] = -im[n - k];
      }
      fft1d(re, im, n, +1);
      for (uint32_t i = 0; i < n; i++)
        out[row * n + i] = re[i];
    }
    free(re);
    free(im);
  }
  free(spec);
  return 0;
}
