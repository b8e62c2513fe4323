/*
 * TEST INFRASTRUCTURE ONLY -- the whole exodus CPU path of sif v0.5.0 on the
 * host cores: the timed CPU baseline and the end-to-end check. See oracle.h.
 *
 * Stages as bin/sif-exodus/pipeline.c:540-700 runs them: CIC (grid.c order),
 * density contrast, chain mesh, forward FFT with the CIC window divided out
 * (exodus_finder.c:515-553), then per rung the top-hat filter and c2r
 * (fft.c:703-888) followed by the finder. The FFT is this oracle's own in
 * double; each smoothed field is rounded to float as the reference stores it.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_desc(const void* pa, const void* pb) {
  const float a = *(const float*)pa, b = *(const float*)pb;
  return (a < b) - (a > b);
}

/* The float path's per-rung smoothing, made when the finder reaches a rung
 * (the reference smooths each rung into the one grid buffer the same way,
 * exodus_finder.c:1704-1760). */
typedef struct {
  const float* spec;
  float* work;
  float* field;
  float* lut;
  const double* axis;
  const float* radii;
  uint32_t n;
  float box;
  double seconds;
} lazy_smooth_t;

static const float* lazy_field(void* ctx, uint32_t rung) {
  lazy_smooth_t* L = ctx;
  const double t0 = now_s();
  orc_filter_lut(ORC_FILTER_TOP_HAT, L->radii[rung], 0.0f, L->box, L->n, L->lut);
  const float norm = 1.0f / (float)((uint64_t)L->n * L->n * L->n);
  const int st = orc_fftf_filter_c2r(L->spec, L->n, L->lut, L->axis, norm, L->work, L->field);
  L->seconds += now_s() - t0;
  return st == 0 ? L->field : NULL;
}

int orc_exodus_pipeline(const float* x, const float* y, const float* z,
  uint64_t n, uint32_t N, float box, uint32_t mesh_cells, const float* radii_in,
  uint32_t n_radii, float threshold, uint32_t max_rungs, int fp32, orc_catalogue_t* cat,
  double* stage_s) {
  const uint64_t total = (uint64_t)N * N * N;
  const uint32_t nz = N / 2 + 1;
  const uint64_t modes = (uint64_t)N * N * nz;
  double t0 = now_s(), t1;
  for (int i = 0; i < 6; i++)
    stage_s[i] = 0.0;

  float* grid = malloc(sizeof(float) * total);
  orc_cic_float(x, y, z, NULL, n, N, box, 0, grid);
  t1 = now_s();
  stage_s[0] = t1 - t0;
  t0 = t1;

  double mean;
  if (orc_density_contrast(grid, total, &mean) != 0) {
    free(grid);
    return -3;
  }
  t1 = now_s();
  stage_s[1] = t1 - t0;
  t0 = t1;

  uint64_t* offsets = malloc(sizeof(uint64_t) * ((uint64_t)mesh_cells *
                                                   mesh_cells * mesh_cells + 1));
  float* mx = malloc(sizeof(float) * n);
  float* my = malloc(sizeof(float) * n);
  float* mz = malloc(sizeof(float) * n);
  double tw;
  orc_chain_mesh(x, y, z, NULL, n, mesh_cells, box, offsets, mx, my, mz, NULL,
    NULL, NULL, &tw);
  t1 = now_s();
  stage_s[2] = t1 - t0;
  t0 = t1;

  float* radii = malloc(sizeof(float) * n_radii);
  memcpy(radii, radii_in, sizeof(float) * n_radii);
  qsort(radii, n_radii, sizeof(float), cmp_desc);
  if (max_rungs && max_rungs < n_radii)
    n_radii = max_rungs;
  double* axis = malloc(sizeof(double) * N);
  orc_window_axis(N, 2, 0, axis);
  float* lut = malloc(sizeof(float) * orc_filter_lut_size(N));
  const orc_mesh_t mesh = {mesh_cells, box, box / (float)mesh_cells, offsets, mx,
    my, mz, NULL, NULL, (double)n, NULL, NULL, NULL};
  int st = 0;

  if (fp32) {
    float* fspec = malloc(sizeof(float) * 2 * modes);
    float* fwork = malloc(sizeof(float) * 2 * modes);
    st = fspec && fwork ? orc_fftf_r2c(grid, N, fspec) : -4;
    t1 = now_s();
    stage_s[3] = t1 - t0;
    t0 = t1;
    if (st == 0) {
      lazy_smooth_t L = {fspec, fwork, grid, lut, axis, radii, N, box, 0.0};
      st = orc_exodus_run_lazy(lazy_field, &L, N, box, &mesh, n, radii, n_radii, threshold,
        0.0f, 0, 0, 0, 1, 8, cat, NULL);
      t1 = now_s();
      stage_s[4] = L.seconds;
      stage_s[5] = t1 - t0 - L.seconds;
    } else {
      st = -5;
    }
    free(fspec);
    free(fwork);
  } else {
    double* dgrid = malloc(sizeof(double) * total);
    for (uint64_t i = 0; i < total; i++)
      dgrid[i] = (double)grid[i];
    double* spec = malloc(sizeof(double) * 2 * modes);
    st = orc_fft_r2c(dgrid, N, spec) != 0 ? -5 : 0;
    t1 = now_s();
    stage_s[3] = t1 - t0;
    t0 = t1;
    float** fields = calloc(n_radii, sizeof(float*));
    double* work = malloc(sizeof(double) * 2 * modes);
    const double norm = 1.0 / (double)total;
    for (uint32_t r = 0; r < n_radii && st == 0; r++) {
      orc_filter_lut(ORC_FILTER_TOP_HAT, radii[r], 0.0f, box, N, lut);
#pragma omp parallel for schedule(static)
      for (uint64_t m = 0; m < modes; m++) {
        const uint32_t ix = (uint32_t)(m / ((uint64_t)N * nz));
        const uint32_t iy = (uint32_t)((m / nz) % N);
        const uint32_t iz = (uint32_t)(m % nz);
        const double s = (double)orc_mode_factor(lut, axis, N, ix, iy, iz);
        work[2 * m] = spec[2 * m] * s;
        work[2 * m + 1] = spec[2 * m + 1] * s;
      }
      orc_fft_c2r(work, N, dgrid);
      fields[r] = malloc(sizeof(float) * total);
      for (uint64_t i = 0; i < total; i++)
        fields[r][i] = (float)(dgrid[i] * norm);
    }
    t1 = now_s();
    stage_s[4] = t1 - t0;
    t0 = t1;
    if (st == 0)
      st = orc_exodus_run(fields, N, box, &mesh, n, radii, n_radii, threshold, 0.0f, 0, 0, 0,
        1, 8, cat, NULL);
    t1 = now_s();
    stage_s[5] = t1 - t0;
    for (uint32_t r = 0; r < n_radii; r++)
      free(fields[r]);
    free(fields);
    free(work);
    free(spec);
    free(dgrid);
  }

  free(lut);
  free(axis);
  free(radii);
  free(grid);
  free(offsets);
  free(mx);
  free(my);
  free(mz);
  return st;
}
