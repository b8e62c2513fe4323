/*
 * TEST INFRASTRUCTURE ONLY -- oracle restatement of sif_profiles
 * (src/measure/profiles.c:144-350, 354-673, sif v0.5.0): per void, the cells
 * of the sphere's bounding box walked x, y, z with the slab, row and cell gap
 * tests, each cell's tracers binned in mesh order into float accumulators,
 * then the profiles bin by bin. See oracle.h for who may use it. Built with
 * -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_F ((float)3.14159265358979323846)

typedef struct {
  const orc_mesh_t* m;
  float box, half_box, inv_box, cell, half_cell;
  uint32_t n;
  int periodic;
} walk_t;

/* profiles.c:302-330 */
static void axis_range(float centre, float r_max, float inv_cell, uint32_t n,
  int periodic, int32_t* lo_out, int32_t* hi_out) {
  double lo = floor(((double)centre - (double)r_max) * (double)inv_cell);
  double hi = floor(((double)centre + (double)r_max) * (double)inv_cell);
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

/* profiles.c:261-300 */
static uint32_t cell_image(const walk_t* W, int32_t idx, float centre,
  float* image_out, float* gap_out) {
  int32_t w = idx;
  if (W->periodic) {
    if (w < 0)
      w += (int32_t)W->n;
    else if (w >= (int32_t)W->n)
      w -= (int32_t)W->n;
  }
  float gap = ((float)w + 0.5f) * W->cell - centre;
  float image = centre;
  if (W->periodic) {
    const float k = roundf(gap * W->inv_box);
    gap -= k * W->box;
    image += k * W->box;
  }
  gap = fabsf(gap) - W->half_cell;
  *image_out = image;
  *gap_out = gap > 0.0f ? gap : 0.0f;
  return (uint32_t)w;
}

int orc_profiles(const orc_mesh_t* m, const float* cx, const float* cy,
  const float* cz, const float* rv, uint64_t n_voids, float ext,
  uint32_t n_bins, uint32_t opt, float* dens, float* vel, uint64_t* count) {
  if (!m || n_bins == 0 || (!dens && !vel) || (vel && !m->vx))
    return -1;
  if (!(ext > 0.0f))
    ext = 5.0f;
  const int differential = (opt >> 8) & 1u;
  const int have_w = m->w != NULL;
  const int vel_w = vel && have_w && ((opt >> 9) & 1u);
  walk_t W;
  W.m = m;
  W.box = m->box_length;
  W.half_box = W.box * 0.5f;
  W.inv_box = 1.0f / W.box;
  W.cell = m->cell_length;
  W.half_cell = W.cell * 0.5f;
  W.n = m->n_cells;
  W.periodic = (opt & 1u) == 0;
  const float inv_cell = 1.0f / m->cell_length;
  const float bin_width = ext / (float)n_bins;
  float* vols = malloc(sizeof(float) * n_bins);
  for (uint32_t j = 0; j < n_bins; j++) {
    const float outer = (float)(j + 1) * bin_width;
    const float inner = differential ? (float)j * bin_width : 0.0f;
    vols[j] = (float)(4.0 / 3.0) * PI_F *
              (outer * outer * outer - inner * inner * inner);
  }
  const double box_vol = (double)W.box * (double)W.box * (double)W.box;
  const float mean_dens = (float)(m->total_weight / box_vol);

#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t i = 0; i < n_voids; i++) {
    const float r_v = rv[i];
    if (!(r_v > 0.0f))
      continue;
    float* mass = calloc(n_bins, sizeof(float));
    float* vrad = calloc(n_bins, sizeof(float));
    uint64_t* cnt = calloc(n_bins, sizeof(uint64_t));
    const float c3[3] = {cx[i], cy[i], cz[i]};
    const float r_max = r_v * ext;
    const float r_max_sq = r_max * r_max;
    const float inv_bin = 1.0f / (r_v * bin_width);
    const int per_cell = !W.periodic || (r_max + m->cell_length <= W.half_box);
    int32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++)
      axis_range(c3[a], r_max, inv_cell, W.n, W.periodic, &lo[a], &hi[a]);
    for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
      float ex, gx;
      const uint32_t wx = cell_image(&W, ix, c3[0], &ex, &gx);
      const float gx2 = gx * gx;
      if (gx2 > r_max_sq)
        continue;
      for (int32_t iy = lo[1]; iy <= hi[1]; iy++) {
        float ey, gy;
        const uint32_t wy = cell_image(&W, iy, c3[1], &ey, &gy);
        const float gxy2 = gx2 + gy * gy;
        if (gxy2 > r_max_sq)
          continue;
        for (int32_t iz = lo[2]; iz <= hi[2]; iz++) {
          float ez, gz;
          const uint32_t wz = cell_image(&W, iz, c3[2], &ez, &gz);
          if (gxy2 + gz * gz > r_max_sq)
            continue;
          const uint64_t flat = (uint64_t)wx * W.n * W.n + (uint64_t)wy * W.n + wz;
          const uint64_t s = m->offsets[flat], e = m->offsets[flat + 1];
          const float ox = per_cell ? ex : c3[0];
          const float oy = per_cell ? ey : c3[1];
          const float oz = per_cell ? ez : c3[2];
          for (uint64_t p = s; p < e; p++) {
            float dx = m->x[p] - ox, dy = m->y[p] - oy, dz = m->z[p] - oz;
            if (!per_cell) { /* profiles.c:211-219 */
              dx -= W.box * (float)(dx > W.half_box) - W.box * (float)(dx < -W.half_box);
              dy -= W.box * (float)(dy > W.half_box) - W.box * (float)(dy < -W.half_box);
              dz -= W.box * (float)(dz > W.half_box) - W.box * (float)(dz < -W.half_box);
            }
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 <= r_max_sq))
              continue;
            const float r = sqrtf(r2);
            const uint32_t b = (uint32_t)(r * inv_bin);
            if (b >= n_bins)
              continue;
            cnt[b]++;
            if (dens || vel_w)
              mass[b] += have_w ? m->w[p] : 1.0f;
            if (vel) {
              const float inv_r = r > 0.0f ? 1.0f / r : 0.0f;
              const float v_r = (m->vx[p] * dx + m->vy[p] * dy + m->vz[p] * dz) * inv_r;
              vrad[b] += vel_w ? m->w[p] * v_r : v_r;
            }
          }
        }
      }
    }
    const float rv3 = r_v * r_v * r_v;
    float cumulative = 0.0f;
    for (uint32_t j = 0; j < n_bins; j++) {
      const uint64_t o = i * n_bins + j;
      if (dens) {
        cumulative += mass[j];
        const float weight = differential ? mass[j] : cumulative;
        const float raw = weight / (vols[j] * rv3);
        dens[o] = raw / mean_dens - 1.0f;
      }
      if (vel) {
        if (vel_w)
          vel[o] = mass[j] > 0.0f ? vrad[j] / mass[j] : 0.0f;
        else
          vel[o] = cnt[j] > 0 ? vrad[j] / (float)cnt[j] : 0.0f;
      }
      if (count)
        count[o] = cnt[j];
    }
    free(mass);
    free(vrad);
    free(cnt);
  }
  free(vols);
  return 0;
}
