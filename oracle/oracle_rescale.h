/*
 * TEST INFRASTRUCTURE ONLY -- helpers of the oracle's rescaling restatement
 * shared by the box (oracle_finder.c) and survey (oracle_survey.c) forms.
 */
#ifndef SIF_ORACLE_RESCALE_H
#define SIF_ORACLE_RESCALE_H

#include "oracle.h"

typedef struct {
  const orc_mesh_t* m;
  const orc_template_t* t;
  float cx, cy, cz;
  int32_t ix, iy, iz;
} query_t;

/* exodus_rescale.c:295-342 */
static inline int query_cell(const query_t* q, uint32_t i, uint64_t* s, uint64_t* e,
  float* ex, float* ey, float* ez) {
  const int32_t N = (int32_t)q->m->n_cells;
  const float L = q->m->box_length;
  int32_t ix = q->ix + q->t->dx[i];
  float cx = q->cx;
  if (ix < 0) {
    ix += N;
    cx += L;
  } else if (ix >= N) {
    ix -= N;
    cx -= L;
  }
  int32_t iy = q->iy + q->t->dy[i];
  float cy = q->cy;
  if (iy < 0) {
    iy += N;
    cy += L;
  } else if (iy >= N) {
    iy -= N;
    cy -= L;
  }
  int32_t iz = q->iz + q->t->dz[i];
  float cz = q->cz;
  if (iz < 0) {
    iz += N;
    cz += L;
  } else if (iz >= N) {
    iz -= N;
    cz -= L;
  }
  const uint64_t flat = (uint64_t)ix * N * N + (uint64_t)iy * N + iz;
  *s = q->m->offsets[flat];
  *e = q->m->offsets[flat + 1];
  *ex = cx;
  *ey = cy;
  *ez = cz;
  return *e > *s;
}

/* exodus_rescale.c:349-355 */
static inline float dist2(float px, float py, float pz, float cx, float cy,
  float cz) {
  const float dx = px - cx, dy = py - cy, dz = pz - cz;
  return dx * dx + dy * dy + dz * dz;
}

/* exodus_rescale.c:360-369 */
static inline uint32_t bin_of(float d2, float rc2, float inv_w, uint32_t nb) {
  const int32_t b = (int32_t)((d2 - rc2) * inv_w);
  if (b < 0)
    return 0;
  if ((uint32_t)b >= nb)
    return nb - 1;
  return (uint32_t)b;
}

typedef struct {
  float d2, w;
} pair_t;

/* exodus_rescale.c:29 sorts the pairs on d2 alone, so the order of equal
 * distances is its quicksort's. Equal distances give one radius whichever
 * passes first (the walk's enclosed weight only falls through them); here
 * they are ordered by weight too, the order the device sort gives, which
 * fixes the reported enclosed count. */
static inline int cmp_pair(const void* pa, const void* pb) {
  const pair_t* a = pa;
  const pair_t* b = pb;
  if (a->d2 != b->d2)
    return (a->d2 > b->d2) - (a->d2 < b->d2);
  const float wa = a->w + 0.0f, wb = b->w + 0.0f; /* -0 as +0 */
  return (wa > wb) - (wa < wb);
}

#endif
