/*
 * sif-amd: stacked radial profiles of voids over the tracers of a chain
 * mesh, the reference's include/sif/measure/profiles.h (sif v0.5.0) for
 * this build. Struct layouts, option bits, defaults and semantics are the
 * reference's, so a caller of sif_profiles() links against this library
 * unchanged; the binning runs on the GPU.
 */
#ifndef SIF_MEASURE_PROFILES_H
#define SIF_MEASURE_PROFILES_H

#include <math.h>
#include <stdbool.h>
#include <stdint.h>

#include "sif/core/macros.h"
#include "sif/structures/catalogue.h"
#include "sif/structures/chain_mesh.h"

SIF_EXTERN_C_BEGIN

/* One density profile per void (profiles.h:30-46): n_bins values a row,
 * rows in catalogue order; radii in units of each void's own radius. */
typedef struct {
  uint64_t n_voids;
  uint32_t n_bins;
  sif_real ext;       /* outer edge of the profile, in void radii */
  bool differential;  /* a bin holds its own shell, not everything inside it */
  sif_real* r_edges;  /* n_bins + 1 edges, shared by the rows */
  sif_real* profiles; /* n_voids * n_bins, row-major */
} sif_density_profiles_t;

/* The radius a density bin's value is at, in void radii (profiles.h:48-69):
 * the outer edge of a cumulative bin, the middle of a differential one. */
static inline sif_real sif_density_profiles_bin_radius(
  const sif_density_profiles_t* profs, uint32_t bin) {
  return profs->differential
           ? (profs->r_edges[bin] + profs->r_edges[bin + 1]) * (sif_real)0.5
           : profs->r_edges[bin + 1];
}

/* One radial velocity profile per void (profiles.h:71-83): a shell mean. */
typedef struct {
  uint64_t n_voids;
  uint32_t n_bins;
  sif_real ext;
  sif_real* r_edges;
  sif_real* v_rad;    /* n_voids * n_bins, row-major */
} sif_velocity_profiles_t;

/* profiles.h:85-95. NULL is accepted. */
SIF_API void sif_density_profiles_free(sif_density_profiles_t* profs);
SIF_API void sif_velocity_profiles_free(sif_velocity_profiles_t* profs);

/* A void's row (profiles.h:97-117), borrowed. */
static inline const sif_real* sif_density_profiles_get(
  const sif_density_profiles_t* profs, uint64_t void_idx) {
  return &profs->profiles[void_idx * profs->n_bins];
}
static inline const sif_real* sif_velocity_profiles_get(
  const sif_velocity_profiles_t* profs, uint64_t void_idx) {
  return &profs->v_rad[void_idx * profs->n_bins];
}

/* The reach when the caller gives none (profiles.h:119-123). */
#define SIF_PROFILES_DEFAULT_EXT ((sif_real)5.0)

/* sif_profiles (profiles.h:125-188, profiles.c:354-673): bins the mesh's
 * tracers around each void of cat, out to ext void radii (ext not positive:
 * SIF_PROFILES_DEFAULT_EXT) in n_bins bins. Densities are normalised by the
 * mesh's mean density (its total weight over the box) and given as a
 * contrast: cumulative, or per shell with SIF_PROFILES_DIFFERENTIAL; weighted
 * when the mesh carries weights. Velocities are the shell mean of the radial
 * velocity, over tracers or, with SIF_PROFILES_VELOCITY_WEIGHTED on a weighted
 * mesh, over weights; they need a mesh with velocities. SIF_PBC_OPEN treats
 * the box as open. *out_dens / *out_vel: NULL to allocate, an existing set to
 * fill; either pointer may be NULL to skip that output. A void of non-positive
 * radius keeps its row as it was. Errors: SIF_ERR_INVALID (no voids, an empty
 * mesh, nothing asked for, n_bins 0, velocities from a mesh without them, sky
 * coordinates), SIF_ERR_ALLOC; on failure both outputs are NULL. Needs
 * sif_system_init(). */
SIF_API SIF_NODISCARD int sif_profiles(const sif_catalogue_t* cat,
  const sif_chain_mesh_t* mesh, sif_real ext, uint32_t n_bins, sif_option opt,
  sif_density_profiles_t** out_dens, sif_velocity_profiles_t** out_vel);

/* Mesh resolution sif_profiles runs fastest at (profiles.h:190-243): about
 * 36 tracers a cell, at most 256 cells a side, never zero. The profiles do
 * not depend on it. */
#define SIF_PROFILES_MESH_PARTICLES_PER_CELL 36.0
#define SIF_PROFILES_MESH_MAX_CELLS 256u

static inline uint32_t sif_profiles_suggest_mesh_cells(uint64_t n_particles) {
  if (n_particles == 0)
    return 1;
  const double n = cbrt((double)n_particles / SIF_PROFILES_MESH_PARTICLES_PER_CELL);
  if (n >= (double)SIF_PROFILES_MESH_MAX_CELLS)
    return SIF_PROFILES_MESH_MAX_CELLS;
  return n >= 1.0 ? (uint32_t)n : 1u;
}

SIF_EXTERN_C_END

#endif /* SIF_MEASURE_PROFILES_H */
