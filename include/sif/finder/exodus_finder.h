/*
 * The exodus spherical void finder. Entry points, extras and variants of
 * include/sif/finder/exodus_finder.h:40-407, :560-723 (sif v0.5.0).
 *
 * The grid must be a density contrast and the mesh must span the same box.
 * The finder smooths on the GPU (Stockham c2r with the top-hat and the CIC
 * window fused), scans and sorts the candidates on the GPU, grows spheres one
 * wavefront per candidate, and commits in candidate order on the host, which
 * makes the catalogue independent of batching and launch shapes.
 */
#ifndef SIF_FINDER_EXODUS_FINDER_H
#define SIF_FINDER_EXODUS_FINDER_H

#include "sif/core/macros.h"
#include "sif/structures/catalogue.h"
#include "sif/structures/chain_mesh.h"
#include "sif/structures/field.h"
#include "sif/structures/grid.h"

SIF_EXTERN_C_BEGIN

/* exodus_finder.h:40-106 */
typedef struct {
  const char* checkpoint;
  const sif_catalogue_t* prior;
  int keep_checkpoint;
  int save_memory;
} sif_finder_exodus_extras_t;

/* exodus_finder.h:124-130 */
typedef enum {
  SIF_EXODUS_EXCURSION_SET = 0,
  SIF_EXODUS_EXCURSION_PEAK,
  SIF_EXODUS_LAGRANGIAN_OVERLAP,
  SIF__EXODUS_N_VARIANTS
} sif_exodus_variant_t;

/* exodus_finder.h:147 */
SIF_API SIF_NODISCARD int sif_finder_exodus_preset(
  sif_exodus_variant_t variant, sif_real* overlap_fraction, sif_option* opt);
SIF_API const char* sif_finder_exodus_variant_name(sif_exodus_variant_t variant);
SIF_API SIF_NODISCARD int sif_finder_exodus_variant_from_name(
  const char* name, sif_exodus_variant_t* variant);

/* exodus_finder.h:184 */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_finder_exodus(sif_grid_t* grid,
  const sif_chain_mesh_t* mesh, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_exodus_variant_t variant,
  const sif_finder_exodus_extras_t* extras);

/* exodus_finder.h:300. Box runs with SIF_FINDER_SEEDS_CELLS on their nodes
 * (excursion-set, lagrangian-overlap); other seeds and rules return
 * SIF_ERR_UNSUPPORTED in this build. */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_finder_exodus_raw(sif_grid_t* grid,
  const sif_chain_mesh_t* mesh, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_real overlap_fraction, sif_option opt,
  const sif_finder_exodus_extras_t* extras);

/* exodus_finder.h:305: sif_finder_exodus_survey_raw with a variant's
 * settings (excursion-set or lagrangian-overlap). */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_finder_exodus_survey(sif_grid_t* data_grid,
  sif_grid_t* random_grid, const sif_chain_mesh_t* data_mesh,
  const sif_chain_mesh_t* random_mesh, const sif_real offset[3], const sif_real* radii,
  uint32_t n_radii, sif_real threshold, sif_exodus_variant_t variant,
  const sif_finder_exodus_extras_t* extras);

/* exodus_finder.h:402: the exodus finder on a survey whose geometry comes
 * from a random catalogue. data_grid holds the data's CIC densities (NOT a
 * density contrast) and is smoothed and restored like a box run's grid;
 * random_grid the randoms' on the same cells (only read here); the meshes
 * share the box, their cells need not match. Every tracer must sit the
 * largest search sphere plus a grid cell (or half the sphere and a mesh cell,
 * whichever is more) from every face (SIF_ERR_INVALID otherwise).
 * SIF_FINDER_SEEDS_CELLS on their nodes only. The catalogue carries the
 * footprint columns and is moved back by -offset (NULL: not moved). */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_finder_exodus_survey_raw(
  sif_grid_t* data_grid, sif_grid_t* random_grid, const sif_chain_mesh_t* data_mesh,
  const sif_chain_mesh_t* random_mesh, const sif_real offset[3], const sif_real* radii,
  uint32_t n_radii, sif_real threshold, sif_real overlap_fraction, sif_option opt,
  const sif_finder_exodus_extras_t* extras);

/* exodus_finder.h:601: the cubic box a survey must be searched in (the
 * padding sized for grid cells up to the smallest radius) and the offset
 * that centres data and randoms in it. */
SIF_API SIF_NODISCARD int sif_finder_exodus_survey_box(const sif_field_t* data,
  const sif_field_t* randoms, const sif_real* radii, uint32_t n_radii, sif_option opt,
  sif_real offset[3], sif_real* box_length);

/* exodus_finder.h:409-531: the checkpoint a run writes after every rung. A
 * 192-byte header, then in host byte order the ladder largest first (n_radii
 * floats) and the voids as four columns of n_voids floats (every x, then every
 * y, z and radius). Float run parameters are stored widened to double. */
#define SIF_FINDER_CHECKPOINT_MAGIC "XCKP"
#define SIF_FINDER_CHECKPOINT_VERSION 3

typedef struct {
  char magic[4];
  uint32_t version;
  uint32_t byte_order;   /* 0x01020304 as written */
  uint32_t is_double;    /* 1 for a 64-bit sif_real */
  uint32_t is_survey;
  uint32_t is_weighted;
  uint32_t options;      /* the options that change the catalogue */
  uint32_t n_radii;
  uint32_t next_rung;    /* rungs done */
  uint32_t grid_n_cells;
  double box_length;
  double threshold;
  double overlap_fraction;
  uint64_t n_particles;
  double total_weight;
  uint64_t random_n_particles;
  double random_total_weight;
  double offset[3];
  uint64_t n_prior;      /* the first n_prior voids came from extras->prior */
  uint32_t prior_crc32;  /* CRC32 of the prior's four columns, 0 for none */
  uint32_t crc32;        /* CRC32 of the header (this field 0) and the rest */
  uint64_t n_voids;
  uint64_t n_mismatched; /* added voids more than twice their rung */
  uint32_t assignment;   /* sif_assignment_t, | 1 << 8 if interlaced */
  char reserved[36];
} sif_finder_exodus_checkpoint_t;

/* exodus_finder.h:533-549: the header and the voids (caller's frame) of a
 * checkpoint; NULL, with the error record set, if the file cannot be read, is
 * not a checkpoint, is of another byte order or precision, or fails its
 * checksum. Older versions are brought to this one's option meanings. */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_finder_exodus_checkpoint_read(
  const char* filepath, sif_finder_exodus_checkpoint_t* header);

#define SIF_FINDER_MESH_PARTICLES_PER_CELL 30.0
#define SIF_FINDER_MESH_MAX_CELLS          1024u

/* exodus_finder.h:654-685: mesh resolution at about 30 tracers per cell,
 * floored so the widest search sphere fits, clamped to [8, 1024]. */
static inline uint32_t sif_finder_suggest_mesh_cells(uint64_t n_particles,
  sif_real box_length, sif_real max_radius, sif_real volume) {
  if (n_particles == 0 || !(box_length > 0.0f) || !isfinite(box_length) ||
      !(volume >= 0.0f) || !isfinite(volume))
    return 0;
  const double side = (double)box_length;
  const double as_box = volume > 0.0f
                          ? (double)n_particles * side * side * side / volume
                          : (double)n_particles;
  double cells = cbrt(as_box / SIF_FINDER_MESH_PARTICLES_PER_CELL);
  if (max_radius > 0.0f) {
    const double reach = 2.0 * (double)max_radius;
    if (reach >= side)
      return 0;
    const double least = side / (side - reach) + 1.0;
    if (cells < least)
      cells = least;
  }
  if (cells < 8.0)
    cells = 8.0;
  if (cells > (double)SIF_FINDER_MESH_MAX_CELLS)
    cells = (double)SIF_FINDER_MESH_MAX_CELLS;
  return (uint32_t)cells;
}

typedef enum {
  SIF_FINDER_UNITS_PHYSICAL = 0,
  SIF_FINDER_UNITS_MPS
} sif_finder_units_t;

/* exodus_finder.h:719 */
SIF_API SIF_NODISCARD uint32_t sif_finder_grid_cells_from_size(
  uint64_t n_particles, sif_real box_length, sif_real cell_size,
  sif_finder_units_t units, sif_real volume);

SIF_EXTERN_C_END

#endif
