/*
 * Particle field: SoA positions, optional velocities and weights. Layout and
 * functions of include/sif/structures/field.h:90-286 (sif v0.5.0), kept
 * byte-compatible so a reference caller can hand its fields over.
 */
#ifndef SIF_STRUCTURES_FIELD_H
#define SIF_STRUCTURES_FIELD_H

#include "sif/core/macros.h"

SIF_EXTERN_C_BEGIN

#define SIF_FIELD_STATE_MORTON_SORTED (1u << 0)
#define SIF_FIELD_STATE_BOUNDS_VALID  (1u << 1)

typedef struct {
  sif_real* _position_block;
  sif_real* _velocity_block;
  sif_real* x;
  sif_real* y;
  sif_real* z;
  sif_real* vx;
  sif_real* vy;
  sif_real* vz;
  sif_real* weights;
  uint64_t* original_indices;
  sif_real min_p[3];
  sif_real max_p[3];
  sif_real centre[3];
  sif_real half_span;
  uint32_t state_flags;
  uint64_t n_particles;
  sif_coordinates_t units;
} sif_field_t;

/* field.h:146 */
SIF_API SIF_NODISCARD sif_field_t* sif_field_alloc(uint64_t n_particles);
/* field.h:159: per-array stride of a packed x/y/z block (64-byte lines). */
SIF_API uint64_t sif_field_padded_n(uint64_t n_particles);
SIF_API void sif_field_free(sif_field_t* field);
SIF_API SIF_NODISCARD int sif_field_reserve_positions(sif_field_t* field);
SIF_API SIF_NODISCARD int sif_field_reserve_velocities(sif_field_t* field);
SIF_API SIF_NODISCARD int sif_field_reserve_weights(sif_field_t* field);
/* field.h:203 */
SIF_API SIF_NODISCARD int sif_field_assign_positions(sif_field_t* field,
  const sif_real* x, const sif_real* y, const sif_real* z);
SIF_API SIF_NODISCARD int sif_field_assign_velocities(sif_field_t* field,
  const sif_real* vx, const sif_real* vy, const sif_real* vz);
SIF_API SIF_NODISCARD int sif_field_assign_weights(
  sif_field_t* field, const sif_real* weights);
/* field.h:266: fold every coordinate into [0, box_length); counts of those
 * exactly on the upper face, folded from outside, and not finite. */
SIF_API SIF_NODISCARD int sif_field_wrap_periodic(sif_field_t* field,
  sif_real box_length, uint64_t* n_boundary, uint64_t* n_wrapped,
  uint64_t* n_nonfinite);
/* field.h:286: add offset to every position. */
SIF_API SIF_NODISCARD int sif_field_translate(
  sif_field_t* field, const sif_real offset[3]);

SIF_EXTERN_C_END

#endif
