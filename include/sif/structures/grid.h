/*
 * Cubic density grid and its CIC assignment. Layout and functions of
 * include/sif/structures/grid.h:86-233 (sif v0.5.0). Values are row-major,
 * z fastest. sif_grid_assign_cic() runs the fixed-point deterministic HIP
 * deposit; the result is bit-identical for any launch shape.
 */
#ifndef SIF_STRUCTURES_GRID_H
#define SIF_STRUCTURES_GRID_H

#include "sif/core/macros.h"
#include "sif/structures/field.h"

SIF_EXTERN_C_BEGIN

typedef enum {
  SIF_GRID_EMPTY = 0,
  SIF_GRID_DENSITY,
  SIF_GRID_DENSITY_CONTRAST
} sif_grid_content_t;

typedef enum {
  SIF_ASSIGNMENT_UNKNOWN = 0,
  SIF_ASSIGNMENT_NGP = 1,
  SIF_ASSIGNMENT_CIC = 2,
  SIF_ASSIGNMENT_TSC = 3,
  SIF_ASSIGNMENT_PCS = 4
} sif_assignment_t;

#define SIF_GRID_INTERLACE (1u << 8)

typedef struct {
  sif_real* values;
  sif_grid_content_t content;
  sif_assignment_t assignment;
  uint32_t interlaced;
  uint32_t n_cells;
  uint64_t total_cells;
  sif_real box_length;
  sif_real cell_length;
  uint32_t p2_mask;
  uint32_t p2_shift;
} sif_grid_t;

/* grid.h:140 */
SIF_API SIF_NODISCARD sif_grid_t* sif_grid_alloc(
  uint32_t n_cells, sif_real box_length);
SIF_API void sif_grid_free(sif_grid_t* grid);
/* grid.h:171: CIC only in this build (other schemes: SIF_ERR_UNSUPPORTED). */
SIF_API SIF_NODISCARD int sif_grid_assign(sif_grid_t* grid,
  const sif_field_t* field, sif_assignment_t scheme, sif_option opt);
/* grid.h:205 */
SIF_API SIF_NODISCARD int sif_grid_assign_cic(
  sif_grid_t* grid, const sif_field_t* field);
/* grid.h:233 */
SIF_API SIF_NODISCARD int sif_grid_to_density_contrast(sif_grid_t* grid);

SIF_EXTERN_C_END

#endif
