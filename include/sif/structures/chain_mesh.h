/*
 * Chain mesh: a cell-sorted copy of a field. Layout and functions of
 * include/sif/structures/chain_mesh.h:40-215 (sif v0.5.0). Built by a GPU
 * counting sort whose cell_offsets and per-cell order (ascending source index)
 * are bit-exact with the reference's canonical mesh.
 */
#ifndef SIF_STRUCTURES_CHAIN_MESH_H
#define SIF_STRUCTURES_CHAIN_MESH_H

#include "sif/core/macros.h"
#include "sif/structures/field.h"

SIF_EXTERN_C_BEGIN

typedef struct {
  uint32_t n_cells;
  uint64_t total_cells;
  sif_real box_length;
  sif_real cell_length;
  sif_real* _position_block;
  sif_real* _velocity_block;
  sif_real* x;
  sif_real* y;
  sif_real* z;
  sif_real* vx;
  sif_real* vy;
  sif_real* vz;
  sif_real* weights;
  double total_weight;
  sif_real* cell_weights;
  uint64_t* original_indices;
  uint64_t n_particles;
  uint64_t* cell_offsets;
} sif_chain_mesh_t;

/* chain_mesh.h:141 */
SIF_API SIF_NODISCARD sif_chain_mesh_t* sif_chain_mesh_alloc(uint32_t n_cells,
  sif_real box_length, const sif_field_t* field, sif_option opt);
/* chain_mesh.h:170 */
SIF_API SIF_NODISCARD sif_chain_mesh_t* sif_chain_mesh_alloc_consume(
  uint32_t n_cells, sif_real box_length, sif_field_t* field, sif_option opt);
SIF_API void sif_chain_mesh_free(sif_chain_mesh_t* mesh);
/* chain_mesh.h:194, :211 */
SIF_API uint64_t sif_chain_mesh_find_nearest_open(
  const sif_chain_mesh_t* mesh, sif_real px, sif_real py, sif_real pz);
SIF_API uint64_t sif_chain_mesh_find_nearest_pbc(
  const sif_chain_mesh_t* mesh, sif_real px, sif_real py, sif_real pz);

SIF_EXTERN_C_END

#endif
