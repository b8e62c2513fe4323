/*
 * Void catalogue: SoA centres and radii over one arena. Layout and functions
 * of include/sif/structures/catalogue.h:33-140 (sif v0.5.0).
 */
#ifndef SIF_STRUCTURES_CATALOGUE_H
#define SIF_STRUCTURES_CATALOGUE_H

#include "sif/core/macros.h"

SIF_EXTERN_C_BEGIN

struct sif_catalogue_meta_entry;

typedef struct {
  sif_real* _block;
  sif_real* cx;
  sif_real* cy;
  sif_real* cz;
  sif_real* radii;
  sif_real* _footprint_block;
  sif_real* footprint;
  sif_real* footprint_shell;
  sif_coordinates_t units;
  struct sif_catalogue_meta_entry* _meta;
  uint32_t _n_meta;
  uint64_t n_voids;
  uint64_t capacity;
} sif_catalogue_t;

#define SIF_CATALOGUE_FOOTPRINT_UNKNOWN ((sif_real)(-1.0))

SIF_API SIF_NODISCARD sif_catalogue_t* sif_catalogue_alloc(
  uint64_t initial_capacity);
SIF_API void sif_catalogue_free(sif_catalogue_t* catalogue);
SIF_API int sif_catalogue_append(
  sif_catalogue_t* catalogue, sif_real x, sif_real y, sif_real z, sif_real r);
SIF_API int sif_catalogue_trim(sif_catalogue_t* catalogue);
/* catalogue.h:166: offset added to every centre; SIF_ERR_INVALID on a sky
 * catalogue. */
SIF_API int sif_catalogue_translate(sif_catalogue_t* catalogue, const sif_real offset[3]);
/* catalogue.h:224-237: footprint columns for the catalogue, every void
 * already in it at SIF_CATALOGUE_FOOTPRINT_UNKNOWN (as is every void
 * appended later); nothing when it has them. */
SIF_API int sif_catalogue_reserve_footprint(sif_catalogue_t* catalogue);

SIF_EXTERN_C_END

#endif
