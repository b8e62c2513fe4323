/*
 * sif-amd: the .xfield particle format of the reference
 * (include/sif/io/field_io.h:1-130, src/io/field_io.c:55-345, sif v0.5.0),
 * byte for byte: a 64-byte header, then the x, y and z blocks, the vx, vy
 * and vz blocks when present, and the weight block when present, each
 * n_particles sif_real in host byte order. Since version 2 the header holds
 * the CRC32 of that payload in write order; version 1 files (no checksum)
 * are still read.
 */
#ifndef SIF_IO_FIELD_IO_H
#define SIF_IO_FIELD_IO_H

#include <stdint.h>

#include "sif/core/macros.h"
#include "sif/structures/field.h"

SIF_EXTERN_C_BEGIN

#define SIF_XFIELD_MAGIC "XFLD"
#define SIF_XFIELD_VERSION 2

/* field_io.h:51-84: fixed width; box_length is a double whatever sif_real
 * is. */
typedef struct {
  char magic[4];
  uint32_t version;
  uint64_t n_particles;
  double box_length;
  uint32_t has_weights;
  uint32_t has_velocities;
  uint32_t is_double;
  uint32_t crc32;  /* of the payload; 0 in version 1 */
  char padding[24];
} sif_xfield_header_t;

/* field_io.h:86-98: a new field from the file, the blocks it holds and no
 * more; *out_box_length (optional) gets the header's box. NULL on failure
 * (unreadable, not an .xfield, a newer version, another precision, empty, a
 * failed checksum). */
SIF_API SIF_NODISCARD sif_field_t* sif_field_read(const char* filepath,
  double* out_box_length);

/* field_io.h:100-111: into an existing field of the same particle count,
 * whose arrays for the file's blocks are reserved. SIF_ERR_INVALID on a NULL
 * argument, SIF_ERR_IO on any file problem or mismatch. */
SIF_API int sif_field_read_into(sif_field_t* field, const char* filepath);

/* field_io.h:113-130: positions, and velocities and weights when the field
 * carries them. SIF_ERR_INVALID on a NULL argument, a field without
 * positions or one of sky coordinates; SIF_ERR_IO when the file cannot be
 * written or flushed. */
SIF_API int sif_field_write(const char* filepath, const sif_field_t* field,
  double box_length);

SIF_EXTERN_C_END

#endif /* SIF_IO_FIELD_IO_H */
