/*
 * sif-amd: the .xgrid grid format of the reference
 * (include/sif/io/grid_io.h:1-124, src/io/grid_io.c:41-275, sif v0.5.0),
 * byte for byte: a 64-byte header, then the n_cells^3 cell values in host
 * byte order. Since version 2 the header holds the CRC32 of the values and
 * what they are (density or contrast); version 1 files are still read.
 */
#ifndef SIF_IO_GRID_IO_H
#define SIF_IO_GRID_IO_H

#include <stdint.h>

#include "sif/core/macros.h"
#include "sif/structures/grid.h"

SIF_EXTERN_C_BEGIN

#define SIF_XGRID_MAGIC "XGRD"
#define SIF_XGRID_VERSION 2

/* grid_io.h:35-84 */
typedef struct {
  char magic[4];
  uint32_t version;
  uint32_t n_cells;
  uint32_t is_double;
  uint64_t total_cells;
  double box_length;
  uint32_t crc32;       /* of the cell values; 0 in version 1 */
  uint32_t content;     /* a sif_grid_content_t */
  uint64_t source_key[2];  /* provenance of a cached grid; 0 when none */
  uint32_t assignment;  /* the scheme in the low byte, 1 in the next if interlaced */
  char padding[4];
} sif_xgrid_header_t;

/* grid_io.h:86-95: a new grid from the file. NULL on failure. */
SIF_API SIF_NODISCARD sif_grid_t* sif_grid_read(const char* filepath);

/* grid_io.h:97-106: into an existing grid of the same cells and box (the
 * box compared exactly). SIF_ERR_INVALID on a NULL argument, SIF_ERR_IO on
 * any file problem or mismatch. The grid's content and assignment are the
 * file's. */
SIF_API int sif_grid_read_into(sif_grid_t* grid, const char* filepath);

/* grid_io.h:108-120: the values as they are, with what they hold.
 * SIF_ERR_INVALID on a NULL argument or a grid without values, SIF_ERR_IO
 * when the file cannot be written or flushed. */
SIF_API int sif_grid_write(const char* filepath, const sif_grid_t* grid);

SIF_EXTERN_C_END

#endif /* SIF_IO_GRID_IO_H */
