/*
 * sif-amd: void catalogues as text, the reference's
 * include/sif/io/catalogue_io.h:1-110 and src/io/catalogue_io.c (sif
 * v0.5.0): "#n=<count>", a comment naming the columns ("cx cy cz r", "ra dec
 * z r" on the sky, " footprint footprint_shell" when the catalogue carries
 * one), then a row a void, values with "%.9g", which gives a float back
 * exactly. This build keeps no catalogue metadata, so it writes none, and
 * "key=value" comments are read and dropped.
 */
#ifndef SIF_IO_CATALOGUE_IO_H
#define SIF_IO_CATALOGUE_IO_H

#include "sif/core/macros.h"
#include "sif/structures/catalogue.h"

SIF_EXTERN_C_BEGIN

/* catalogue_io.h:38-47. SIF_ERR_INVALID on a NULL argument, SIF_ERR_IO when
 * the file cannot be written or flushed. */
SIF_API int sif_catalogue_write_ascii(const char* filepath,
  const sif_catalogue_t* catalogue);

/* catalogue_io.h:49-83: comments are lines starting with '#', anywhere.
 * Before the first row: "n=N" sizes the catalogue (without it the rows are
 * counted), a comment naming the centre and radius columns ("cx cy cz" or
 * "x y z", or "ra dec z"; "r" or "radius"; optionally "footprint
 * footprint_shell") places them, in any order. Without names a row is "cx cy
 * cz r", or six values with the footprint; a first line holding the count
 * alone (older files) is read as the count. fmt, when given, places the
 * columns itself: "x y z" ("cx cy cz") or "ra dec z", "r", and "*" for a
 * column not read, e.g. "* x y z * r". Values are separated by blanks or
 * commas. Rows must be exactly as many as the count. NULL on failure. */
SIF_API SIF_NODISCARD sif_catalogue_t* sif_catalogue_read_ascii(
  const char* filepath, const char* fmt);

SIF_EXTERN_C_END

#endif /* SIF_IO_CATALOGUE_IO_H */
