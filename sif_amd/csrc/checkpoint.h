// The exodus checkpoint (checkpoint.hip): the reference's format and rules.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "sif/finder/exodus_finder.h"

namespace sif_amd {

// four columns of n floats: x, y, z, radius
struct Columns {
  const float* col[4];
  uint64_t n;
};

// exodus_checkpoint.c:57-70: CRC32 of a catalogue's four columns
// The running CRC32 of utils/crc32.h: start from 0xFFFFFFFF, update over
// each buffer in turn, finish by XOR with 0xFFFFFFFF.
uint32_t crc32_update(uint32_t crc, const void* data, size_t n);
uint32_t catalogue_crc(const float* x, const float* y, const float* z, const float* r,
  uint64_t n);
// exodus-<16 hex digits>.ckpt, a name no earlier run has used
void checkpoint_name(char* out, size_t size);
// the header completed (magic, version, marks, count, checksum) and the
// file replaced whole
int checkpoint_write(const char* path, sif_finder_exodus_checkpoint_t* h, const float* radii,
  const Columns& v);
int checkpoint_load(const char* path, sif_finder_exodus_checkpoint_t* out,
  std::vector<float>* radii, std::vector<float> cols[4]);
// SIF_OK when `got` describes the run `want`, else SIF_ERR_INVALID naming what
// differs
int checkpoint_match(const char* path, const sif_finder_exodus_checkpoint_t& want,
  const float* want_radii, const sif_finder_exodus_checkpoint_t& got, const float* got_radii);

}  // namespace sif_amd
