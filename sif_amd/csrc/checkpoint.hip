// The exodus checkpoint (include/sif/finder/exodus_finder.h:409-549): the
// run's description and its catalogue after each rung, so that a run can be
// resumed from the last rung it finished and return what it would have
// uninterrupted. The format is the reference's (exodus_checkpoint.c), byte
// for byte: a 192-byte header, the ladder, the voids as four columns, and a
// CRC32 (IEEE 802.3, reflected, as sif_crc32) of the header with its crc32
// field zero followed by the rest. Host code only.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "checkpoint.h"
#include "common.h"
#include "sif/finder/exodus_finder.h"

namespace sif_amd {

static_assert(sizeof(sif_finder_exodus_checkpoint_t) == 192,
  "the checkpoint header is 192 bytes (exodus_finder.h:427-531)");

namespace {

constexpr uint32_t BYTE_ORDER_MARK = 0x01020304u;

struct Crc32 {
  uint32_t table[256];
  Crc32() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++)
        c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
      table[i] = c;
    }
  }
  uint32_t update(uint32_t crc, const void* data, size_t n) const {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    for (size_t i = 0; i < n; i++)
      crc = table[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
    return crc;
  }
};

const Crc32& crc_table() {
  static const Crc32 t;
  return t;
}

uint32_t file_crc(const sif_finder_exodus_checkpoint_t& h, const float* radii,
  const Columns& v) {
  sif_finder_exodus_checkpoint_t z = h;
  z.crc32 = 0;
  const Crc32& t = crc_table();
  uint32_t c = 0xFFFFFFFFu;
  c = t.update(c, &z, sizeof(z));
  c = t.update(c, radii, (size_t)h.n_radii * sizeof(float));
  for (int k = 0; k < 4; k++)
    c = t.update(c, v.col[k], (size_t)v.n * sizeof(float));
  return c ^ 0xFFFFFFFFu;
}

int read_failed(const char* path, const char* what) {
  error_set(SIF_ERR_IO, "checkpoint: %s: %s", path, what);
  return SIF_ERR_IO;
}

}  // namespace

uint32_t crc32_update(uint32_t crc, const void* data, size_t n) {
  return crc_table().update(crc, data, n);
}

uint32_t catalogue_crc(const float* x, const float* y, const float* z, const float* r,
  uint64_t n) {
  const Crc32& t = crc_table();
  uint32_t c = 0xFFFFFFFFu;
  const float* cols[4] = {x, y, z, r};
  for (int k = 0; k < 4; k++)
    c = t.update(c, cols[k], (size_t)n * sizeof(float));
  return c ^ 0xFFFFFFFFu;
}

void checkpoint_name(char* out, size_t size) {
  uint64_t id = 0;
  const int fd = open("/dev/urandom", O_RDONLY);
  if (fd >= 0) {
    if (read(fd, &id, sizeof(id)) != (ssize_t)sizeof(id))
      id = 0;
    close(fd);
  }
  if (id == 0)
    id = ((uint64_t)getpid() << 32) ^ (uint64_t)time(nullptr);
  snprintf(out, size, "exodus-%016" PRIx64 ".ckpt", id);
}

int checkpoint_write(const char* path, sif_finder_exodus_checkpoint_t* h, const float* radii,
  const Columns& v) {
  memcpy(h->magic, SIF_FINDER_CHECKPOINT_MAGIC, 4);
  h->version = SIF_FINDER_CHECKPOINT_VERSION;
  h->byte_order = BYTE_ORDER_MARK;
  h->is_double = 0;
  h->n_voids = v.n;
  memset(h->reserved, 0, sizeof(h->reserved));
  h->crc32 = file_crc(*h, radii, v);
  // written beside the target and renamed over it: a reader sees the old
  // checkpoint or the new one, never half of one
  const std::string tmp = std::string(path) + ".tmp." + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) {
    error_set(SIF_ERR_IO, "checkpoint: %s: cannot create: %s", tmp.c_str(), strerror(errno));
    return SIF_ERR_IO;
  }
  bool ok = fwrite(h, sizeof(*h), 1, f) == 1 &&
            fwrite(radii, sizeof(float), h->n_radii, f) == h->n_radii;
  for (int k = 0; k < 4 && ok; k++)
    ok = fwrite(v.col[k], sizeof(float), (size_t)v.n, f) == (size_t)v.n;
  const int err = ok ? 0 : errno;
  if (fclose(f) != 0 && ok) {
    ok = false;
  }
  if (!ok || rename(tmp.c_str(), path) != 0) {
    const int e = err ? err : errno;
    unlink(tmp.c_str());
    error_set(SIF_ERR_IO, "checkpoint: %s: cannot write: %s", path, strerror(e));
    return SIF_ERR_IO;
  }
  return SIF_OK;
}

// exodus_finder.h:463-486: an older header brought to this version's option
// meanings (v1: CIC grids; up to v2: bit 10 was the Gaussian refinement, and
// minima implied it, the extended reach and displacement)
static void upgrade(sif_finder_exodus_checkpoint_t* h) {
  if (h->version < 2)
    h->assignment = SIF_ASSIGNMENT_CIC;
  if (h->version < 3) {
    const uint32_t old_refine = 1u << 10;
    if (h->options & old_refine)
      h->options |= SIF_FINDER_REFINE_GAUSSIAN;
    if (h->options & SIF_FINDER_SEEDS_MINIMA)
      h->options |= SIF_FINDER_REFINE_GAUSSIAN | SIF_FINDER_EXTEND_REACH | SIF_FINDER_DISPLACE;
    h->options &= ~old_refine;
  }
}

int checkpoint_load(const char* path, sif_finder_exodus_checkpoint_t* out,
  std::vector<float>* radii, std::vector<float> cols[4]) {
  FILE* f = fopen(path, "rb");
  if (!f) {
    error_set(SIF_ERR_IO, "checkpoint: %s: cannot open: %s", path, strerror(errno));
    return SIF_ERR_IO;
  }
  struct stat sb;
  int st = SIF_ERR_IO;
  sif_finder_exodus_checkpoint_t h;
  char what[256];
  do {
    if (fstat(fileno(f), &sb) != 0) {
      st = read_failed(path, "cannot stat");
      break;
    }
    const uint64_t bytes = (uint64_t)sb.st_size;
    if (bytes < sizeof(h) || fread(&h, sizeof(h), 1, f) != 1) {
      snprintf(what, sizeof(what), "%" PRIu64 " bytes, shorter than a header (%zu)", bytes,
        sizeof(h));
      st = read_failed(path, what);
      break;
    }
    if (memcmp(h.magic, SIF_FINDER_CHECKPOINT_MAGIC, 4) != 0) {
      st = read_failed(path, "not an exodus checkpoint (no XCKP at byte 0)");
      break;
    }
    if (h.byte_order != BYTE_ORDER_MARK) {
      st = read_failed(path, "written on a machine of the other byte order");
      break;
    }
    if (h.version == 0 || h.version > SIF_FINDER_CHECKPOINT_VERSION) {
      snprintf(what, sizeof(what), "version %u (expected 1 to %u)", h.version,
        (unsigned)SIF_FINDER_CHECKPOINT_VERSION);
      st = read_failed(path, what);
      break;
    }
    if (h.is_double != 0) {
      st = read_failed(path, "FP64 values, this build is FP32");
      break;
    }
    if (h.n_radii == 0 || h.next_rung > h.n_radii || h.n_prior > h.n_voids) {
      snprintf(what, sizeof(what),
        "header: %u rungs, %u done, %" PRIu64 " prior voids of %" PRIu64 " (inconsistent)",
        h.n_radii, h.next_rung, h.n_prior, h.n_voids);
      st = read_failed(path, what);
      break;
    }
    // sized against the file before anything is allocated from the header
    const uint64_t expect = sizeof(h) + ((uint64_t)h.n_radii + 4u * h.n_voids) * 4u;
    if (h.n_voids > UINT64_MAX / 64u || bytes != expect) {
      snprintf(what, sizeof(what), "%" PRIu64 " bytes, header says %" PRIu64
        " (%u rungs, %" PRIu64 " voids)", bytes, expect, h.n_radii, h.n_voids);
      st = read_failed(path, what);
      break;
    }
    radii->resize(h.n_radii);
    bool ok = fread(radii->data(), 4, h.n_radii, f) == h.n_radii;
    Columns v;
    v.n = h.n_voids;
    for (int k = 0; k < 4 && ok; k++) {
      cols[k].resize(h.n_voids);
      ok = fread(cols[k].data(), 4, (size_t)h.n_voids, f) == (size_t)h.n_voids;
      v.col[k] = cols[k].data();
    }
    if (!ok) {
      st = read_failed(path, "short read");
      break;
    }
    const uint32_t got = file_crc(h, radii->data(), v);
    if (got != h.crc32) {
      snprintf(what, sizeof(what), "corrupt: checksum %08x (expected %08x)", got, h.crc32);
      st = read_failed(path, what);
      break;
    }
    upgrade(&h);  // after the checksum, which covers the header as written
    *out = h;
    st = SIF_OK;
  } while (false);
  fclose(f);
  return st;
}

int checkpoint_match(const char* path, const sif_finder_exodus_checkpoint_t& want,
  const float* want_radii, const sif_finder_exodus_checkpoint_t& got, const float* got_radii) {
  char what[384];
  auto mismatch = [&](const char* fmt, auto... args) {
    char d[320];
    snprintf(d, sizeof(d), fmt, args...);
    snprintf(what, sizeof(what), "checkpoint: %s: not this run: %s", path, d);
    error_set(SIF_ERR_INVALID, "%s", what);
    return SIF_ERR_INVALID;
  };
  if (got.is_survey != want.is_survey)
    return mismatch("saved from %s (this run is %s)",
      got.is_survey ? "a survey" : "a periodic box", want.is_survey ? "a survey" : "a periodic box");
  if (got.is_weighted != want.is_weighted)
    return mismatch("saved from %s tracers (this run's are %s)",
      got.is_weighted ? "weighted" : "unweighted", want.is_weighted ? "weighted" : "unweighted");
  if (got.n_radii != want.n_radii)
    return mismatch("ladder of %u radii (expected %u)", got.n_radii, want.n_radii);
  for (uint32_t i = 0; i < want.n_radii; i++)
    if (got_radii[i] != want_radii[i])
      return mismatch("radius %u of the ladder, largest first: %g (expected %g)", i,
        (double)got_radii[i], (double)want_radii[i]);
  if (got.threshold != want.threshold)
    return mismatch("threshold %g (expected %g)", got.threshold, want.threshold);
  if (got.overlap_fraction != want.overlap_fraction)
    return mismatch("overlap fraction %g (expected %g)", got.overlap_fraction,
      want.overlap_fraction);
  if ((got.options & SIF__FINDER_SEARCH_MASK) != (want.options & SIF__FINDER_SEARCH_MASK))
    return mismatch("search factor %g (expected %g)",
      (double)sif__finder_search_factor(got.options),
      (double)sif__finder_search_factor(want.options));
  const uint32_t diff = got.options ^ want.options;
  if (diff & SIF_FINDER_KEEP_CIC_WINDOW)
    return mismatch("CIC window %s (this run %s it)",
      (got.options & SIF_FINDER_KEEP_CIC_WINDOW) ? "kept" : "deconvolved",
      (want.options & SIF_FINDER_KEEP_CIC_WINDOW) ? "keeps" : "deconvolves");
  if (diff & (SIF__FINDER_SEEDS_MASK | SIF__FINDER_REFINE_MASK))
    return mismatch("centres from options %#x (this run's: %#x)",
      got.options & (SIF__FINDER_SEEDS_MASK | SIF__FINDER_REFINE_MASK),
      want.options & (SIF__FINDER_SEEDS_MASK | SIF__FINDER_REFINE_MASK));
  if (diff & SIF_FINDER_EXTEND_REACH)
    return mismatch("reach %s (this run's: %s)",
      (got.options & SIF_FINDER_EXTEND_REACH) ? "extended" : "fixed",
      (want.options & SIF_FINDER_EXTEND_REACH) ? "extended" : "fixed");
  if (diff & SIF_FINDER_DISPLACE)
    return mismatch("displacement %s (this run's: %s)",
      (got.options & SIF_FINDER_DISPLACE) ? "on" : "off",
      (want.options & SIF_FINDER_DISPLACE) ? "on" : "off");
  if (diff & SIF_FINDER_LAGRANGIAN_OVERLAP)
    return mismatch("overlap on %s radii (this run's: %s)",
      (got.options & SIF_FINDER_LAGRANGIAN_OVERLAP) ? "Lagrangian" : "Eulerian",
      (want.options & SIF_FINDER_LAGRANGIAN_OVERLAP) ? "Lagrangian" : "Eulerian");
  if (got.grid_n_cells != want.grid_n_cells || got.box_length != want.box_length)
    return mismatch("grid of %u cells over %g (expected %u over %g)", got.grid_n_cells,
      got.box_length, want.grid_n_cells, want.box_length);
  if (diff & SIF_FINDER_ALIASING_WINDOW)
    return mismatch("window deconvolved %s (this run's: %s)",
      (got.options & SIF_FINDER_ALIASING_WINDOW) ? "allowing for aliasing" : "in full",
      (want.options & SIF_FINDER_ALIASING_WINDOW) ? "allowing for aliasing" : "in full");
  if (got.assignment != want.assignment)
    return mismatch("grid assignment %#x (this run's: %#x)", got.assignment, want.assignment);
  // a total weight is a parallel sum: equal to 1e-9, not bit for bit
  const bool weight_ok = std::fabs(got.total_weight - want.total_weight) <=
                         1e-9 * std::fmax(std::fabs(got.total_weight),
                                  std::fabs(want.total_weight));
  if (got.n_particles != want.n_particles || !weight_ok)
    return mismatch("%" PRIu64 " tracers of total weight %.17g (expected %" PRIu64 " of %.17g)",
      got.n_particles, got.total_weight, want.n_particles, want.total_weight);
  if (want.is_survey) {  // exodus_checkpoint.c:417-430
    const bool rw_ok = std::fabs(got.random_total_weight - want.random_total_weight) <=
                       1e-9 * std::fmax(std::fabs(got.random_total_weight),
                                std::fabs(want.random_total_weight));
    if (got.random_n_particles != want.random_n_particles || !rw_ok)
      return mismatch("%" PRIu64 " randoms of total weight %.17g (expected %" PRIu64
        " of %.17g)", got.random_n_particles, got.random_total_weight,
        want.random_n_particles, want.random_total_weight);
    for (int k = 0; k < 3; k++)
      if (got.offset[k] != want.offset[k])
        return mismatch("offset[%d] %g (expected %g)", k, got.offset[k], want.offset[k]);
  }
  if (got.n_prior != want.n_prior || got.prior_crc32 != want.prior_crc32)
    return mismatch("started from %" PRIu64 " prior voids, checksum %08x (this run's: %" PRIu64
      ", %08x)", got.n_prior, got.prior_crc32, want.n_prior, want.prior_crc32);
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" sif_catalogue_t* sif_finder_exodus_checkpoint_read(const char* filepath,
  sif_finder_exodus_checkpoint_t* header) {
  if (!filepath) {
    error_set(SIF_ERR_INVALID, "filepath: NULL");
    return nullptr;
  }
  sif_finder_exodus_checkpoint_t h;
  std::vector<float> radii, cols[4];
  if (checkpoint_load(filepath, &h, &radii, cols) != SIF_OK)
    return nullptr;
  sif_catalogue_t* cat = sif_catalogue_alloc(h.n_voids ? h.n_voids : 1);
  if (!cat) {
    error_set(SIF_ERR_ALLOC, "checkpoint: %s: cannot allocate %" PRIu64 " voids", filepath,
      h.n_voids);
    return nullptr;
  }
  // a survey's voids sit in its box: back to the caller's frame
  const float back[3] = {(float)-h.offset[0], (float)-h.offset[1], (float)-h.offset[2]};
  for (uint64_t i = 0; i < h.n_voids; i++)
    if (sif_catalogue_append(cat, cols[0][i] + (h.is_survey ? back[0] : 0.0f),
          cols[1][i] + (h.is_survey ? back[1] : 0.0f),
          cols[2][i] + (h.is_survey ? back[2] : 0.0f), cols[3][i]) != SIF_OK) {
      sif_catalogue_free(cat);
      return nullptr;
    }
  if (header)
    *header = h;
  return cat;
}
