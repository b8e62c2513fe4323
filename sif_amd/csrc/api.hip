// The sif v0.5.0 C API of the exodus path, on host structs, over the HIP
// stages: field (include/sif/structures/field.h:146-286), grid (grid.h:140-233),
// chain mesh (chain_mesh.h:141-215), catalogue (catalogue.h:110-153) and the
// exodus finder (exodus_finder.h:147-723).
//
// The structs keep the reference layouts and own host memory, so a caller of
// the reference sees no difference. Each call moves what it needs to the
// device, runs there, and brings the result back before returning.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <unistd.h>

#include "checkpoint.h"
#include "common.h"
#include "internal.h"
#include "sif/core/system.h"
#include "sif/finder/exodus_finder.h"
#include "sif/measure/profiles.h"
#include "sif/structures/catalogue.h"
#include "sif/structures/chain_mesh.h"
#include "sif/structures/field.h"
#include "sif/structures/grid.h"
#include "sif/utils/logger.h"
#include "sif_amd.h"

namespace sif_amd {

namespace {

constexpr size_t LINE = 64;

void* aligned_alloc_bytes(size_t bytes) {
  void* p = nullptr;
  const size_t padded = (bytes + LINE - 1) / LINE * LINE;
  if (posix_memalign(&p, LINE, padded ? padded : LINE) != 0)
    return nullptr;
  return p;
}

// A device copy of host arrays, released on scope exit.
struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p)
      (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
      p = nullptr;
      error_set(SIF_ERR_ALLOC, "device allocation of %zu bytes failed", bytes);
      return SIF_ERR_ALLOC;
    }
    return SIF_OK;
  }
  int upload(const void* src, size_t bytes) {
    int st = alloc(bytes);
    if (st != SIF_OK)
      return st;
    if (bytes)
      SIF_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream()));
    return SIF_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

int need_init() {
  if (!initialised()) {
    error_set(SIF_ERR_UNINITIALISED, "sif_system_init() has not been called");
    return SIF_ERR_UNINITIALISED;
  }
  return SIF_OK;
}

int field_check(const sif_field_t* f, const char* name) {
  if (!f) {
    error_set(SIF_ERR_INVALID, "%s: NULL", name);
    return SIF_ERR_INVALID;
  }
  if (f->n_particles == 0) {
    error_set(SIF_ERR_INVALID, "%s: 0 particles", name);
    return SIF_ERR_INVALID;
  }
  if (!f->x || !f->y || !f->z) {
    error_set(SIF_ERR_INVALID, "%s: no positions (%llu particles declared)", name,
      (unsigned long long)f->n_particles);
    return SIF_ERR_INVALID;
  }
  if (f->units != SIF_COORDINATES_CARTESIAN) {
    error_set(SIF_ERR_INVALID, "%s: sky coordinates (ra, dec, z), not positions", name);
    return SIF_ERR_INVALID;
  }
  return SIF_OK;
}

// Upload positions (and weights) and run the in-box check of field.c:889-935.
int upload_field(const sif_field_t* f, float box, Dev& dx, Dev& dy, Dev& dz, Dev& dw) {
  const size_t b = f->n_particles * sizeof(float);
  int st = dx.upload(f->x, b);
  if (st == SIF_OK)
    st = dy.upload(f->y, b);
  if (st == SIF_OK)
    st = dz.upload(f->z, b);
  if (st == SIF_OK && f->weights)
    st = dw.upload(f->weights, b);
  if (st != SIF_OK)
    return st;
  uint64_t bad = 0, edge = 0, first = 0;
  st = check_in_box(dx.as<float>(), dy.as<float>(), dz.as<float>(), f->n_particles, box,
    &bad, &edge, &first);
  if (st != SIF_OK)
    return st;
  if (bad) {
    error_set(SIF_ERR_RANGE,
      "field: %llu of %llu particles outside the box [0, %g) (first: particle %llu at "
      "(%g, %g, %g)); fold a periodic box back into it, or translate a survey into one",
      (unsigned long long)bad, (unsigned long long)f->n_particles, (double)box,
      (unsigned long long)first, (double)f->x[first], (double)f->y[first],
      (double)f->z[first]);
    return SIF_ERR_RANGE;
  }
  return SIF_OK;
}

}  // namespace
}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

/* --- field --- */

uint64_t sif_field_padded_n(uint64_t n) {
  const uint64_t per_line = LINE / sizeof(sif_real);
  return (n + per_line - 1) & ~(per_line - 1);
}

sif_field_t* sif_field_alloc(uint64_t n_particles) {
  sif_field_t* f = (sif_field_t*)calloc(1, sizeof(sif_field_t));
  if (!f) {
    error_set(SIF_ERR_ALLOC, "field: cannot allocate");
    return nullptr;
  }
  f->n_particles = n_particles;
  f->units = SIF_COORDINATES_CARTESIAN;
  return f;
}

void sif_field_free(sif_field_t* f) {
  if (!f)
    return;
  free(f->_position_block);
  free(f->_velocity_block);
  free(f->weights);
  free(f->original_indices);
  free(f);
}

static int reserve_block(sif_field_t* f, sif_real** block, sif_real** a, sif_real** b,
  sif_real** c) {
  if (!f || f->n_particles == 0) {
    error_set(SIF_ERR_INVALID, "field: empty");
    return SIF_ERR_INVALID;
  }
  if (*block)
    return SIF_OK;
  const uint64_t pn = sif_field_padded_n(f->n_particles);
  *block = (sif_real*)aligned_alloc_bytes(3 * pn * sizeof(sif_real));
  if (!*block) {
    error_set(SIF_ERR_ALLOC, "field: cannot allocate %llu particles",
      (unsigned long long)f->n_particles);
    return SIF_ERR_ALLOC;
  }
  *a = *block;
  *b = *block + pn;
  *c = *block + 2 * pn;
  return SIF_OK;
}

int sif_field_reserve_positions(sif_field_t* f) {
  return reserve_block(f, &f->_position_block, &f->x, &f->y, &f->z);
}

int sif_field_reserve_velocities(sif_field_t* f) {
  return reserve_block(f, &f->_velocity_block, &f->vx, &f->vy, &f->vz);
}

int sif_field_reserve_weights(sif_field_t* f) {
  if (!f || f->n_particles == 0) {
    error_set(SIF_ERR_INVALID, "field: empty");
    return SIF_ERR_INVALID;
  }
  if (f->weights)
    return SIF_OK;
  f->weights = (sif_real*)aligned_alloc_bytes(f->n_particles * sizeof(sif_real));
  if (!f->weights) {
    error_set(SIF_ERR_ALLOC, "field: cannot allocate weights");
    return SIF_ERR_ALLOC;
  }
  return SIF_OK;
}

int sif_field_assign_positions(sif_field_t* f, const sif_real* x, const sif_real* y,
  const sif_real* z) {
  if (!x || !y || !z) {
    error_set(SIF_ERR_INVALID, "positions: NULL");
    return SIF_ERR_INVALID;
  }
  int st = sif_field_reserve_positions(f);
  if (st != SIF_OK)
    return st;
  memcpy(f->x, x, f->n_particles * sizeof(sif_real));
  memcpy(f->y, y, f->n_particles * sizeof(sif_real));
  memcpy(f->z, z, f->n_particles * sizeof(sif_real));
  f->state_flags = 0;
  free(f->original_indices);
  f->original_indices = nullptr;
  return SIF_OK;
}

int sif_field_assign_velocities(sif_field_t* f, const sif_real* vx, const sif_real* vy,
  const sif_real* vz) {
  if (!vx || !vy || !vz) {
    error_set(SIF_ERR_INVALID, "velocities: NULL");
    return SIF_ERR_INVALID;
  }
  int st = sif_field_reserve_velocities(f);
  if (st != SIF_OK)
    return st;
  for (uint64_t i = 0; i < f->n_particles; i++) {
    const uint64_t s = f->original_indices ? f->original_indices[i] : i;
    f->vx[i] = vx[s];
    f->vy[i] = vy[s];
    f->vz[i] = vz[s];
  }
  return SIF_OK;
}

int sif_field_assign_weights(sif_field_t* f, const sif_real* w) {
  if (!w) {
    error_set(SIF_ERR_INVALID, "weights: NULL");
    return SIF_ERR_INVALID;
  }
  int st = sif_field_reserve_weights(f);
  if (st != SIF_OK)
    return st;
  for (uint64_t i = 0; i < f->n_particles; i++)
    f->weights[i] = w[f->original_indices ? f->original_indices[i] : i];
  return SIF_OK;
}

// field.c:332-400
int sif_field_wrap_periodic(sif_field_t* f, sif_real box, uint64_t* n_boundary,
  uint64_t* n_wrapped, uint64_t* n_nonfinite) {
  int st = field_check(f, "field");
  if (st != SIF_OK)
    return st;
  if (!(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "box_length: %g (expected > 0)", (double)box);
    return SIF_ERR_INVALID;
  }
  uint64_t boundary = 0, wrapped = 0, nonfinite = 0;
  auto wrap = [&](sif_real v) -> sif_real {
    if (v >= 0.0f && v < box)
      return v;
    if (!std::isfinite(v)) {
      nonfinite++;
      return v;
    }
    if (v == box) {
      boundary++;
      return 0.0f;
    }
    wrapped++;
    sif_real w = fmodf(v, box);
    if (w < 0.0f)
      w += box;
    if (w >= box)
      w = 0.0f;
    return w;
  };
  for (uint64_t i = 0; i < f->n_particles; i++) {
    f->x[i] = wrap(f->x[i]);
    f->y[i] = wrap(f->y[i]);
    f->z[i] = wrap(f->z[i]);
  }
  if (boundary || wrapped)
    f->state_flags &= ~(SIF_FIELD_STATE_BOUNDS_VALID | SIF_FIELD_STATE_MORTON_SORTED);
  if (n_boundary)
    *n_boundary = boundary;
  if (n_wrapped)
    *n_wrapped = wrapped;
  if (n_nonfinite)
    *n_nonfinite = nonfinite;
  if (nonfinite) {
    error_set(SIF_ERR_RANGE, "field: %llu coordinates are not finite",
      (unsigned long long)nonfinite);
    return SIF_ERR_RANGE;
  }
  return SIF_OK;
}

int sif_field_translate(sif_field_t* f, const sif_real offset[3]) {
  int st = field_check(f, "field");
  if (st != SIF_OK)
    return st;
  if (!offset) {
    error_set(SIF_ERR_INVALID, "offset: NULL");
    return SIF_ERR_INVALID;
  }
  for (uint64_t i = 0; i < f->n_particles; i++) {
    f->x[i] += offset[0];
    f->y[i] += offset[1];
    f->z[i] += offset[2];
  }
  f->state_flags &= ~SIF_FIELD_STATE_BOUNDS_VALID;
  return SIF_OK;
}

/* --- catalogue --- */

static int catalogue_resize(sif_catalogue_t* c, uint64_t cap) {
  if (cap < 1)
    cap = 1;
  const uint64_t per_line = LINE / sizeof(sif_real);
  const uint64_t stride = (cap + per_line - 1) & ~(per_line - 1);
  sif_real* block = (sif_real*)aligned_alloc_bytes(4 * stride * sizeof(sif_real));
  if (!block) {
    error_set(SIF_ERR_ALLOC, "catalogue: cannot grow to %llu voids", (unsigned long long)cap);
    return SIF_ERR_ALLOC;
  }
  const uint64_t keep = c->n_voids < cap ? c->n_voids : cap;
  if (c->_block && keep) {
    memcpy(block, c->cx, keep * sizeof(sif_real));
    memcpy(block + stride, c->cy, keep * sizeof(sif_real));
    memcpy(block + 2 * stride, c->cz, keep * sizeof(sif_real));
    memcpy(block + 3 * stride, c->radii, keep * sizeof(sif_real));
  }
  sif_real* fblock = nullptr;
  if (c->_footprint_block) {  // the footprint columns follow the same stride
    fblock = (sif_real*)aligned_alloc_bytes(2 * stride * sizeof(sif_real));
    if (!fblock) {
      free(block);
      error_set(SIF_ERR_ALLOC, "catalogue: cannot grow to %llu voids", (unsigned long long)cap);
      return SIF_ERR_ALLOC;
    }
    if (keep) {
      memcpy(fblock, c->footprint, keep * sizeof(sif_real));
      memcpy(fblock + stride, c->footprint_shell, keep * sizeof(sif_real));
    }
    free(c->_footprint_block);
    c->_footprint_block = fblock;
    c->footprint = fblock;
    c->footprint_shell = fblock + stride;
  }
  free(c->_block);
  c->_block = block;
  c->cx = block;
  c->cy = block + stride;
  c->cz = block + 2 * stride;
  c->radii = block + 3 * stride;
  c->capacity = cap;
  c->n_voids = keep;
  return SIF_OK;
}

sif_catalogue_t* sif_catalogue_alloc(uint64_t initial_capacity) {
  sif_catalogue_t* c = (sif_catalogue_t*)calloc(1, sizeof(sif_catalogue_t));
  if (!c) {
    error_set(SIF_ERR_ALLOC, "catalogue: cannot allocate");
    return nullptr;
  }
  c->units = SIF_COORDINATES_CARTESIAN;
  if (catalogue_resize(c, initial_capacity) != SIF_OK) {
    free(c);
    return nullptr;
  }
  return c;
}

void sif_catalogue_free(sif_catalogue_t* c) {
  if (!c)
    return;
  free(c->_block);
  free(c->_footprint_block);
  free(c);
}

int sif_catalogue_append(sif_catalogue_t* c, sif_real x, sif_real y, sif_real z, sif_real r) {
  if (!c) {
    error_set(SIF_ERR_INVALID, "catalogue: NULL");
    return SIF_ERR_INVALID;
  }
  if (c->n_voids >= c->capacity) {
    int st = catalogue_resize(c, c->capacity << 1);
    if (st != SIF_OK)
      return st;
  }
  const uint64_t k = c->n_voids++;
  c->cx[k] = x;
  c->cy[k] = y;
  c->cz[k] = z;
  c->radii[k] = r;
  if (c->footprint) {  // catalogue.h:224-236: not measured yet
    c->footprint[k] = SIF_CATALOGUE_FOOTPRINT_UNKNOWN;
    c->footprint_shell[k] = SIF_CATALOGUE_FOOTPRINT_UNKNOWN;
  }
  return SIF_OK;
}

int sif_catalogue_reserve_footprint(sif_catalogue_t* c) {
  if (!c) {
    error_set(SIF_ERR_INVALID, "catalogue: NULL");
    return SIF_ERR_INVALID;
  }
  if (c->footprint)
    return SIF_OK;
  const uint64_t per_line = LINE / sizeof(sif_real);
  const uint64_t stride = (c->capacity + per_line - 1) & ~(per_line - 1);
  sif_real* fblock = (sif_real*)aligned_alloc_bytes(2 * stride * sizeof(sif_real));
  if (!fblock) {
    error_set(SIF_ERR_ALLOC, "catalogue: cannot allocate the footprint");
    return SIF_ERR_ALLOC;
  }
  c->_footprint_block = fblock;
  c->footprint = fblock;
  c->footprint_shell = fblock + stride;
  for (uint64_t i = 0; i < c->n_voids; i++)
    c->footprint[i] = c->footprint_shell[i] = SIF_CATALOGUE_FOOTPRINT_UNKNOWN;
  return SIF_OK;
}

int sif_catalogue_trim(sif_catalogue_t* c) {
  if (!c) {
    error_set(SIF_ERR_INVALID, "catalogue: NULL");
    return SIF_ERR_INVALID;
  }
  if (c->n_voids == c->capacity)
    return SIF_OK;
  return catalogue_resize(c, c->n_voids ? c->n_voids : 1);
}

/* --- grid --- */

sif_grid_t* sif_grid_alloc(uint32_t n_cells, sif_real box_length) {
  if (n_cells == 0) {
    error_set(SIF_ERR_INVALID, "n_cells: 0 (expected > 0)");
    return nullptr;
  }
  if (!(box_length > 0.0f)) {
    error_set(SIF_ERR_INVALID, "box_length: %g (expected > 0)", (double)box_length);
    return nullptr;
  }
  sif_grid_t* g = (sif_grid_t*)calloc(1, sizeof(sif_grid_t));
  if (!g) {
    error_set(SIF_ERR_ALLOC, "grid: cannot allocate");
    return nullptr;
  }
  g->n_cells = n_cells;
  g->total_cells = (uint64_t)n_cells * n_cells * n_cells;
  g->box_length = box_length;
  g->cell_length = box_length / n_cells;
  g->p2_mask = ((n_cells & (n_cells - 1)) == 0) ? n_cells - 1 : 0;
  g->p2_shift = g->p2_mask ? (uint32_t)__builtin_ctz(n_cells) : 0;
  g->content = SIF_GRID_EMPTY;
  g->assignment = SIF_ASSIGNMENT_UNKNOWN;
  g->values = (sif_real*)aligned_alloc_bytes(g->total_cells * sizeof(sif_real));
  if (!g->values) {
    error_set(SIF_ERR_ALLOC, "grid: cannot allocate %u^3 cells", n_cells);
    free(g);
    return nullptr;
  }
  memset(g->values, 0, g->total_cells * sizeof(sif_real));
  return g;
}

void sif_grid_free(sif_grid_t* g) {
  if (!g)
    return;
  free(g->values);
  free(g);
}

int sif_grid_assign(sif_grid_t* g, const sif_field_t* f, sif_assignment_t scheme,
  sif_option opt) {
  if (!g || !g->values) {
    error_set(SIF_ERR_INVALID, "grid: %s", !g ? "NULL" : "no cell values");
    return SIF_ERR_INVALID;
  }
  if (scheme < SIF_ASSIGNMENT_NGP || scheme > SIF_ASSIGNMENT_PCS) {
    error_set(SIF_ERR_INVALID, "scheme: %d (expected NGP, CIC, TSC or PCS)", (int)scheme);
    return SIF_ERR_INVALID;
  }
  if (g->n_cells < 4) {
    error_set(SIF_ERR_INVALID, "n_cells: %u (expected at least 4 to assign onto)", g->n_cells);
    return SIF_ERR_INVALID;
  }
  if (scheme != SIF_ASSIGNMENT_CIC || (opt & SIF_GRID_INTERLACE)) {
    error_set(SIF_ERR_UNSUPPORTED, "assignment: only CIC without interlacing in this build");
    return SIF_ERR_UNSUPPORTED;
  }
  int st = field_check(f, "field");
  if (st == SIF_OK)
    st = ensure_init();
  if (st != SIF_OK)
    return st;
  Dev dx, dy, dz, dw, dv, acc;
  st = upload_field(f, g->box_length, dx, dy, dz, dw);
  if (st == SIF_OK)
    st = dv.alloc(g->total_cells * sizeof(float));
  if (st == SIF_OK)
    st = acc.alloc(g->total_cells * sizeof(int64_t));
  if (st == SIF_OK)
    st = cic_deposit(dx.as<float>(), dy.as<float>(), dz.as<float>(),
      f->weights ? dw.as<float>() : nullptr, f->n_particles, g->n_cells, g->box_length,
      dv.as<float>(), acc.as<int64_t>());
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(g->values, dv.p, g->total_cells * sizeof(float),
    hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  g->content = SIF_GRID_DENSITY;
  g->assignment = SIF_ASSIGNMENT_CIC;
  g->interlaced = 0;
  return SIF_OK;
}

int sif_grid_assign_cic(sif_grid_t* g, const sif_field_t* f) {
  return sif_grid_assign(g, f, SIF_ASSIGNMENT_CIC, 0);
}

int sif_grid_to_density_contrast(sif_grid_t* g) {
  if (!g || !g->values) {
    error_set(SIF_ERR_INVALID, "grid: %s", !g ? "NULL" : "no cell values");
    return SIF_ERR_INVALID;
  }
  if (g->content == SIF_GRID_DENSITY_CONTRAST) {
    error_set(SIF_ERR_INVALID,
      "grid: already a density contrast (converting again would divide by a mean of "
      "about 0)");
    return SIF_ERR_INVALID;
  }
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  Dev dv;
  st = dv.upload(g->values, g->total_cells * sizeof(float));
  if (st != SIF_OK)
    return st;
  double mean = 0.0;
  st = grid_mean(dv.as<float>(), g->total_cells, &mean);
  if (st != SIF_OK)
    return st;
  if (!(mean > 0.0)) {
    error_set(SIF_ERR_RANGE, "grid: mean density %g (expected > 0)%s", mean,
      mean == 0.0  ? "; nothing has been deposited onto it"
      : mean < 0.0 ? "; negative weights outweigh the positive ones"
                   : "; some cell values are not finite");
    return SIF_ERR_RANGE;
  }
  st = density_contrast(dv.as<float>(), g->total_cells, mean);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(g->values, dv.p, g->total_cells * sizeof(float),
    hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  g->content = SIF_GRID_DENSITY_CONTRAST;
  return SIF_OK;
}

/* --- chain mesh --- */

void sif_chain_mesh_free(sif_chain_mesh_t* m) {
  if (!m)
    return;
  free(m->_position_block);
  free(m->_velocity_block);
  free(m->weights);
  free(m->cell_weights);
  free(m->original_indices);
  free(m->cell_offsets);
  free(m);
}

static sif_chain_mesh_t* mesh_build_host(uint32_t n_cells, sif_real box,
  const sif_field_t* f, sif_option opt) {
  int st = field_check(f, "field");
  if (st == SIF_OK && n_cells == 0) {
    error_set(SIF_ERR_INVALID, "n_cells: 0 (expected > 0)");
    st = SIF_ERR_INVALID;
  }
  if (st == SIF_OK && n_cells > 1024) {
    error_set(SIF_ERR_UNSUPPORTED, "n_cells: %u (this build builds meshes up to 1024)", n_cells);
    st = SIF_ERR_UNSUPPORTED;
  }
  if (st == SIF_OK && !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "box_length: %g (expected > 0)", (double)box);
    st = SIF_ERR_INVALID;
  }
  if (st == SIF_OK)
    st = ensure_init();
  if (st != SIF_OK)
    return nullptr;
  const uint64_t n = f->n_particles;
  const uint64_t cells = (uint64_t)n_cells * n_cells * n_cells;
  Dev dx, dy, dz, dw;
  st = upload_field(f, box, dx, dy, dz, dw);
  if (st != SIF_OK)
    return nullptr;
  Dev doff, dmx, dmy, dmz, dmw, dorig, dcw;
  st = doff.alloc((cells + 1) * 8);
  if (st == SIF_OK) st = dmx.alloc(n * 4);
  if (st == SIF_OK) st = dmy.alloc(n * 4);
  if (st == SIF_OK) st = dmz.alloc(n * 4);
  if (st == SIF_OK) st = dorig.alloc(n * 8);
  if (st == SIF_OK && f->weights) st = dmw.alloc(n * 4);
  if (st == SIF_OK && f->weights) st = dcw.alloc(cells * 4);
  if (st != SIF_OK)
    return nullptr;
  double total_weight = 0.0;
  st = mesh_build(dx.as<float>(), dy.as<float>(), dz.as<float>(),
    f->weights ? dw.as<float>() : nullptr, n, n_cells, box, doff.as<uint64_t>(),
    dmx.as<float>(), dmy.as<float>(), dmz.as<float>(), dmw.as<float>(), dorig.as<uint64_t>(),
    dcw.as<float>(), &total_weight);
  if (st != SIF_OK)
    return nullptr;
  sif_chain_mesh_t* m = (sif_chain_mesh_t*)calloc(1, sizeof(sif_chain_mesh_t));
  if (!m) {
    error_set(SIF_ERR_ALLOC, "chain mesh: cannot allocate");
    return nullptr;
  }
  m->n_cells = n_cells;
  m->total_cells = cells;
  m->box_length = box;
  m->cell_length = box / (sif_real)n_cells;
  m->n_particles = n;
  m->total_weight = total_weight;
  const uint64_t pn = sif_field_padded_n(n);
  m->_position_block = (sif_real*)aligned_alloc_bytes(3 * pn * sizeof(sif_real));
  m->cell_offsets = (uint64_t*)aligned_alloc_bytes((cells + 1) * sizeof(uint64_t));
  m->original_indices = (uint64_t*)aligned_alloc_bytes(n * sizeof(uint64_t));
  if (f->weights) {
    m->weights = (sif_real*)aligned_alloc_bytes(n * sizeof(sif_real));
    m->cell_weights = (sif_real*)aligned_alloc_bytes(cells * sizeof(sif_real));
  }
  if (!m->_position_block || !m->cell_offsets || !m->original_indices ||
      (f->weights && (!m->weights || !m->cell_weights))) {
    error_set(SIF_ERR_ALLOC, "chain mesh: cannot allocate its arrays");
    sif_chain_mesh_free(m);
    return nullptr;
  }
  m->x = m->_position_block;
  m->y = m->_position_block + pn;
  m->z = m->_position_block + 2 * pn;
  hipStream_t s = stream();
  bool ok = hipMemcpyAsync(m->x, dmx.p, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(m->y, dmy.p, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(m->z, dmz.p, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipMemcpyAsync(m->cell_offsets, doff.p, (cells + 1) * 8, hipMemcpyDeviceToHost,
              s) == hipSuccess &&
            hipMemcpyAsync(m->original_indices, dorig.p, n * 8, hipMemcpyDeviceToHost, s) ==
              hipSuccess;
  if (ok && f->weights)
    ok = hipMemcpyAsync(m->weights, dmw.p, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(m->cell_weights, dcw.p, cells * 4, hipMemcpyDeviceToHost, s) ==
           hipSuccess;
  ok = ok && hipStreamSynchronize(s) == hipSuccess;
  if (!ok) {
    error_set(SIF_ERR_INVALID, "chain mesh: device to host copy failed");
    sif_chain_mesh_free(m);
    return nullptr;
  }
  if (f->vx) {
    m->_velocity_block = (sif_real*)aligned_alloc_bytes(3 * pn * sizeof(sif_real));
    if (!m->_velocity_block) {
      error_set(SIF_ERR_ALLOC, "chain mesh: cannot allocate velocities");
      sif_chain_mesh_free(m);
      return nullptr;
    }
    m->vx = m->_velocity_block;
    m->vy = m->_velocity_block + pn;
    m->vz = m->_velocity_block + 2 * pn;
    for (uint64_t p = 0; p < n; p++) {
      const uint64_t i = m->original_indices[p];
      m->vx[p] = f->vx[i];
      m->vy[p] = f->vy[i];
      m->vz[p] = f->vz[i];
    }
  }
  if (opt & SIF_MESH_DROP_INDICES) {
    free(m->original_indices);
    m->original_indices = nullptr;
  }
  return m;
}

sif_chain_mesh_t* sif_chain_mesh_alloc(uint32_t n_cells, sif_real box_length,
  const sif_field_t* field, sif_option opt) {
  return mesh_build_host(n_cells, box_length, field, opt);
}

sif_chain_mesh_t* sif_chain_mesh_alloc_consume(uint32_t n_cells, sif_real box_length,
  sif_field_t* field, sif_option opt) {
  sif_chain_mesh_t* m = mesh_build_host(n_cells, box_length, field, opt);
  if (!m)
    return nullptr;
  // chain_mesh.c:270-290: the field gives its storage up and is left empty.
  free(field->_position_block);
  free(field->_velocity_block);
  free(field->weights);
  free(field->original_indices);
  field->_position_block = field->_velocity_block = nullptr;
  field->x = field->y = field->z = nullptr;
  field->vx = field->vy = field->vz = nullptr;
  field->weights = nullptr;
  field->original_indices = nullptr;
  field->n_particles = 0;
  field->state_flags = 0;
  return m;
}

// Nearest particle by a shell walk over the mesh cells: exact, and the first
// of equal distances in mesh order (chain_mesh.c:966-1068).
static uint64_t nearest_slot(const sif_chain_mesh_t* m, sif_real px, sif_real py,
  sif_real pz, int periodic) {
  if (!m || m->n_particles == 0)
    return UINT64_MAX;
  const int32_t n = (int32_t)m->n_cells;
  const float l = m->cell_length, L = m->box_length;
  auto d2 = [&](uint64_t p) {
    float dx = m->x[p] - px, dy = m->y[p] - py, dz = m->z[p] - pz;
    if (periodic) {
      dx = fabsf(dx);
      dy = fabsf(dy);
      dz = fabsf(dz);
      dx = dx > 0.5f * L ? L - dx : dx;
      dy = dy > 0.5f * L ? L - dy : dy;
      dz = dz > 0.5f * L ? L - dz : dz;
    }
    return dx * dx + dy * dy + dz * dz;
  };
  const int32_t cx = std::min(std::max((int32_t)floorf(px / l), 0), n - 1);
  const int32_t cy = std::min(std::max((int32_t)floorf(py / l), 0), n - 1);
  const int32_t cz = std::min(std::max((int32_t)floorf(pz / l), 0), n - 1);
  float best = INFINITY;
  uint64_t best_p = UINT64_MAX;
  const int32_t max_r = periodic ? n / 2 + 1 : n;
  for (int32_t R = 0; R <= max_r; R++) {
    for (int32_t i = -R; i <= R; i++)
      for (int32_t j = -R; j <= R; j++)
        for (int32_t k = -R; k <= R; k++) {
          if (std::max(std::max(abs(i), abs(j)), abs(k)) != R)
            continue;
          int32_t a = cx + i, b = cy + j, c = cz + k;
          if (periodic) {
            a = (a % n + n) % n;
            b = (b % n + n) % n;
            c = (c % n + n) % n;
          } else if (a < 0 || a >= n || b < 0 || b >= n || c < 0 || c >= n) {
            continue;
          }
          const uint64_t cell = ((uint64_t)a * n + b) * n + c;
          for (uint64_t p = m->cell_offsets[cell]; p < m->cell_offsets[cell + 1]; p++) {
            const float d = d2(p);
            if (d < best) {
              best = d;
              best_p = p;
            }
          }
        }
    // every point of a cell beyond shell R is at least R cells from the
    // query's own cell boundary
    if (best_p != UINT64_MAX && sqrtf(best) <= (float)R * l)
      break;
  }
  return best_p;
}

uint64_t sif_chain_mesh_find_nearest_open(const sif_chain_mesh_t* m, sif_real px, sif_real py,
  sif_real pz) {
  if (!m || !m->original_indices) {
    error_set(SIF_ERR_INVALID, "find_nearest_open: mesh: %s",
      !m ? "NULL" : "built without its index map (SIF_MESH_DROP_INDICES)");
    return UINT64_MAX;
  }
  const uint64_t s = nearest_slot(m, px, py, pz, 0);
  return s == UINT64_MAX ? UINT64_MAX : m->original_indices[s];
}

uint64_t sif_chain_mesh_find_nearest_pbc(const sif_chain_mesh_t* m, sif_real px, sif_real py,
  sif_real pz) {
  if (!m || !m->original_indices) {
    error_set(SIF_ERR_INVALID, "find_nearest_pbc: mesh: %s",
      !m ? "NULL" : "built without its index map (SIF_MESH_DROP_INDICES)");
    return UINT64_MAX;
  }
  const uint64_t s = nearest_slot(m, px, py, pz, 1);
  return s == UINT64_MAX ? UINT64_MAX : m->original_indices[s];
}

/* --- finder --- */

static const struct {
  const char* name;
  sif_real overlap_fraction;
  sif_option options;
} VARIANTS[SIF__EXODUS_N_VARIANTS] = {
  {"excursion-set", 0.0f, SIF_FINDER_SEEDS_CELLS},
  {"excursion-peak", 1.0f,
    SIF_FINDER_SEEDS_MINIMA | SIF_FINDER_EXTEND_REACH | SIF_FINDER_DISPLACE},
  {"lagrangian-overlap", 0.0f, SIF_FINDER_SEEDS_CELLS | SIF_FINDER_LAGRANGIAN_OVERLAP},
};

int sif_finder_exodus_preset(sif_exodus_variant_t v, sif_real* f, sif_option* opt) {
  if ((unsigned)v >= (unsigned)SIF__EXODUS_N_VARIANTS) {
    error_set(SIF_ERR_INVALID, "variant: %d (expected 0 to %d)", (int)v,
      SIF__EXODUS_N_VARIANTS - 1);
    return SIF_ERR_INVALID;
  }
  if (f)
    *f = VARIANTS[v].overlap_fraction;
  if (opt)
    *opt = VARIANTS[v].options;
  return SIF_OK;
}

const char* sif_finder_exodus_variant_name(sif_exodus_variant_t v) {
  return (unsigned)v < (unsigned)SIF__EXODUS_N_VARIANTS ? VARIANTS[v].name : nullptr;
}

int sif_finder_exodus_variant_from_name(const char* name, sif_exodus_variant_t* v) {
  for (int k = 0; name && k < SIF__EXODUS_N_VARIANTS; k++)
    if (strcmp(name, VARIANTS[k].name) == 0) {
      if (v)
        *v = (sif_exodus_variant_t)k;
      return SIF_OK;
    }
  error_set(SIF_ERR_INVALID,
    "variant: \"%s\" (expected excursion-set, excursion-peak or lagrangian-overlap)",
    name ? name : "(null)");
  return SIF_ERR_INVALID;
}

sif_catalogue_t* sif_finder_exodus(sif_grid_t* grid, const sif_chain_mesh_t* mesh,
  const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_exodus_variant_t variant,
  const sif_finder_exodus_extras_t* extras) {
  sif_real f;
  sif_option opt;
  if (sif_finder_exodus_preset(variant, &f, &opt) != SIF_OK)
    return nullptr;
  return sif_finder_exodus_raw(grid, mesh, radii, n_radii, threshold, f, opt, extras);
}

// A run's checkpoint and where it stands (exodus_finder.c:1384-1393).
struct Checkpointing {
  bool active = false;    // written after every rung
  bool on_disk = false;   // this run wrote one, or resumed one
  bool progress = false;  // the one on disk holds a finished rung
  bool keep = false;      // left in place on success, for the caller
  char path[4096] = {0};
  sif_finder_exodus_checkpoint_t header;
};

}  // extern "C"

namespace {

// What a survey run adds to the drop-in (exodus_survey.c:386-535).
struct SurveyIn {
  sif_grid_t* random_grid;
  const sif_chain_mesh_t* random_mesh;
  const sif_real* offset;  // NULL: (0, 0, 0)
};

// exodus_finder.c:1145-1191 (sif__exodus_check_extras): a prior catalogue
// the finder can place -- Cartesian, radii positive and finite, centres
// (moved by the offset) inside [0, box).
int check_extras(const sif_finder_exodus_extras_t* extras, const sif_real* offset, float box) {
  const sif_catalogue_t* prior = extras ? extras->prior : nullptr;
  if (!prior)
    return SIF_OK;
  if (prior->units != SIF_COORDINATES_CARTESIAN) {
    error_set(SIF_ERR_INVALID, "prior: centres on the sky (expected Cartesian coordinates)");
    return SIF_ERR_INVALID;
  }
  for (uint64_t i = 0; i < prior->n_voids; i++) {
    const sif_real r = prior->radii[i];
    if (!(r > 0.0f) || !std::isfinite(r)) {
      error_set(SIF_ERR_INVALID, "prior: void %llu: radius %g (expected positive and finite)",
        (unsigned long long)i, (double)r);
      return SIF_ERR_INVALID;
    }
    const sif_real c[3] = {prior->cx[i] + (offset ? offset[0] : 0.0f),
      prior->cy[i] + (offset ? offset[1] : 0.0f), prior->cz[i] + (offset ? offset[2] : 0.0f)};
    for (int k = 0; k < 3; k++)
      if (!(c[k] >= 0.0f && c[k] < box)) {
        error_set(SIF_ERR_INVALID, "prior: void %llu: centre (%g, %g, %g)%s outside the box "
          "[0, %g)", (unsigned long long)i, (double)prior->cx[i], (double)prior->cy[i],
          (double)prior->cz[i], offset ? ", moved by the offset," : "", (double)box);
        return SIF_ERR_INVALID;
      }
  }
  return SIF_OK;
}

// sif__exodus_run's drop-in half (exodus_finder.c:1435-1575, 2430-2492):
// where the run starts (a checkpoint resumed, or the prior voids), the
// upload of the host grid(s) and mesh(es), the run on the device, the
// checkpoint after every rung, and the grid handed back. `sv` set: a survey.
sif_catalogue_t* exodus_dropin(sif_grid_t* grid, const sif_chain_mesh_t* mesh,
  const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_real overlap_fraction,
  sif_option options, const sif_finder_exodus_extras_t* extras, const SurveyIn* sv) {
  const char* named = extras ? extras->checkpoint : nullptr;
  if (named && (options & SIF_FINDER_NO_CHECKPOINT)) {
    error_set(SIF_ERR_INVALID, "checkpoint: %s, with SIF_FINDER_NO_CHECKPOINT set", named);
    return nullptr;
  }
  if (named && strlen(named) + 32 >= 4096) {
    error_set(SIF_ERR_INVALID, "checkpoint: path of %zu characters (expected under %d)",
      strlen(named), 4096 - 32);
    return nullptr;
  }
  int st = need_init();
  if (st != SIF_OK)
    return nullptr;
  const sif_real* offset = sv ? sv->offset : nullptr;

  // exodus_finder.c:1435-1575: where the run starts. A checkpoint the caller
  // named and that exists is resumed once its header describes this run;
  // otherwise the run starts from the prior voids (moved into the box by the
  // offset) and writes its first checkpoint now (only a named one is worth
  // failing over). Checkpointed voids are in the box's frame.
  std::vector<float> sorted;
  sorted_ladder(radii, n_radii, sorted);
  const sif_catalogue_t* prior = extras ? extras->prior : nullptr;
  sif_catalogue_t* in_box = nullptr;  // the prior voids, moved by the offset
  if (prior && offset) {
    in_box = sif_catalogue_alloc(prior->n_voids ? prior->n_voids : 1);
    if (!in_box) {
      error_set(SIF_ERR_ALLOC, "prior: cannot allocate %llu voids",
        (unsigned long long)prior->n_voids);
      return nullptr;
    }
    for (uint64_t i = 0; i < prior->n_voids; i++)
      sif_catalogue_append(in_box, prior->cx[i] + offset[0], prior->cy[i] + offset[1],
        prior->cz[i] + offset[2], prior->radii[i]);
  }
  const sif_catalogue_t* start = in_box ? in_box : prior;
  Checkpointing ck;
  ck.keep = extras && extras->keep_checkpoint;
  sif_catalogue_t* resumed = nullptr;
  uint32_t first_rung = 0;
  uint64_t mismatched0 = 0;
  auto fail = [&]() -> sif_catalogue_t* {
    sif_catalogue_free(resumed);
    sif_catalogue_free(in_box);
    return nullptr;
  };
  if (!(options & SIF_FINDER_NO_CHECKPOINT)) {
    sif_finder_exodus_checkpoint_t& h = ck.header;
    memset(&h, 0, sizeof(h));
    h.is_survey = sv ? 1u : 0u;
    h.is_weighted = mesh->weights != nullptr;
    h.options = options & (SIF__FINDER_SEARCH_MASK | SIF_FINDER_KEEP_CIC_WINDOW |
                            SIF__FINDER_SEEDS_MASK | SIF__FINDER_REFINE_MASK |
                            SIF_FINDER_EXTEND_REACH | SIF_FINDER_DISPLACE |
                            SIF_FINDER_ALIASING_WINDOW | SIF_FINDER_LAGRANGIAN_OVERLAP);
    h.n_radii = n_radii;
    h.grid_n_cells = grid->n_cells;
    h.assignment = ((uint32_t)grid->assignment & 0xffu) | ((grid->interlaced ? 1u : 0u) << 8);
    h.box_length = (double)grid->box_length;
    h.threshold = (double)threshold;
    h.overlap_fraction = (double)overlap_fraction;
    h.n_particles = mesh->n_particles;
    h.total_weight = mesh->total_weight;
    if (sv) {  // exodus_finder.c:1478-1484
      h.random_n_particles = sv->random_mesh->n_particles;
      h.random_total_weight = sv->random_mesh->total_weight;
      for (int k = 0; offset && k < 3; k++)
        h.offset[k] = (double)offset[k];
    }
    if (prior && prior->n_voids > 0) {
      h.n_prior = prior->n_voids;
      h.prior_crc32 = catalogue_crc(prior->cx, prior->cy, prior->cz, prior->radii,
        prior->n_voids);
    }
    if (named)
      snprintf(ck.path, sizeof(ck.path), "%s", named);
    else
      checkpoint_name(ck.path, sizeof(ck.path));
    if (named && access(named, F_OK) == 0) {
      sif_finder_exodus_checkpoint_t got;
      std::vector<float> got_radii, cols[4];
      st = checkpoint_load(named, &got, &got_radii, cols);
      if (st == SIF_OK)
        st = checkpoint_match(named, h, sorted.data(), got, got_radii.data());
      if (st == SIF_OK) {
        resumed = sif_catalogue_alloc(got.n_voids ? got.n_voids : 1);
        if (!resumed) {
          error_set(SIF_ERR_ALLOC, "checkpoint: %s: cannot allocate its voids", named);
          st = SIF_ERR_ALLOC;
        }
        for (uint64_t v = 0; st == SIF_OK && v < got.n_voids; v++)
          st = sif_catalogue_append(resumed, cols[0][v], cols[1][v], cols[2][v], cols[3][v]);
      }
      if (st != SIF_OK)
        return fail();
      first_rung = got.next_rung;
      mismatched0 = got.n_mismatched;
      h.next_rung = got.next_rung;
      h.n_mismatched = got.n_mismatched;
      ck.active = ck.on_disk = ck.progress = true;
    } else {
      const uint64_t np = start ? start->n_voids : 0;
      Columns v{{start ? start->cx : nullptr, start ? start->cy : nullptr,
                  start ? start->cz : nullptr, start ? start->radii : nullptr}, np};
      st = checkpoint_write(ck.path, &h, sorted.data(), v);
      if (st == SIF_OK) {
        ck.active = ck.on_disk = true;
      } else if (named) {
        return fail();
      } else {
        sif_error_clear();  // running without a checkpoint
        st = SIF_OK;
      }
    }
  }

  const uint64_t cells = grid->total_cells;
  const uint64_t mcells = mesh->total_cells;
  const uint64_t n = mesh->n_particles;
  Dev dv, doff, dmx, dmy, dmz, dmw, dcw;
  Dev rv, roff, rmx, rmy, rmz, rmw, rcw;  // a survey's randoms
  st = dv.upload(grid->values, cells * 4);
  if (st == SIF_OK) st = doff.upload(mesh->cell_offsets, (mcells + 1) * 8);
  if (st == SIF_OK) st = dmx.upload(mesh->x, n * 4);
  if (st == SIF_OK) st = dmy.upload(mesh->y, n * 4);
  if (st == SIF_OK) st = dmz.upload(mesh->z, n * 4);
  if (st == SIF_OK && mesh->weights) st = dmw.upload(mesh->weights, n * 4);
  if (st == SIF_OK && mesh->weights) st = dcw.upload(mesh->cell_weights, mcells * 4);
  if (sv) {
    const sif_chain_mesh_t* rm = sv->random_mesh;
    const uint64_t rn = rm->n_particles, rcells = rm->total_cells;
    if (st == SIF_OK) st = rv.upload(sv->random_grid->values, cells * 4);
    if (st == SIF_OK) st = roff.upload(rm->cell_offsets, (rcells + 1) * 8);
    if (st == SIF_OK) st = rmx.upload(rm->x, rn * 4);
    if (st == SIF_OK) st = rmy.upload(rm->y, rn * 4);
    if (st == SIF_OK) st = rmz.upload(rm->z, rn * 4);
    if (st == SIF_OK && rm->weights) st = rmw.upload(rm->weights, rn * 4);
    if (st == SIF_OK && rm->weights) st = rcw.upload(rm->cell_weights, rcells * 4);
  }
  RunSpec R;
  R.d_delta = dv.as<float>();
  R.n = grid->n_cells;
  R.box = grid->box_length;
  R.d_offsets = doff.as<uint64_t>();
  R.d_mx = dmx.as<float>();
  R.d_my = dmy.as<float>();
  R.d_mz = dmz.as<float>();
  if (mesh->weights) {
    R.d_mw = dmw.as<float>();
    R.d_cw = dcw.as<float>();
  }
  R.total_weight = mesh->weights ? mesh->total_weight : (double)n;
  R.mesh_n = mesh->n_cells;
  R.n_particles = n;
  sorted_ladder(radii, n_radii, R.radii);
  R.threshold = threshold;
  R.f = overlap_fraction;
  R.opt = options;
  R.restore = !(options & SIF_FINDER_CONSUME_GRID);
  R.prior = resumed ? resumed : start;
  R.first_rung = first_rung;
  R.mismatched0 = mismatched0;
  if (st == SIF_OK && sv) {
    const sif_chain_mesh_t* rm = sv->random_mesh;
    sif_amd_mesh_t dm{doff.as<uint64_t>(), dmx.as<float>(), dmy.as<float>(), dmz.as<float>(),
      mesh->weights ? dmw.as<float>() : nullptr, mesh->weights ? dcw.as<float>() : nullptr,
      mesh->weights ? mesh->total_weight : (double)n, mesh->n_cells, n};
    sif_amd_mesh_t rmd{roff.as<uint64_t>(), rmx.as<float>(), rmy.as<float>(), rmz.as<float>(),
      rm->weights ? rmw.as<float>() : nullptr, rm->weights ? rcw.as<float>() : nullptr,
      rm->weights ? rm->total_weight : (double)rm->n_particles, rm->n_cells,
      rm->n_particles};
    st = survey_prepare(R, dm, rmd, rv.as<float>(), radii, n_radii, options);
  }
  if (st != SIF_OK) {
    if (ck.on_disk && !ck.progress)
      unlink(ck.path);
    return fail();
  }
  if (ck.active)
    R.on_rung = [&](uint32_t done, const float* x, const float* y, const float* z,
                  const float* r, uint64_t nv, uint64_t mism) {
      // a failed write leaves the previous checkpoint standing; the next rung
      // tries again
      ck.header.next_rung = done;
      ck.header.n_mismatched = mism;
      if (checkpoint_write(ck.path, &ck.header, R.radii.data(), Columns{{x, y, z, r}, nv}) ==
          SIF_OK)
        ck.progress = true;
      else
        sif_error_clear();
    };
  sif_catalogue_t* cat = nullptr;
  st = exodus_run(R, &cat, nullptr, nullptr);
  sif_catalogue_free(resumed);
  resumed = nullptr;
  // exodus_finder.c:1560-1575: a finished run has no use for its checkpoint
  // (unless the caller keeps it); a failed one keeps it once it holds a rung
  if (ck.on_disk && !(st != SIF_OK && ck.progress) && !(st == SIF_OK && ck.keep))
    unlink(ck.path);
  // the grid comes back restored (or holding the last smoothing under
  // SIF_FINDER_CONSUME_GRID), on failure too; a survey's random grid is only
  // ever read
  if (hipMemcpyAsync(grid->values, dv.p, cells * 4, hipMemcpyDeviceToHost, stream()) !=
        hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess) {
    if (st == SIF_OK)
      error_set(SIF_ERR_INVALID, "exodus: the grid could not be copied back");
    sif_catalogue_free(cat);
    return fail();
  }
  if (st != SIF_OK) {
    sif_catalogue_free(cat);
    return fail();
  }
  sif_catalogue_trim(cat);
  // exodus_survey.c:538-560: back to the caller's frame, the prior voids as
  // the caller gave them (x + o - o need not be x)
  if (offset) {
    const sif_real back[3] = {-offset[0], -offset[1], -offset[2]};
    sif_catalogue_translate(cat, back);
    if (prior && prior->n_voids > 0) {
      const size_t bytes = (size_t)prior->n_voids * sizeof(sif_real);
      memcpy(cat->cx, prior->cx, bytes);
      memcpy(cat->cy, prior->cy, bytes);
      memcpy(cat->cz, prior->cz, bytes);
    }
  }
  sif_catalogue_free(in_box);
  return cat;
}

// exodus_survey.c:321-342 on a host mesh, and the padding check of 367-385
int survey_check_margin_host(const sif_chain_mesh_t* m, const char* what, float margin,
  const char* hint) {
  float lo[3] = {m->box_length, m->box_length, m->box_length}, hi[3] = {0.0f, 0.0f, 0.0f};
  for (uint64_t p = 0; p < m->n_particles; p++) {
    lo[0] = fminf(lo[0], m->x[p]);
    lo[1] = fminf(lo[1], m->y[p]);
    lo[2] = fminf(lo[2], m->z[p]);
    hi[0] = fmaxf(hi[0], m->x[p]);
    hi[1] = fmaxf(hi[1], m->y[p]);
    hi[2] = fmaxf(hi[2], m->z[p]);
  }
  const float box = m->box_length;
  for (int k = 0; k < 3; k++)
    if (lo[k] < margin || hi[k] > box - margin) {
      error_set(SIF_ERR_INVALID,
        "%s: span [%g, %g] on axis %d of a box of %g, %g from its face (needs %g of padding "
        "on every side%s)", what, (double)lo[k], (double)hi[k], k, (double)box,
        (double)fminf(lo[k], box - hi[k]), (double)margin, hint);
      return SIF_ERR_INVALID;
    }
  return SIF_OK;
}

}  // namespace

extern "C" {

sif_catalogue_t* sif_finder_exodus_raw(sif_grid_t* grid, const sif_chain_mesh_t* mesh,
  const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_real overlap_fraction,
  sif_option options, const sif_finder_exodus_extras_t* extras) {
  if (!grid || !grid->values || !mesh) {
    error_set(SIF_ERR_INVALID, "%s",
      !grid ? "grid: NULL" : !grid->values ? "grid: no values" : "mesh: NULL");
    return nullptr;
  }
  int st = check_args(radii, n_radii, threshold, overlap_fraction);
  if (st != SIF_OK)
    return nullptr;
  if (!mesh->x || mesh->n_particles == 0) {
    error_set(SIF_ERR_INVALID, "mesh: no particles");
    return nullptr;
  }
  if (mesh->box_length != grid->box_length) {
    error_set(SIF_ERR_INVALID, "mesh: box of %g, grid: box of %g (must match)",
      (double)mesh->box_length, (double)grid->box_length);
    return nullptr;
  }
  if (mesh->weights &&
      check_weights_host(mesh->weights, mesh->n_particles, mesh->cell_weights,
        mesh->total_weight) != SIF_OK)
    return nullptr;
  if ((st = check_options(options)) != SIF_OK)
    return nullptr;
  if (check_extras(extras, nullptr, grid->box_length) != SIF_OK)
    return nullptr;
  return exodus_dropin(grid, mesh, radii, n_radii, threshold, overlap_fraction, options,
    extras, nullptr);
}

// exodus_survey.c:386-535, in the reference's order of checks
sif_catalogue_t* sif_finder_exodus_survey_raw(sif_grid_t* data_grid, sif_grid_t* random_grid,
  const sif_chain_mesh_t* data_mesh, const sif_chain_mesh_t* random_mesh,
  const sif_real offset[3], const sif_real* radii, uint32_t n_radii, sif_real threshold,
  sif_real overlap_fraction, sif_option options, const sif_finder_exodus_extras_t* extras) {
  const char* missing = !data_grid                          ? "data grid: NULL"
                        : !data_grid->values                ? "data grid: no values"
                        : !random_grid                      ? "random grid: NULL"
                        : !random_grid->values              ? "random grid: no values"
                        : !data_mesh                        ? "data mesh: NULL"
                        : !random_mesh                      ? "random mesh: NULL"
                        : !data_mesh->x || data_mesh->n_particles == 0 ? "data mesh: no particles"
                        : !random_mesh->x || random_mesh->n_particles == 0
                          ? "random mesh: no particles"
                          : nullptr;
  if (missing) {
    error_set(SIF_ERR_INVALID, "%s", missing);
    return nullptr;
  }
  if (check_args(radii, n_radii, threshold, overlap_fraction) != SIF_OK ||
      survey_check_options(options) != SIF_OK)
    return nullptr;
  for (int k = 0; offset && k < 3; k++)
    if (!std::isfinite(offset[k])) {
      error_set(SIF_ERR_INVALID, "offset[%d]: %g, expected finite", k, (double)offset[k]);
      return nullptr;
    }
  // the contrast is formed against the randoms: a grid already turned into
  // one against the box mean is the wrong input
  if (data_grid->content == SIF_GRID_DENSITY_CONTRAST ||
      random_grid->content == SIF_GRID_DENSITY_CONTRAST) {
    error_set(SIF_ERR_INVALID, "%s grid: holds a density contrast; a survey takes the CIC "
      "densities as assign_cic leaves them",
      data_grid->content == SIF_GRID_DENSITY_CONTRAST ? "data" : "random");
    return nullptr;
  }
  if (data_grid->n_cells != random_grid->n_cells ||
      data_grid->box_length != random_grid->box_length) {
    error_set(SIF_ERR_INVALID, "data grid: %u cells over %g, random grid: %u over %g (must "
      "match)", data_grid->n_cells, (double)data_grid->box_length, random_grid->n_cells,
      (double)random_grid->box_length);
    return nullptr;
  }
  if (data_mesh->box_length != data_grid->box_length ||
      random_mesh->box_length != data_grid->box_length) {
    error_set(SIF_ERR_INVALID, "meshes: boxes of %g (data) and %g (randoms), grids: %g (must "
      "match)", (double)data_mesh->box_length, (double)random_mesh->box_length,
      (double)data_grid->box_length);
    return nullptr;
  }
  if ((data_mesh->weights && check_weights_host(data_mesh->weights, data_mesh->n_particles,
                               data_mesh->cell_weights, data_mesh->total_weight) != SIF_OK) ||
      (random_mesh->weights &&
        check_weights_host(random_mesh->weights, random_mesh->n_particles,
          random_mesh->cell_weights, random_mesh->total_weight) != SIF_OK))
    return nullptr;
  if (!(data_mesh->total_weight > 0.0) || !(random_mesh->total_weight > 0.0)) {
    error_set(SIF_ERR_INVALID, "%s mesh: total weight is zero",
      data_mesh->total_weight > 0.0 ? "random" : "data");
    return nullptr;
  }
  const float cell = data_grid->box_length / (float)data_grid->n_cells;
  float r_max = radii[0], r_min = radii[0];
  for (uint32_t i = 1; i < n_radii; i++) {
    r_max = radii[i] > r_max ? radii[i] : r_max;
    r_min = radii[i] < r_min ? radii[i] : r_min;
  }
  const float sf = sif__finder_search_factor(options);
  const float r_search = r_max + fmaxf((sf - 1.0f) * r_max, 3.0f * cell);
  const float mesh_cell = fmaxf(data_mesh->box_length / (float)data_mesh->n_cells,
    random_mesh->box_length / (float)random_mesh->n_cells);
  const float margin = fmaxf(r_search, 0.5f * (r_search + mesh_cell)) + cell;
  char hint[192] = ", as survey_box sizes it";
  if (cell > r_min)
    snprintf(hint, sizeof(hint),
      "; survey_box allows for grid cells up to the smallest radius, %g, and these are %g",
      (double)r_min, (double)cell);
  if (survey_check_margin_host(random_mesh, "randoms", margin, hint) != SIF_OK ||
      survey_check_margin_host(data_mesh, "data", margin, hint) != SIF_OK)
    return nullptr;
  if (check_extras(extras, offset, data_grid->box_length) != SIF_OK)
    return nullptr;
  SurveyIn sv{random_grid, random_mesh, offset};
  return exodus_dropin(data_grid, data_mesh, radii, n_radii, threshold, overlap_fraction,
    options, extras, &sv);
}

sif_catalogue_t* sif_finder_exodus_survey(sif_grid_t* data_grid, sif_grid_t* random_grid,
  const sif_chain_mesh_t* data_mesh, const sif_chain_mesh_t* random_mesh,
  const sif_real offset[3], const sif_real* radii, uint32_t n_radii, sif_real threshold,
  sif_exodus_variant_t variant, const sif_finder_exodus_extras_t* extras) {
  sif_real f;
  sif_option opt;
  if (sif_finder_exodus_preset(variant, &f, &opt) != SIF_OK) {
    error_set(SIF_ERR_INVALID, "variant: %d", (int)variant);
    return nullptr;
  }
  return sif_finder_exodus_survey_raw(data_grid, random_grid, data_mesh, random_mesh, offset,
    radii, n_radii, threshold, f, opt, extras);
}

// exodus_survey.c:602-703: a box just large enough for the survey, centred,
// with the padding the finder asks for at grid cells up to the smallest radius.
int sif_finder_exodus_survey_box(const sif_field_t* data, const sif_field_t* randoms,
  const sif_real* radii, uint32_t n_radii, sif_option opt, sif_real offset[3],
  sif_real* box_length) {
  const char* names[2] = {"data", "randoms"};
  const sif_field_t* given[2] = {data, randoms};
  for (int i = 0; i < 2; i++) {
    const sif_field_t* f = given[i];
    if (!f || !f->x || !f->y || !f->z || f->n_particles == 0) {
      error_set(SIF_ERR_INVALID, "%s: %s", names[i],
        !f ? "NULL" : f->n_particles == 0 ? "no particles" : "no positions");
      return SIF_ERR_INVALID;
    }
    if (f->units != SIF_COORDINATES_CARTESIAN) {
      error_set(SIF_ERR_INVALID, "%s: positions on the sky (expected Cartesian coordinates)",
        names[i]);
      return SIF_ERR_INVALID;
    }
  }
  if (!radii || n_radii == 0 || !offset || !box_length) {
    error_set(SIF_ERR_INVALID, "%s",
      !radii || n_radii == 0 ? "radii: none given" : "offset, box_length: NULL");
    return SIF_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n_radii; i++)
    if (!(radii[i] > 0.0f) || !std::isfinite(radii[i])) {
      error_set(SIF_ERR_INVALID, "radii[%u] = %g: not a positive radius", i,
        (double)radii[i]);
      return SIF_ERR_INVALID;
    }
  // exodus_survey.c:567-600: the bounding box, a coordinate that is not finite
  // counted (fminf and fmaxf would pass over it)
  auto extent = [&](const sif_field_t* f, const char* what, float lo[3], float hi[3]) {
    uint64_t bad = 0;
    for (uint64_t p = 0; p < f->n_particles; p++) {
      const float v[3] = {f->x[p], f->y[p], f->z[p]};
      bad += !(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]));
      for (int k = 0; k < 3; k++) {
        lo[k] = fminf(lo[k], v[k]);
        hi[k] = fmaxf(hi[k], v[k]);
      }
    }
    if (bad) {
      error_set(SIF_ERR_INVALID, "%s: %llu of %llu with a coordinate that is not finite", what,
        (unsigned long long)bad, (unsigned long long)f->n_particles);
      return SIF_ERR_INVALID;
    }
    return SIF_OK;
  };
  float lo[3] = {randoms->x[0], randoms->y[0], randoms->z[0]};
  float hi[3] = {lo[0], lo[1], lo[2]};
  if (extent(randoms, "randoms", lo, hi) != SIF_OK)
    return SIF_ERR_INVALID;
  if (extent(data, "data", lo, hi) != SIF_OK)
    return SIF_ERR_INVALID;
  float ext = 0.0f, r_min = radii[0], r_max = radii[0];
  for (int k = 0; k < 3; k++)
    ext = fmaxf(ext, hi[k] - lo[k]);
  for (uint32_t i = 1; i < n_radii; i++) {
    r_min = fminf(r_min, radii[i]);
    r_max = fmaxf(r_max, radii[i]);
  }
  // exodus_survey.c:655-690: margin = r_search + cell at a cell of r_min,
  // with 1% slack on each side
  const float f = sif__finder_search_factor(opt);
  const float cell = r_min;
  const float margin = r_max + fmaxf((f - 1.0f) * r_max, 3.0f * cell) + cell;
  const float L = ext + 2.0f * 1.01f * margin;
  for (int k = 0; k < 3; k++)
    offset[k] = 0.5f * L - 0.5f * (lo[k] + hi[k]);
  *box_length = L;
  return SIF_OK;
}

int sif_catalogue_translate(sif_catalogue_t* c, const sif_real offset[3]) {
  if (!c || !offset) {
    error_set(SIF_ERR_INVALID, "%s: NULL", !c ? "catalogue" : "offset");
    return SIF_ERR_INVALID;
  }
  if (c->units != SIF_COORDINATES_CARTESIAN) {
    error_set(SIF_ERR_INVALID, "catalogue: sky coordinates (ra, dec, z), not positions; a "
      "translation would move angles");
    return SIF_ERR_INVALID;
  }
  for (uint64_t i = 0; i < c->n_voids; i++) {
    c->cx[i] += offset[0];
    c->cy[i] += offset[1];
    c->cz[i] += offset[2];
  }
  return SIF_OK;
}

uint32_t sif_finder_grid_cells_from_size(uint64_t n_particles, sif_real box_length,
  sif_real cell_size, sif_finder_units_t units, sif_real volume) {
  // exodus_finder.c:2504-2560
  if (!std::isfinite(box_length) || !(box_length > 0.0f)) {
    error_set(SIF_ERR_INVALID, "box_length: %g, expected positive and finite", (double)box_length);
    return 0;
  }
  if (!std::isfinite(cell_size) || !(cell_size > 0.0f)) {
    error_set(SIF_ERR_INVALID, "cell_size: %g, expected positive and finite", (double)cell_size);
    return 0;
  }
  if (units != SIF_FINDER_UNITS_PHYSICAL && units != SIF_FINDER_UNITS_MPS) {
    error_set(SIF_ERR_INVALID, "units: %d", (int)units);
    return 0;
  }
  if (units == SIF_FINDER_UNITS_MPS && n_particles == 0) {
    error_set(SIF_ERR_INVALID, "n_particles: 0, a cell size in mps needs tracers");
    return 0;
  }
  if (units == SIF_FINDER_UNITS_MPS && (!std::isfinite(volume) || volume < 0.0f)) {
    error_set(SIF_ERR_INVALID, "volume: %g", (double)volume);
    return 0;
  }
  double cell = (double)cell_size;
  if (units == SIF_FINDER_UNITS_MPS) {
    const double b = (double)box_length;
    const double v = volume > 0.0f ? (double)volume : b * b * b;
    cell *= cbrt(v / (double)n_particles);
  }
  const double cells = ceil((double)box_length / cell * (1.0 - 1e-5));
  if (cells > (double)UINT32_MAX) {
    error_set(SIF_ERR_RANGE, "cell_size: %g cells per side is too many", cells);
    return 0;
  }
  return cells < 1.0 ? 1u : (uint32_t)cells;
}

// --- stacked profiles (profiles.h, profiles.c:24-92, 354-673) ---

void sif_density_profiles_free(sif_density_profiles_t* p) {
  if (!p)
    return;
  free(p->r_edges);
  free(p->profiles);
  free(p);
}

void sif_velocity_profiles_free(sif_velocity_profiles_t* p) {
  if (!p)
    return;
  free(p->r_edges);
  free(p->v_rad);
  free(p);
}

int sif_profiles(const sif_catalogue_t* cat, const sif_chain_mesh_t* mesh, sif_real ext,
  uint32_t n_bins, sif_option opt, sif_density_profiles_t** out_dens,
  sif_velocity_profiles_t** out_vel) {
  // profiles.c:97-142, in its order
  if (!cat || cat->n_voids == 0) {
    error_set(SIF_ERR_INVALID, "no voids to profile");
    return SIF_ERR_INVALID;
  }
  if (!mesh || mesh->n_particles == 0 || !(mesh->box_length > 0.0f)) {
    error_set(SIF_ERR_INVALID, "invalid or empty chain mesh");
    return SIF_ERR_INVALID;
  }
  if (!out_dens && !out_vel) {
    error_set(SIF_ERR_INVALID, "both outputs are NULL; there is nothing to compute");
    return SIF_ERR_INVALID;
  }
  if (n_bins == 0) {
    error_set(SIF_ERR_INVALID, "n_bins must be non-zero");
    return SIF_ERR_INVALID;
  }
  if (out_vel && !mesh->vx) {
    error_set(SIF_ERR_INVALID,
      "velocity profiles were requested but the mesh carries no velocities");
    return SIF_ERR_INVALID;
  }
  if (cat->units != SIF_COORDINATES_CARTESIAN) {
    error_set(SIF_ERR_INVALID,
      "the catalogue holds sky coordinates; profiles are measured around Cartesian "
      "centres, before sif_catalogue_to_sky()");
    return SIF_ERR_INVALID;
  }
  int st = need_init();
  if (st != SIF_OK)
    return st;
  TracerProfileSpec S;
  tracer_profile_spec(&S, mesh->n_cells, mesh->box_length, mesh->total_weight, ext, n_bins,
    opt, mesh->weights != nullptr, out_vel != nullptr);
  S.cell_length = mesh->cell_length;
  const uint64_t nv = cat->n_voids, rows = nv * n_bins;
  // a set the caller passed is refilled; NULL asks for a new one, zeroed
  // (profiles.c:393-412)
  bool made_dens = false, made_vel = false;
  auto fail = [&](int status) {
    if (out_dens && *out_dens) {
      sif_density_profiles_free(*out_dens);
      *out_dens = nullptr;
    }
    if (out_vel && *out_vel) {
      sif_velocity_profiles_free(*out_vel);
      *out_vel = nullptr;
    }
    return status;
  };
  if (out_dens && !*out_dens) {
    sif_density_profiles_t* p =
      (sif_density_profiles_t*)calloc(1, sizeof(sif_density_profiles_t));
    if (p) {
      p->n_voids = nv;
      p->n_bins = n_bins;
      p->ext = S.ext;
      p->differential = S.differential != 0;
      p->r_edges = (sif_real*)aligned_alloc_bytes((n_bins + 1) * sizeof(sif_real));
      p->profiles = (sif_real*)aligned_alloc_bytes(rows * sizeof(sif_real));
      if (p->profiles)
        memset(p->profiles, 0, rows * sizeof(sif_real));
    }
    if (!p || !p->r_edges || !p->profiles) {
      sif_density_profiles_free(p);
      error_set(SIF_ERR_ALLOC, "OOM allocating density profiles");
      return SIF_ERR_ALLOC;
    }
    *out_dens = p;
    made_dens = true;
  }
  if (out_vel && !*out_vel) {
    sif_velocity_profiles_t* p =
      (sif_velocity_profiles_t*)calloc(1, sizeof(sif_velocity_profiles_t));
    if (p) {
      p->n_voids = nv;
      p->n_bins = n_bins;
      p->ext = S.ext;
      p->r_edges = (sif_real*)aligned_alloc_bytes((n_bins + 1) * sizeof(sif_real));
      p->v_rad = (sif_real*)aligned_alloc_bytes(rows * sizeof(sif_real));
      if (p->v_rad)
        memset(p->v_rad, 0, rows * sizeof(sif_real));
    }
    if (!p || !p->r_edges || !p->v_rad) {
      sif_velocity_profiles_free(p);
      error_set(SIF_ERR_ALLOC, "OOM allocating velocity profiles");
      return fail(SIF_ERR_ALLOC);
    }
    *out_vel = p;
    made_vel = true;
  }
  (void)made_dens;
  (void)made_vel;
  if (out_dens)
    memcpy((*out_dens)->r_edges, S.edges.data(), (n_bins + 1) * sizeof(sif_real));
  if (out_vel) {
    // velocity edges are the cumulative binning's (profiles.c:431-432)
    for (uint32_t j = 0; j < n_bins; j++)
      (*out_vel)->r_edges[j] = (float)j * S.bin_width;
    (*out_vel)->r_edges[n_bins] = S.ext;
  }

  const uint64_t n = mesh->n_particles, cells = mesh->total_cells;
  Dev doff, dmx, dmy, dmz, dmw, dvx, dvy, dvz, dcx, dcy, dcz, drv, dd, dv;
  st = doff.upload(mesh->cell_offsets, (cells + 1) * 8);
  if (st == SIF_OK) st = dmx.upload(mesh->x, n * 4);
  if (st == SIF_OK) st = dmy.upload(mesh->y, n * 4);
  if (st == SIF_OK) st = dmz.upload(mesh->z, n * 4);
  if (st == SIF_OK && mesh->weights) st = dmw.upload(mesh->weights, n * 4);
  if (st == SIF_OK && out_vel) st = dvx.upload(mesh->vx, n * 4);
  if (st == SIF_OK && out_vel) st = dvy.upload(mesh->vy, n * 4);
  if (st == SIF_OK && out_vel) st = dvz.upload(mesh->vz, n * 4);
  if (st == SIF_OK) st = dcx.upload(cat->cx, nv * 4);
  if (st == SIF_OK) st = dcy.upload(cat->cy, nv * 4);
  if (st == SIF_OK) st = dcz.upload(cat->cz, nv * 4);
  if (st == SIF_OK) st = drv.upload(cat->radii, nv * 4);
  // the rows as they are, so that those of voids without a radius stay so
  if (st == SIF_OK && out_dens) st = dd.upload((*out_dens)->profiles, rows * 4);
  if (st == SIF_OK && out_vel) st = dv.upload((*out_vel)->v_rad, rows * 4);
  if (st != SIF_OK)
    return fail(st);
  S.offsets = doff.as<uint64_t>();
  S.mx = dmx.as<float>();
  S.my = dmy.as<float>();
  S.mz = dmz.as<float>();
  S.mw = mesh->weights ? dmw.as<float>() : nullptr;
  S.vx = out_vel ? dvx.as<float>() : nullptr;
  S.vy = out_vel ? dvy.as<float>() : nullptr;
  S.vz = out_vel ? dvz.as<float>() : nullptr;
  S.cx = dcx.as<float>();
  S.cy = dcy.as<float>();
  S.cz = dcz.as<float>();
  S.rv = drv.as<float>();
  S.n_voids = (uint32_t)nv;
  S.dens = out_dens ? dd.as<float>() : nullptr;
  S.vel = out_vel ? dv.as<float>() : nullptr;
  if (nv > 0xffffffffull) {
    error_set(SIF_ERR_RANGE, "profiles: %llu voids (at most 2^32 - 1)", (unsigned long long)nv);
    return fail(SIF_ERR_RANGE);
  }
  st = tracer_profiles(S);
  if (st == SIF_OK && out_dens &&
      hipMemcpyAsync((*out_dens)->profiles, dd.p, rows * 4, hipMemcpyDeviceToHost, stream()) !=
        hipSuccess)
    st = SIF_ERR_INVALID;
  if (st == SIF_OK && out_vel &&
      hipMemcpyAsync((*out_vel)->v_rad, dv.p, rows * 4, hipMemcpyDeviceToHost, stream()) !=
        hipSuccess)
    st = SIF_ERR_INVALID;
  if (st == SIF_OK && hipStreamSynchronize(stream()) != hipSuccess)
    st = SIF_ERR_INVALID;
  if (st != SIF_OK) {
    if (st == SIF_ERR_INVALID)
      error_set(SIF_ERR_INVALID, "profiles: the device run failed");
    return fail(st);
  }
  return SIF_OK;
}

}  // extern "C"
