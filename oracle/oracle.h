/*
 * TEST INFRASTRUCTURE ONLY -- the oracle for the sif-amd exodus hot path.
 *
 * A CPU restatement, in plain C99, of the reference algorithm (sif v0.5.0)
 * for every stage of the path: field check/fold, CIC deposit, density
 * contrast, chain mesh, filter/window tables, 3D FFT smoothing, candidate
 * scan and sort, mesh template, sphere growth, witness, overlap tests, mask
 * painting and the sequential exodus commit. Each function cites the
 * reference file:line it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the timed CPU baseline. Nothing in
 * sif_amd/ links or calls it.
 *
 * Each function is pinned by checks that a plausible mistake here would fail:
 * the reference's own test vectors, closed forms, brute forces over every
 * tracer or node, and invariants, named in the tests under tests/ that call
 * it. Parity unpinned: FFTW's output bits (FFTW is absent; the smoothed field
 * is checked against an FP64 transform instead), and so the end-to-end
 * catalogue of the reference's float path with FFTW's bits.
 */
#ifndef SIF_ORACLE_H
#define SIF_ORACLE_H

#include <stdint.h>

/* --- field --- */
float orc_box_fold(float v, float box);
/* Returns the number of particles outside [-tol, box + tol]; n_edge receives
 * those accepted only through the tolerance, first_bad the first refused. */
uint64_t orc_check_in_box(const float* x, const float* y, const float* z,
  uint64_t n, float box, uint64_t* n_edge, uint64_t* first_bad);

/* --- CIC --- */
/* Float deposit in the reference's slab/chunk/tile order; out = density. */
void orc_cic_float(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box, uint64_t tile,
  float* out);
/* Fixed-point deposit: per-corner float weight rounded to 2^-s units and
 * summed exactly in int64, the model of the GPU kernel; s from the count and
 * the largest |w| (orc_cic_unit_exponent), returned. */
int orc_cic_unit_exponent(uint64_t n, float wmax);
int orc_cic_fixed(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box, int64_t* acc);
void orc_cic_fixed_density(const int64_t* acc, uint32_t n_cells, float box, int s,
  float* out, double* mean_out);
/* Textbook double-precision CIC, no shared arithmetic with the above. */
void orc_cic_brute_double(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box, double* out);
/* Reference density contrast (64 ordered double blocks). 0 or -1 on a
 * non-positive mean. */
int orc_density_contrast(float* values, uint64_t total_cells, double* mean_out);
/* The same map with a mean supplied by the caller. */
void orc_density_contrast_with_mean(float* values, uint64_t total_cells,
  double mean);

/* --- chain mesh --- */
void orc_chain_mesh(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box,
  uint64_t* offsets, float* mx, float* my, float* mz, float* mw,
  uint64_t* original_indices, float* cell_weights, double* total_weight);
uint64_t orc_chain_mesh_cell(float x, float y, float z, uint32_t n_cells,
  float box);

/* --- filter and window --- */
#define ORC_FILTER_NONE     0
#define ORC_FILTER_TOP_HAT  1
#define ORC_FILTER_GAUSSIAN 2
uint32_t orc_filter_lut_size(uint32_t n_cells);
void orc_filter_lut(int filter, float r, float sigma, float box,
  uint32_t n_cells, float* lut);
/* mode 0: 1/W (full), 1: W/S (aliasing-aware); order 0 fills 1.0 */
void orc_window_axis(uint32_t n_cells, int order, int mode, double* axis);
double orc_window_corner_gain(int order, int mode, int interlaced);
/* Per-mode factor as the reference applies it: (float)(lut * wx * wy * wz). */
float orc_mode_factor(const float* lut, const double* axis, uint32_t n,
  uint32_t ix, uint32_t iy, uint32_t iz);

/* --- 3D FFT in double (power-of-two n), spectrum [n][n][n/2+1] --- */
int orc_fft_r2c(const double* in, uint32_t n, double* spec /* 2 per mode */);
int orc_fft_c2r(const double* spec, uint32_t n, double* out /* unscaled */);

/* --- 3D FFT in float (power-of-two n >= 4), the CPU baseline's: spectrum
 * [n][n][n/2+1] interleaved complex; the c2r multiplies by the mode factors
 * (lut null for none) and by norm --- */
int orc_fftf_r2c(const float* in, uint32_t n, float* spec);
int orc_fftf_filter_c2r(const float* spec, uint32_t n, const float* lut, const double* axis,
  float norm, float* work, float* out);

/* --- candidates --- */
typedef struct {
  float delta;
  uint32_t refined;
  uint64_t flat_idx;
} orc_candidate_t;
uint64_t orc_scan_candidates(const float* delta, const uint64_t* mask_words,
  uint64_t total_cells, float threshold, orc_candidate_t* out);
/* sort.h:56-99 three-way quicksort, middle pivot: the reference tie order. */
void orc_sort_candidates_ref(orc_candidate_t* c, uint64_t n);
/* (delta ascending, flat ascending): the GPU tie rule. */
void orc_sort_candidates_flat(orc_candidate_t* c, uint64_t n);

/* --- template and growth --- */
#define ORC_CELL_CORE    1
#define ORC_CELL_STRADDLES 2
#define ORC_CELL_EDGE    3
#define ORC_CELL_SHELL   4
typedef struct {
  int32_t* dx;
  int32_t* dy;
  int32_t* dz;
  uint8_t* type;
  float* min_d2;
  float* max_d2;
  uint32_t count;
} orc_template_t;
int orc_template_build(orc_template_t* t, uint32_t mesh_cells, float cell_length,
  float rmin, float r_search);
void orc_template_free(orc_template_t* t);

typedef struct {
  uint32_t n_cells;
  float box_length;
  float cell_length;
  const uint64_t* offsets;
  const float* x;
  const float* y;
  const float* z;
  /* a weighted mesh (chain_mesh.h:80-95): tracer weights, per-cell sums (a
   * float each, summed in double) and the total; NULL on an unweighted one */
  const float* w;
  const float* cell_w;
  double total_weight;
  /* velocities in mesh order, or NULL */
  const float* vx;
  const float* vy;
  const float* vz;
} orc_mesh_t;

/* sif_profiles (src/measure/profiles.c:354-673) for voids (cx, cy, cz, rv)
 * over the mesh: float accumulators per void, the reference's cell walk and
 * order. opt: SIF_PBC_OPEN (bit 0), SIF_PROFILES_DIFFERENTIAL (bit 8),
 * SIF_PROFILES_VELOCITY_WEIGHTED (bit 9). dens / vel: n_voids * n_bins or
 * NULL; count (optional): the tracers binned, n_voids * n_bins. Rows of voids
 * of non-positive radius are not written. Returns 0, or -1 on a bad
 * argument. */
int orc_profiles(const orc_mesh_t* m, const float* cx, const float* cy,
  const float* cz, const float* rv, uint64_t n_voids, float ext,
  uint32_t n_bins, uint32_t opt, float* dens, float* vel, uint64_t* count);

#define ORC_RESCALE_OK         0
#define ORC_RESCALE_EMPTY      1
#define ORC_RESCALE_BEYOND     2
#define ORC_RESCALE_BELOW_RUNG 3
#define ORC_RESCALE_DEGENERATE 4
#define ORC_RESCALE_NO_RANDOMS 6
/* Unweighted find_exact_radius_impl. window = bins resolved per pass. n_in
 * (optional) receives the enclosed count at the crossing. */
float orc_find_radius(const orc_mesh_t* m, float cx, float cy, float cz,
  float r_search, float threshold, float vol_factor, float rmin,
  const orc_template_t* tpl, uint32_t window, uint8_t* reason, uint64_t* n_in);
/* The weighted one (exodus_rescale.c:1029-1037), on m->w and m->cell_w, the
 * density tests against mass_factor. n_in: the tracers enclosed. */
float orc_find_radius_weighted(const orc_mesh_t* m, float cx, float cy,
  float cz, float r_search, float threshold, float vol_factor,
  float mass_factor, float rmin, const orc_template_t* tpl, uint32_t window,
  uint8_t* reason, uint64_t* n_in);
int orc_witness(const orc_mesh_t* m, float cx, float cy, float cz, float rho2,
  float r_search, float threshold, float vol_factor, const orc_template_t* tpl,
  int closed);
/* test_finders.c:433-465, double, from every particle. */
double orc_brute_force_radius(const float* x, const float* y, const float* z,
  uint64_t n, float box, double cx, double cy, double cz, double rung,
  double r_search, double threshold);

/* --- overlap and mask --- */
int orc_overlap_quick(const uint64_t* mask_words, uint32_t n_cells, uint32_t ix,
  uint32_t iy, uint32_t iz, uint32_t r);
void orc_mark_sphere(uint64_t* mask_words, float cx, float cy, float cz,
  float r, uint32_t n_cells, float cell_length);
/* utils.c:850-960 for nq queries against nv voids (reference 32^3 list,
 * reference max_r given); out[q] = 1 on an overlap. */
void orc_overlap_exact_many(const float* vcx, const float* vcy, const float* vcz,
  const float* vr, uint64_t nv, float max_r, const float* qx, const float* qy,
  const float* qz, const float* qr, uint64_t nq, float box, float f, float scale,
  uint8_t* out);

/* --- sequential exodus driver (box, cells on nodes, unweighted) --- */
typedef struct {
  float* cx;
  float* cy;
  float* cz;
  float* r;
  uint64_t* n_in;
  uint32_t* rung;
  uint64_t n;
  uint64_t capacity;
} orc_catalogue_t;

typedef struct {
  uint64_t candidates, masked, proxy, overlap, witness, rescale_fail, accepted,
    displaced;
} orc_rung_stats_t;

/* utils.c:406-552 with minima_only and no refinement: the cells of
 * orc_scan_candidates that are local minima of the periodic n^3 field. */
uint64_t orc_scan_minima(const float* delta, uint32_t n, const uint64_t* mask,
  float t, orc_candidate_t* out);

/* search_opt bit 9 (SIF_FINDER_SEEDS_MINIMA): the candidates are the minima */
#define ORC_SEEDS_MINIMA (1u << 9)
/* search_opt bit 21 (SIF_FINDER_EXTEND_REACH): a growth that ends beyond the
 * rung's reach is measured again at the extended levels */
#define ORC_EXTEND_REACH (1u << 21)
/* search_opt bit 22 (SIF_FINDER_DISPLACE): a void is refused only by a larger
 * one and displaces the smaller ones it collides with */
#define ORC_DISPLACE (1u << 22)

/* fields[i] is the smoothed density contrast for the i-th radius of the
 * descending ladder (n_cells^3 floats each). tie_rule 0 = reference
 * quicksort, 1 = (delta, flat). use_witness mirrors the reference's witness
 * refusal (exodus_finder.c:2021-2035). Returns 0, or a negative status. */
int orc_exodus_run(float* const* fields, uint32_t n_cells, float box,
  const orc_mesh_t* mesh, uint64_t n_particles, const float* radii_desc,
  uint32_t n_radii, float threshold, float overlap_fraction, int lagrangian,
  uint32_t search_opt, int tie_rule, int use_witness, uint32_t window,
  orc_catalogue_t* cat, orc_rung_stats_t* stats);
void orc_catalogue_free(orc_catalogue_t* cat);

/* Whole CPU path on the host cores (OpenMP): CIC (reference float order),
 * contrast, mesh, r2c, per rung top-hat + window + c2r, then orc_exodus_run.
 * fp32 0: the transforms in double, each field rounded to float, all rungs
 * smoothed before the finder; fp32 1: the transforms in float (the
 * reference's precision) and each rung smoothed as the finder reaches it,
 * into one buffer, as the reference does. Per-stage seconds go to
 * stage_s[6]: cic, contrast, mesh, forward fft, smoothing (all rungs),
 * finder (less the smoothing). */
int orc_exodus_pipeline(const float* x, const float* y, const float* z,
  uint64_t n, uint32_t n_cells, float box, uint32_t mesh_cells,
  const float* radii, uint32_t n_radii, float threshold, uint32_t max_rungs,
  int fp32, orc_catalogue_t* cat, double* stage_s);

/* orc_exodus_run with each rung's field made when the run reaches it:
 * field(ctx, i) returns rung i's field (valid until the next call). */
typedef const float* (*orc_field_fn)(void* ctx, uint32_t rung);
int orc_exodus_run_lazy(orc_field_fn field, void* ctx, uint32_t n, float box,
  const orc_mesh_t* mesh, uint64_t n_particles, const float* radii,
  uint32_t n_radii, float t, float f, int lagrangian, uint32_t search_opt,
  int tie_rule, int use_witness, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats);

/* --- survey: src/finder/exodus_survey.c, exodus_rescale.c:1909-2317 ---
 * Occupancy masks are one byte a grid node (0/1), flat (ix n + iy) n + iz. */
uint32_t orc_survey_node(float p, float inv_l, uint32_t n);
double orc_survey_randoms_per_node(double m);
/* The closing of occ over a (2k + 1)^3 box, nodes past a face unset. */
void orc_survey_close(uint8_t* occ, uint32_t n, int32_t k);
/* The footprint of the randoms, closed; returns k (0: nothing marked). */
int orc_survey_occupancy(const float* rx, const float* ry, const float* rz, uint64_t nr,
  uint32_t n, float cell_length, uint8_t* occ, double* lambda, uint64_t* marked,
  uint64_t* observed);
void orc_survey_contrast(float* data, const float* randoms, const uint8_t* occ, float alpha,
  uint64_t total);
void orc_survey_void_footprint(const uint8_t* occ, uint32_t n, float l, float box, float cx,
  float cy, float cz, float r, float* inside, float* shell);
/* The survey radius against the random mesh rm (its own template rtpl);
 * n_in: the data tracers enclosed. */
float orc_find_radius_survey(const orc_mesh_t* dm, const orc_mesh_t* rm,
  const orc_template_t* dtpl, const orc_template_t* rtpl, float cx, float cy, float cz,
  float r_search, float K, float rmin, uint32_t window, uint8_t* reason, uint64_t* n_in);
int orc_exodus_run_survey(float* const* fields, uint32_t n, float box,
  const orc_mesh_t* data_mesh, const orc_mesh_t* random_mesh, float survey_K,
  const float* radii_desc, uint32_t n_radii, float threshold, float overlap_fraction,
  int lagrangian, uint32_t search_opt, int tie_rule, uint32_t window, orc_catalogue_t* cat,
  orc_rung_stats_t* stats);

int orc_num_threads(void);

#endif
