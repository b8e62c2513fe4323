/*
 * sif-amd device-level C ABI: the stages of the exodus path on data that is
 * already resident in HBM. Plain pointers and sizes only; every pointer named
 * d_* (or documented as device) is device memory on the current HIP device.
 * Each call runs on the library stream and returns once its results are
 * complete. Errors: a negative SIF_ERR_* status, with sif_error_message().
 *
 * These are the entry points the host-struct API of include/sif/ (the drop-in
 * for sif v0.5.0) is built on, exported so that callers holding device
 * buffers -- the bench, a torch autograd function, tests -- skip the host
 * round trip. Each cites the reference function whose work it does.
 */
#ifndef SIF_AMD_H
#define SIF_AMD_H

#include <stdint.h>

#include "sif/core/macros.h"
#include "sif/structures/catalogue.h"

SIF_EXTERN_C_BEGIN

/* --- plumbing (no reference counterpart: device memory for callers without
 * a framework of their own) --- */
SIF_API int sif_amd_device_count(void);
SIF_API int sif_amd_malloc(void** d_ptr, uint64_t bytes);
SIF_API int sif_amd_free(void* d_ptr);
SIF_API int sif_amd_memcpy_h2d(void* d_dst, const void* src, uint64_t bytes);
SIF_API int sif_amd_memcpy_d2h(void* dst, const void* d_src, uint64_t bytes);
SIF_API int sif_amd_memset(void* d_dst, int value, uint64_t bytes);
SIF_API int sif_amd_synchronize(void);
/* Device memory in use on the library's device (every allocation of the
 * process, the library's cached work buffers included) and its size, from
 * hipMemGetInfo; no counterpart in the reference. */
SIF_API int sif_amd_device_memory(uint64_t* used, uint64_t* total);
/* A plain device-to-device copy kernel (16-byte loads and stores, bytes a
 * multiple of 16, both pointers 16-byte aligned): traffic of known size, the
 * yardstick rocprof's FETCH_SIZE and WRITE_SIZE counters are scaled by. */
SIF_API int sif_amd_copy(void* d_dst, const void* d_src, uint64_t bytes);

/* Per-stage device time (HIP events on the library stream, recorded without
 * synchronising and resolved when read). Stage order:
 * cic, contrast, mesh, fft_forward, smooth, scan, prefilter, grow, commit,
 * retest, sort, host, profiles, commit_order, commit_index. commit, retest,
 * host, commit_order and commit_index are host time: the commit's
 * per-candidate tests against the rung's voids and the batch (on the host
 * pool), the re-tests that assemble each growth batch, the rest of the
 * driver's own work per rung (the void index, the survivors' copy, the
 * filter tables), the commit's ordered accept loop, and the commit's
 * preparation (the candidates' nodes and the index of the batch's own
 * voids). profiles is sif_profiles' binning. */
#define SIF_AMD_TIMING_STAGES 15
SIF_API int sif_amd_timing_enable(int on);
SIF_API void sif_amd_timing_reset(void);
SIF_API int sif_amd_timing_read(double* ms, uint64_t* launches, int n_stages);
/* Algorithmic bytes of the timed launches per stage, where the stage knows
 * them (the minimal HBM traffic of the work done; 0 elsewhere). */
SIF_API int sif_amd_timing_bytes(double* bytes, int n_stages);

/* --- field --- */
/* sif__field_check_in_box (src/structures/field.c:889-935). */
SIF_API int sif_amd_check_in_box(const float* d_x, const float* d_y,
  const float* d_z, uint64_t n, float box, uint64_t* n_bad, uint64_t* n_edge,
  uint64_t* first_bad);

/* --- grid --- */
/* Bytes of device scratch sif_amd_grid_assign_cic() needs. */
SIF_API uint64_t sif_amd_cic_work_bytes(uint32_t n_cells);
/* grid_compute_cic for CIC (src/structures/grid.c:679-746): density per unit
 * volume into d_values (n_cells^3 floats). d_w may be NULL. */
SIF_API int sif_amd_grid_assign_cic(const float* d_x, const float* d_y,
  const float* d_z, const float* d_w, uint64_t n, uint32_t n_cells, float box,
  float* d_values, void* d_work);
/* The deterministic mean of a grid (65536-cell blocks, double). */
SIF_API int sif_amd_grid_mean(const float* d_values, uint64_t total_cells,
  double* mean);
/* sif_grid_to_density_contrast (grid.c:1177-1244), in place. */
SIF_API int sif_amd_grid_to_density_contrast(float* d_values,
  uint64_t total_cells, double* mean_out);

/* No counterpart in the reference: sif_amd_grid_to_density_contrast for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The mean's block
 * sums go to a fixed scratch array in device memory that the library holds
 * per device until the process ends, so grids of at most 2^31 cells
 * (SIF_ERR_RANGE beyond); sif_system_init() must have run
 * (SIF_ERR_UNINITIALISED). Same mean and contrast, bit for bit, as
 * sif_amd_grid_to_density_contrast. Calls on one device share the scratch:
 * calls that overlap on a device race. */
SIF_API int sif_amd_grid_to_density_contrast_noalloc(float* d_values,
  uint64_t total_cells, double* mean_out);

/* --- chain mesh --- */
/* chain_mesh_create + canonical order (src/structures/chain_mesh.c:680-832,
 * :364-542): d_offsets has n_cells^3 + 1 entries; d_w, d_mw,
 * d_original_indices and d_cell_weights may be NULL. */
SIF_API int sif_amd_chain_mesh_build(const float* d_x, const float* d_y,
  const float* d_z, const float* d_w, uint64_t n, uint32_t n_cells, float box,
  uint64_t* d_offsets, float* d_mx, float* d_my, float* d_mz, float* d_mw,
  uint64_t* d_original_indices, float* d_cell_weights, double* total_weight);

/* --- FFT smoothing --- */
#define SIF_AMD_FILTER_NONE     0
#define SIF_AMD_FILTER_TOP_HAT  1
#define SIF_AMD_FILTER_GAUSSIAN 2
/* A spectrum is n x-planes of complex64: plane x holds n rows of the pitch,
 * row (x, y) the n/2 + 1 modes kz = 0 .. n/2 of FFTW's r2c layout first, then
 * padding up to the pitch (n/2 + 1 rounded up to a multiple of 16, so rows
 * start on 128-byte lines); each plane ends with 48 more complex values of
 * padding. sif_amd_spectrum_bytes() is its size (n planes), also that of the
 * FFT work buffer. */
SIF_API uint32_t sif_amd_spectrum_pitch(uint32_t n);
SIF_API uint64_t sif_amd_spectrum_bytes(uint32_t n);
/* sif__fft_grid_forward (src/math/fft.c:790-818). */
SIF_API int sif_amd_fft_r2c(const float* d_values, uint32_t n, float* d_spectrum,
  void* d_work);
/* sif__fft_apply_filter_smoothed + sif__fft_grid_backward (fft.c:703-888):
 * window_order 0 applies no window correction, else 1/W (mode 0) or W/S
 * (mode 1) of order window_order. */
SIF_API int sif_amd_fft_filter_c2r(const float* d_spectrum, uint32_t n,
  float box, int filter, float r, float sigma, int window_order,
  int window_mode, float* d_out, void* d_work);

/* --- candidates --- */
/* sif__finder_scan_candidates (src/finder/utils.c:130-213): cells with
 * delta <= threshold and a clear mask bit (d_mask_words may be NULL), sorted
 * (delta, flat). Writes min(count, capacity) entries; *count is the total. */
SIF_API int sif_amd_scan_candidates(const float* d_delta,
  const uint64_t* d_mask_words, uint64_t total_cells, float threshold,
  float* d_out_delta, uint32_t* d_out_flat, uint64_t capacity, uint64_t* count);

/* sif__finder_scan_refined with minima_only and no refinement
 * (src/finder/utils.c:228-262, 406-552), the SIF_FINDER_SEEDS_MINIMA scan: of
 * those cells, the local minima of the periodic n_cells^3 field -- strictly
 * below the 26 neighbours, an equal neighbour of lower flat index winning --
 * on their nodes, sorted (delta, flat). Same output contract. */
SIF_API int sif_amd_scan_minima(const float* d_delta, const uint64_t* d_mask_words,
  uint32_t n_cells, float threshold, float* d_out_delta, uint32_t* d_out_flat,
  uint64_t capacity, uint64_t* count);

/* --- sphere growth --- */
/* sif__exodus_find_radius (src/finder/exodus_rescale.c:642-1027), unweighted,
 * one result per centre: radius (or -1), reason (0 ok, 1 empty, 2 beyond the
 * search sphere, 3 below the rung, 4 degenerate, 5 refinement overflow) and
 * the enclosed count at the crossing. d_n_in may be NULL. */
SIF_API int sif_amd_find_radius(const uint64_t* d_offsets, const float* d_mx,
  const float* d_my, const float* d_mz, uint32_t mesh_cells, float box,
  const float* d_cx, const float* d_cy, const float* d_cz, uint32_t n,
  float rmin, float r_search, float threshold, float vol_factor,
  float* d_radius, uint32_t* d_reason, uint64_t* d_n_in);

/* sif__exodus_find_radius_weighted (src/finder/exodus_rescale.c:1029-1037):
 * the same on weighted tracers -- d_mw and d_cell_weights (device) as
 * sif_amd_chain_mesh_build gives them, the density tests against
 * mass_factor, (4/3) pi total_weight / box^3 (exodus_finder.c:1639-1642);
 * vol_factor still sizes the histogram. Weights are not checked here (see
 * sif_amd_exodus_device_weighted). n_in receives the tracers enclosed. */
SIF_API int sif_amd_find_radius_weighted(const uint64_t* d_offsets, const float* d_mx,
  const float* d_my, const float* d_mz, const float* d_mw, const float* d_cell_weights,
  uint32_t mesh_cells, float box, const float* d_cx, const float* d_cy, const float* d_cz,
  uint32_t n, float rmin, float r_search, float threshold, float vol_factor,
  float mass_factor, float* d_radius, uint32_t* d_reason, uint64_t* d_n_in);

/* --- survey (src/finder/exodus_survey.c) --- */

/* One chain mesh on the device, as sif_amd_chain_mesh_build leaves it:
 * offsets (n_cells^3 + 1), positions, and for a weighted mesh the tracer and
 * cell weights (NULL on an unweighted one). */
typedef struct {
  const uint64_t* offsets;
  const float* x;
  const float* y;
  const float* z;
  const float* weights;
  const float* cell_weights;
  double total_weight;
  uint32_t n_cells;
  uint64_t n_particles;
} sif_amd_mesh_t;

/* sif__exodus_find_radius_survey (src/finder/exodus_rescale.c:2172-2317):
 * the largest data distance d inside r_search with D(<d) <= K R(<d), D and R
 * the data and random weight enclosed (R interpolated in d^2 within each
 * histogram bin), K = (1 + threshold) alpha. Reasons as sif_amd_find_radius's,
 * 6 for no randoms inside r_search; n_in (may be NULL) the data tracers
 * enclosed. Both meshes share the box; their cells need not match. */
SIF_API int sif_amd_find_radius_survey(const sif_amd_mesh_t* data,
  const sif_amd_mesh_t* randoms, float box, const float* d_cx, const float* d_cy,
  const float* d_cz, uint32_t n, float rmin, float r_search, float K, float* d_radius,
  uint32_t* d_reason, uint64_t* d_n_in);

/* sif__exodus_survey_occupancy (src/finder/exodus_survey.c:148-207): the
 * nodes of an n_cells^3 grid over `box` that the n positions reach (node
 * nearest each), closed over (2k + 1)^3 nodes, k in [1, 4] the smallest that
 * expects 128 randoms at lambda randoms a node. d_occupied: n_cells^3 bytes,
 * 0 or 1, flat (ix n + iy) n + iz. k (0 when nothing is marked), lambda,
 * marked and observed node counts may each be NULL. */
SIF_API int sif_amd_survey_occupancy(const float* d_x, const float* d_y, const float* d_z,
  uint64_t n, uint32_t n_cells, float box, uint8_t* d_occupied, int32_t* k, double* lambda,
  uint64_t* marked, uint64_t* observed);

/* sif__exodus_survey_contrast (exodus_survey.c:223-234), in place:
 * data / (alpha randoms) - 1 on observed nodes with randoms expected, 1e30
 * elsewhere. */
SIF_API int sif_amd_survey_contrast(float* d_data, const float* d_randoms,
  const uint8_t* d_occupied, float alpha, uint64_t total_cells);

/* sif__exodus_survey_footprint (exodus_survey.c:247-318): the observed
 * fraction of each void's sphere and of its shell out to twice its radius.
 * Void columns and results are host arrays. */
SIF_API int sif_amd_survey_footprint(const uint8_t* d_occupied, uint32_t n_cells, float box,
  const float* cx, const float* cy, const float* cz, const float* r, uint64_t n_voids,
  float* inside, float* shell);

/* sif__exodus_template_build (exodus_rescale.c:175-283), host only: the
 * per-rung stencil the growth kernel walks, as (dx, dy, dz, class) and
 * (min d^2, max d^2) per entry; *count is the total, entries past capacity
 * are not written. */
SIF_API int sif_amd_template(uint32_t mesh_cells, float box, float rmin,
  float r_search, int32_t* cells4, float* bounds2, uint32_t capacity,
  uint32_t* count);

/* --- the whole finder on device-resident data --- */

/* Per-rung counters of a run, as sif__finder_log_radius prints them
 * (src/finder/utils.c:776-808). */
typedef struct {
  float radius;
  uint64_t candidates, masked, proxy, overlap, rescale_fail, accepted;
  uint64_t evaluated; /* growths the GPU ran for this rung */
} sif_amd_rung_stats_t;

/* sif__mark_sphere (src/finder/utils.c:965-1033): paints n_voids spheres
 * (radius scale * d_r[i]) into the n_cells^3-bit mask d_mask_words (device,
 * (n_cells^3 + 63) / 64 words), OR-ing into what it holds. */
SIF_API int sif_amd_mark_spheres(uint64_t* d_mask_words, uint32_t n_cells, float box,
  const float* d_cx, const float* d_cy, const float* d_cz, const float* d_r,
  uint32_t n_voids, float scale);

/* sif__exodus_run (src/finder/exodus_finder.c:1585-2492) for a periodic box,
 * SIF_FINDER_SEEDS_CELLS on their nodes, unweighted: the grid d_delta (a
 * density contrast, n_cells^3) and the mesh arrays are device memory. On a
 * power-of-two grid (16 to 1024) without SIF_FINDER_SEEDS_MINIMA the rungs
 * list their candidate cells instead of smoothing the grid, which is left as
 * it was; otherwise it is left as the last rung's smoothing unless restore
 * (then transformed back, to float precision). n_in (optional, host,
 * capacity of the result) receives each void's enclosed count. stats
 * (optional, host, n_radii entries) receives the per-rung counters in ladder
 * order (largest first). */
SIF_API sif_catalogue_t* sif_amd_exodus_device(float* d_delta,
  uint32_t n_cells, float box, const uint64_t* d_offsets, const float* d_mx,
  const float* d_my, const float* d_mz, uint32_t mesh_cells,
  uint64_t n_particles, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_real overlap_fraction, sif_option options,
  int restore, uint64_t** n_in, sif_amd_rung_stats_t* stats);

/* sif_finder_exodus_survey_raw (src/finder/exodus_survey.c:386-535) on
 * device-resident data: d_data the data's CIC density grid (NOT a density
 * contrast; smoothed in place, restored when `restore`), d_randoms the
 * randoms' on the same grid (read only), the two meshes over the same box
 * (their cells need not match). The survey must sit at least the largest
 * search sphere plus a cell from every face (SIF_ERR_INVALID otherwise, the
 * padding needed in the message). Seeds on their nodes only. The catalogue
 * carries the footprint columns; coordinates are the box's. */
SIF_API sif_catalogue_t* sif_amd_exodus_survey_device(float* d_data,
  const float* d_randoms, uint32_t n_cells, float box, const sif_amd_mesh_t* data,
  const sif_amd_mesh_t* randoms, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_real overlap_fraction, sif_option options, int restore,
  uint64_t** n_in, sif_amd_rung_stats_t* stats);

/* The same on weighted tracers (sif__exodus_find_radius_weighted,
 * exodus_rescale.c:1029-1037; the density tests against the mean weight
 * density, exodus_finder.c:1634-1642): d_mw and d_cell_weights as
 * sif_amd_chain_mesh_build gives them (device), total_weight the mesh's
 * summed weight. The weights are checked as exodus_finder.c:1051-1106 checks
 * them (finite and non-negative by their bits, a positive total):
 * SIF_ERR_INVALID otherwise. n_in receives the tracers each void encloses. */
SIF_API sif_catalogue_t* sif_amd_exodus_device_weighted(float* d_delta,
  uint32_t n_cells, float box, const uint64_t* d_offsets, const float* d_mx,
  const float* d_my, const float* d_mz, const float* d_mw,
  const float* d_cell_weights, double total_weight, uint32_t mesh_cells,
  uint64_t n_particles, const sif_real* radii, uint32_t n_radii,
  sif_real threshold, sif_real overlap_fraction, sif_option options,
  int restore, uint64_t** n_in, sif_amd_rung_stats_t* stats);

/* --- profiles of a catalogue over the tracers --- */
/* sif_profiles (src/measure/profiles.c:354-673) on device arrays: the chain
 * mesh (d_offsets, n^3 + 1 entries; positions; d_mw weights or NULL; d_vx,
 * d_vy, d_vz velocities, needed only with d_vel) of a box of side box and
 * total weight total_weight (the tracer count when unweighted), and n_voids
 * voids (d_cx, d_cy, d_cz, d_rv). ext (not positive: 5), n_bins and options
 * (SIF_PBC_*, SIF_PROFILES_*) as sif_profiles takes them. d_dens and d_vel
 * (n_voids * n_bins floats, either may be NULL) receive the rows; a void of
 * non-positive radius leaves its rows as they were. An unweighted density
 * profile is the reference's bit for bit; weighted sums and velocities are
 * summed in double in a fixed order. */
SIF_API int sif_amd_tracer_profiles(const uint64_t* d_offsets, const float* d_mx,
  const float* d_my, const float* d_mz, const float* d_mw, const float* d_vx,
  const float* d_vy, const float* d_vz, uint32_t mesh_cells, float box,
  double total_weight, const float* d_cx, const float* d_cy, const float* d_cz,
  const float* d_rv, uint32_t n_voids, float ext, uint32_t n_bins, sif_option options,
  float* d_dens, float* d_vel);

/* --- profiles of a grid, and the adjoint of the profile chain --- */
/* sif_profiles (src/measure/profiles.c:354-673), density part, cumulative
 * (profiles.c:332-350, 619-635), of a density-contrast grid d_delta
 * (n_cells^3, node g at g * box / n_cells) whose every node is a mass element
 * (1 + delta) rho_bar l^3, binned as profiles.c:202-223 bins a tracer:
 * d_profiles[c * n_bins + j] = l^3 sum over nodes of bins <= j of
 * (1 + delta) / ((4/3) pi ((j+1) ext / n_bins)^3 rv^3) - 1, for the centres
 * (d_cx, d_cy, d_cz) of radius d_rv out to ext * rv; a centre of radius <= 0
 * gets a zero row. 1 <= n_bins <= 32. */
SIF_API int sif_amd_grid_profiles(const float* d_delta, uint32_t n_cells,
  float box, const float* d_cx, const float* d_cy, const float* d_cz,
  const float* d_rv, uint32_t n_centres, float ext, uint32_t n_bins,
  float* d_profiles);
/* Its transpose (no reference counterpart: the adjoint of the profiles):
 * d_grad_delta[g] = sum over c, j of d_grad_profiles[c * n_bins + j]
 * * d profiles[c, j] / d delta[g], every node written. n_cells a multiple of
 * 16. Deterministic: no float atomics. */
SIF_API int sif_amd_grid_profiles_adjoint(const float* d_grad_profiles,
  uint32_t n_cells, float box, const float* d_cx, const float* d_cy,
  const float* d_cz, const float* d_rv, uint32_t n_centres, float ext,
  uint32_t n_bins, float* d_grad_delta);
/* The transpose of the CIC deposit (no reference counterpart; the deposit is
 * grid_compute_cic, src/structures/grid.c:585-650): for the density per unit
 * volume rho = sif_amd_grid_assign_cic(x, y, z) and a grid d_grad_rho, the
 * gradient of sum_g d_grad_rho[g] rho[g] with respect to each particle's
 * position, into d_gx, d_gy, d_gz. */
SIF_API int sif_amd_cic_adjoint(const float* d_x, const float* d_y,
  const float* d_z, uint64_t n, uint32_t n_cells, float box,
  const float* d_grad_rho, float* d_gx, float* d_gy, float* d_gz);

/* The chain's loss and its gradient (no reference counterpart): *loss =
 * sum over i < n of (d_p[i] - d_target[i])^2 (double, deterministic) and, when
 * d_grad is not NULL, d_grad[i] = scale * 2 (d_p[i] - d_target[i]). */
SIF_API int sif_amd_squared_residual(const float* d_p, const float* d_target,
  uint64_t n, float scale, float* d_grad, double* loss);

/* The speculation window of the exodus runs that follow in this process
 * (exodus_finder.c:183-213, 2331-2349: a batch that accepts nothing widens
 * the window by the floor, one that accepts divides it down, within
 * [floor, ceiling]). The reference fixes both at compile time and the
 * catalogue does not depend on them; this build's defaults are 8192 and
 * 32768, sized to fill the GPU. 1 <= floor <= ceiling <= 32768; (0, 0)
 * restores the defaults. Lets small inputs cross batch boundaries. Not
 * synchronised, like the reference's settings table. */
SIF_API int sif_amd_set_batch_window(uint32_t floor, uint32_t ceiling);
/* Whether later exodus runs in this process may list each rung's candidate
 * cells in the smoothing instead of writing the smoothed field (on by
 * default; power-of-two grids without SIF_FINDER_SEEDS_MINIMA). The
 * catalogue is the same either way; off, the grid is smoothed in place and
 * restored as the reference does. No reference counterpart. Not
 * synchronised. */
SIF_API int sif_amd_set_cell_lists(int on);
/* The device copy of the catalogue (positions, radii and the overlap index)
 * that exodus runs on this thread keep between runs: frees it, and sets the
 * number of voids it is next allocated for (0: the default, 65536). It grows
 * by doubling as runs need; a caller that knows the catalogue's size can
 * reserve it whole, and a small room makes a run grow it many times. The
 * catalogue does not depend on it. No reference counterpart. Call between
 * runs only. */
SIF_API int sif_amd_set_void_room(uint32_t voids);

/* --- the batched form (no reference counterpart: many independent boxes,
 * each the catalogue sif_finder_exodus_raw would give it) --- */

/* One box: its particles (device, n_particles each), its grid and mesh
 * resolutions (grid cells per side: any n whose prime factors are at most
 * 31; mesh cells per side at most 1024), its side and its radius ladder
 * (host, any order). Unweighted. */
typedef struct {
  const float* d_x;
  const float* d_y;
  const float* d_z;
  uint64_t n_particles;
  uint32_t n_cells;
  uint32_t mesh_cells;
  float box;
  const sif_real* radii;
  uint32_t n_radii;
} sif_amd_box_t;

/* How a batch runs; no field changes any bit of any catalogue. 0 in a field
 * keeps its default: workers 8 (boxes side by side, each on its own stream
 * pair and buffers, 1 to 64), cic_tile 8 (the CIC deposit's tile edge, 8 or
 * 16 nodes), split_waves 4096 (the wave count a growth launch of few
 * candidates splits their first passes towards, 64 to 65536), batch_floor
 * and batch_ceiling as sif_amd_set_batch_window. */
typedef struct {
  uint32_t workers;
  uint32_t cic_tile;
  uint32_t split_waves;
  uint32_t batch_floor;
  uint32_t batch_ceiling;
} sif_amd_tuning_t;

/* Every box's exodus catalogue (CIC deposit, density contrast, chain mesh,
 * sif__exodus_run of exodus_finder.c:1585-2492 over its ladder), each the
 * same bits as the box run alone: catalogues[b] receives box b's (the
 * caller frees each with sif_catalogue_free). tuning may be NULL. On an
 * error no catalogue is returned and the message names the box. */
SIF_API int sif_amd_exodus_batch(const sif_amd_box_t* boxes, uint32_t n_boxes,
  sif_real threshold, sif_real overlap_fraction, sif_option options,
  const sif_amd_tuning_t* tuning, sif_catalogue_t** catalogues);

/* Times a few tunings (boxes side by side, the growth split, the CIC tile)
 * on `probe` boxes spread over the sizes (0: 8), checks that each gives
 * the same catalogues, and returns the fastest. */
SIF_API int sif_amd_batch_autotune(const sif_amd_box_t* boxes, uint32_t n_boxes,
  sif_real threshold, sif_real overlap_fraction, sif_option options, uint32_t probe,
  sif_amd_tuning_t* best);

/* Per-rung smoothed fields as the finder scans them, for parity checks:
 * rung i (descending ladder) into d_fields + i * n_cells^3. */
SIF_API int sif_amd_exodus_fields(const float* d_delta, uint32_t n_cells,
  float box, const sif_real* radii, uint32_t n_radii, sif_option options,
  float* d_fields);

SIF_EXTERN_C_END

#endif
