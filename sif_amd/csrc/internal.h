// sif-amd internal: the stage functions the host driver strings together.
// Every one enqueues on the library stream; the caller synchronises.
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "common.h"
#include "overlap.h"
#include "sif/structures/catalogue.h"
#include "sif_amd.h"

namespace sif_amd {

// cic.hip
int check_in_box(const float* x, const float* y, const float* z, uint64_t n,
  float box, uint64_t* n_bad, uint64_t* n_edge, uint64_t* first_bad);
int cic_deposit(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t N, float box, float* values, int64_t* acc);
int grid_mean(const float* values, uint64_t total, double* mean);
// The CIC tile edge (8 or 16 nodes): a tuning that changes no bit.
uint32_t cic_tile_edge();
int set_cic_tile_edge(uint32_t t);
int density_contrast(float* values, uint64_t total, double mean);

// mesh.hip
int mesh_build(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t nc, float box, uint64_t* offsets, float* mx, float* my,
  float* mz, float* mw, uint64_t* orig, float* cw, double* total_weight);

// fft.hip
// Complex values per row of a spectrum or FFT work buffer: the n/2 + 1 modes,
// rounded up to 16 (a 128-byte line).
__host__ __device__ inline uint32_t spectrum_pitch(uint32_t n) {
  return ((n / 2 + 1) + 15u) & ~15u;
}
// Complex values per x-plane of a spectrum (n rows of the pitch, and a pad
// that keeps successive planes off the same memory channels).
uint64_t spectrum_plane(uint32_t n);
int fft_r2c(const float* values, uint32_t n, float2* spec, float2* work);
int fft_filter_c2r(const float2* spec, uint32_t n, float box, int filter,
  float r, float sigma, int window_order, int window_mode, float* out,
  float2* work);
// The inverse passes with a device LUT (indexed by |k|^2, filter_lut_size(n)
// entries) and window axis (n doubles); either may be null.
int fft_c2r_lut(const float2* spec, uint32_t n, const float* d_lut, const double* d_axis,
  float* out, float2* work, hipStream_t s);
// A smoothing whose field is only scanned need not be written: the rows
// pass lists each block's cells at or below the threshold instead, in cell
// order. Block b covers cells [b LIST_BLOCK, (b + 1) LIST_BLOCK) of the flat
// grid; its list holds count[b] entries, the cell's offset in the block
// (local) and the order-preserving key of its value (float_key), at
// [b LIST_BLOCK, b LIST_BLOCK + count[b]) of keys and local.
constexpr uint32_t LIST_BLOCK = 4096;
struct CellLists {
  uint32_t* keys = nullptr;
  uint16_t* local = nullptr;
  uint32_t* count = nullptr;
  uint64_t blocks = 0;
};
// Whether the rows pass can list cells for an n^3 grid (the power-of-two
// lengths of the register kernels, 16 to 1024).
bool fft_lists_supported(uint32_t n);
uint64_t list_blocks(uint32_t n);
// fft_c2r_lut with the field replaced by the lists of cells <= t.
int fft_c2r_lut_lists(const float2* spec, uint32_t n, const float* d_lut, const double* d_axis,
  float t, const CellLists& lists, float2* work, hipStream_t s);
uint64_t filter_lut_size(uint32_t n);
void build_filter_lut(int filter, float r, float sigma, float box, uint32_t n,
  std::vector<float>& lut);
void build_filter_lut_range(int filter, float r, float sigma, float box, uint32_t k2_begin,
  uint32_t k2_end, float* lut);
void build_window_axis(uint32_t n, int order, int mode, std::vector<double>& axis);

// scan.hip
// minima_n: 0 for every cell below the threshold, the grid's side for the
// local minima among them (SIF_FINDER_SEEDS_MINIMA).
int scan_candidates(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t minima_n, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count);
// The same candidates in flat order, unsorted; then the sort on its own.
// key_differ (optional) receives the bits on which the listed keys differ.
int scan_candidates_flat(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t minima_n, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count,
  uint32_t* key_differ = nullptr);
int sort_candidates(uint32_t* keys, uint32_t* flat, uint64_t n,
  uint32_t key_differ = 0xffffffffu);
// scan_candidates_flat (minima_n 0) from the lists a listing smoothing made
// of the same field: the listed cells whose mask bit is clear.
int scan_lists_flat(const CellLists& lists, const uint64_t* mask, uint64_t total,
  uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count, uint32_t* key_differ);
// The same scans in two steps, so that the host can work while the counting
// pass runs: scan_begin enqueues it and the copy of its results, scan_finish
// waits for them and enqueues the writing pass when the count fits cap. One
// job at a time per host thread.
struct ScanJob {
  const float* delta = nullptr;      // the field, when lists is null
  const CellLists* lists = nullptr;  // or a listing smoothing's lists
  const uint64_t* mask = nullptr;
  uint64_t total = 0;
  float t = 0.0f;
  uint32_t minima_n = 0;
  // The writing pass's buffers, when known as the scan begins: it is then
  // issued at once, and writes only if the count fits cap (checked on the
  // device), so that scan_finish need not wait for the count to issue it.
  uint32_t* keys = nullptr;
  uint32_t* flat = nullptr;
  uint64_t cap = 0;
  // results, after scan_finish
  uint64_t count = 0;
  uint32_t key_differ = 0xffffffffu;
  uint64_t listed = 0;
  bool begun = false;
};
int scan_begin(ScanJob& J);
int scan_finish(ScanJob& J, uint32_t* keys, uint32_t* flat, uint64_t cap);
int keys_to_delta(const uint32_t* keys, uint64_t n, float* out);

// grow.hip
// The catalogue as a rung begins, on the device: one CSR cell list per radius
// class (a class is scanned with its own largest radius, so the many small
// voids are not scanned at the reach of the few large ones), every list over
// the reference's 32^3 cells.
constexpr int VOID_CLASSES = 4;
struct VoidIndex {
  const uint32_t* cell_start;  // VOID_CLASSES * 32^3 + 1 entries, class-major
  // each void's (x, y, z, r), grouped by (class, reference cell): a scan
  // reads a void in one 16-byte load, with no id to look up first
  const float4* voids;
  // one bit a (class, cell): set when the cell lists a void of the class;
  // word (class 32 + wa) 32 + wb holds the row's 32 cells along z
  const uint32_t* occupied;
  uint32_t n;                  // voids
  uint32_t gen;                // the arrays' generation it was built at (host check)
  float inv_l;                 // (float)32 / box
  float max_r;                 // largest radius in the index
  float class_max_r[VOID_CLASSES];  // < 0 for an empty class
};
// A weighted mesh's tracer and cell weights (device) and its mean weight
// density, (4/3) pi total_weight / box^3 (exodus_finder.c:1639-1642).
struct MeshWeights {
  const float* w = nullptr;
  const float* cell = nullptr;
  float mass_factor = 0.0f;
};
// One mesh of a survey run, on the device: cell offsets, positions, and the
// tracer and cell weights (NULL on an unweighted mesh).
struct SurveyMesh {
  const uint64_t* offsets = nullptr;
  const float *x = nullptr, *y = nullptr, *z = nullptr;
  const float* w = nullptr;
  const float* cw = nullptr;
  uint32_t mesh_n = 0;
};
// --- survey (survey.hip) ---
// What an unobserved node reads as on the scanned field (utils.h:161-166):
// far above any threshold, never a candidate.
constexpr float SURVEY_OUTSIDE = 1e30f;
double survey_randoms_per_node(double m);
// The observed nodes of an nn^3 grid over `box` (one byte a node): those the
// n positions reach, closed over (2k + 1)^3 nodes.
int survey_occupancy(const float* x, const float* y, const float* z, uint64_t n, uint32_t nn,
  float box, uint8_t* occ, int32_t* k, double* lambda, uint64_t* marked, uint64_t* observed);
// data <- data / (alpha randoms) - 1 on observed nodes, SURVEY_OUTSIDE elsewhere.
int survey_contrast(float* data, const float* randoms, const uint8_t* occ, float alpha,
  uint64_t total, hipStream_t s);
// The observed fraction of each void's sphere and of its shell to 2r (host
// columns in and out).
int survey_footprint(const uint8_t* occ, uint32_t nn, float box, const float* cx,
  const float* cy, const float* cz, const float* r, uint64_t nv, float* inside, float* shell);
// The bounding box of n device positions.
int mesh_extent(const float* x, const float* y, const float* z, uint64_t n, float lo[3],
  float hi[3]);

// sif__exodus_find_radius_survey (exodus_rescale.c:2172-2317) for n centres
// against the data mesh D and the random mesh Rm, K = (1 + t) alpha; reasons
// as grow_batch's, 6 for no randoms inside r_search.
int grow_survey(const SurveyMesh& D, const SurveyMesh& Rm, float box, const float* cx,
  const float* cy, const float* cz, uint32_t n, float rmin, float r_search, float K,
  float* radius, uint32_t* reason, uint64_t* n_in);
// The wave count a small growth launch's split first pass aims at (a
// tuning that changes no bit).
uint32_t grow_split_waves();
int set_grow_split_waves(uint32_t waves);
int grow_batch(const uint64_t* offsets, const float* mx, const float* my,
  const float* mz, uint32_t mesh_n, float box, const float* cx, const float* cy,
  const float* cz, uint32_t n, float rmin, float r_search, float threshold,
  float vol_factor, float* radius, uint32_t* reason, uint64_t* n_in,
  uint64_t* algo_bytes = nullptr, const MeshWeights* weights = nullptr);
// SIF_FINDER_EXTEND_REACH (exodus_finder.c:2060-2072): the candidates whose
// growth ended beyond its reach, grown again at the levels first .. levels-1
// (core ext_rmin[e], reach ext_reach[e], host arrays) until none is; their
// radius, reason and count are overwritten in place.
int grow_extend(const uint64_t* offsets, const float* mx, const float* my, const float* mz,
  uint32_t mesh_n, float box, const float* cx, const float* cy, const float* cz, uint32_t n,
  const float* ext_rmin, const float* ext_reach, uint32_t first, uint32_t levels,
  float threshold, float vol_factor, float* radius, uint32_t* reason, uint64_t* n_in,
  const MeshWeights* weights);
int node_centres(const uint32_t* flat, uint32_t n, uint32_t grid_n, float l,
  float* cx, float* cy, float* cz);
// keep[k] and the refusal counts (masked, proxy, overlap) added to d_counts.
int prefilter(uint32_t* flat, uint32_t n, uint32_t grid_n, float l, float box,
  const uint64_t* mask, int32_t proxy, float radius, float f, float scale,
  const VoidIndex& V, uint32_t* keep, uint64_t* d_counts, uint8_t* why = nullptr);
// Of n prefiltered candidates (flat nodes, bit 31 aside, and their prefilter
// verdicts why: 0 kept, 1 masked, 2 proxy, 3 overlap), those whose node the
// mask now covers, counted by verdict into d_out[4] (added to).
int count_masked_since(const uint32_t* flat, const uint8_t* why, uint64_t n,
  const uint64_t* mask, uint64_t* d_out);
// The class-split CSR of nv device-resident voids: d_start has
// VOID_CLASSES * 32^3 + 1 entries, d_voids nv (each void's x, y, z, r).
int build_void_index(const float* vx, const float* vy, const float* vz, const float* vr,
  uint32_t nv, float box, uint32_t* d_start, float4* d_voids,
  uint32_t* d_occupied);
// Per grown sphere: its overlap state against the index (bits 0-1) and, in
// GROWN_POLE_BIT (overlap.h), whether the pole probe at the grown radius hits
// the mask.
int grown_overlap(const float* cx, const float* cy, const float* cz, const uint32_t* flat,
  const float* radius, uint32_t n, const VoidIndex& V, float max_r, float box, float f,
  float scale, const uint64_t* mask, uint32_t grid_n, float l, uint32_t* state);
// sif__mark_sphere of nv voids (radius times scale) into the device mask.
int paint_voids(const float* vx, const float* vy, const float* vz, const float* vr,
  uint32_t nv, float scale, float l, uint32_t grid_n, uint64_t* mask);

// primitives.hip compaction helpers used by the driver
int compact_u32(const uint32_t* in, const uint32_t* keep, uint64_t n,
  uint32_t* out, uint64_t* count);
// Two arrays compacted by the same keep flags with one scan.
int compact2_u32(const uint32_t* a, const uint32_t* b, const uint32_t* keep, uint64_t n,
  uint32_t* out_a, uint32_t* out_b, uint64_t* count);
// compact2_u32 in two steps: the compaction and the copy of its count to
// host memory enqueued, then the wait for the count. One at a time per host
// thread.
int compact2_u32_begin(const uint32_t* a, const uint32_t* b, const uint32_t* keep, uint64_t n,
  uint32_t* out_a, uint32_t* out_b);
int compact2_u32_finish(uint64_t* count);

// One exodus run on device-resident data (exodus.hip).
struct RunSpec {
  float* d_delta;
  uint32_t n;
  float box;
  const uint64_t* d_offsets;
  const float *d_mx, *d_my, *d_mz;
  // a weighted mesh: tracer and cell weights (device) and the summed weight;
  // d_mw NULL on an unweighted one
  const float* d_mw = nullptr;
  const float* d_cw = nullptr;
  double total_weight = 0.0;
  uint32_t mesh_n;
  uint64_t n_particles;
  std::vector<float> radii;  // descending
  float threshold, f;
  sif_option opt;
  bool restore;
  const sif_catalogue_t* prior;
  // a resumed run: the first rung to run, and the mismatch count so far
  uint32_t first_rung = 0;
  uint64_t mismatched0 = 0;
  // after each rung: rungs done, the catalogue's columns, the mismatch count
  // (added voids more than twice the rung that found them)
  std::function<void(uint32_t, const float*, const float*, const float*, const float*,
    uint64_t, uint64_t)> on_rung;
  // A survey run (exodus_finder.c:1592-1660): the randoms' CIC grid (device,
  // read only; d_delta then holds the data's CIC density, not a contrast),
  // their mesh, the observed nodes and alpha, the data weight over the
  // randoms'. The catalogue then carries the footprint columns.
  const float* d_random = nullptr;
  SurveyMesh rmesh;
  const uint8_t* d_occupied = nullptr;
  float alpha = 0.0f;
};
int exodus_run(const RunSpec& R, sif_catalogue_t** out_cat, uint64_t** n_in_out,
  sif_amd_rung_stats_t* stats_out);
int sorted_ladder(const sif_real* radii, uint32_t n_radii, std::vector<float>& out);
int check_options(sif_option opt);
// The growth batch window now in force (sif_amd_set_batch_window).
int exodus_batch_window(uint32_t* floor, uint32_t* ceiling);
int check_args(const sif_real* radii, uint32_t n_radii, sif_real threshold, sif_real f);
int check_options(sif_option opt);
// exodus_finder.c:1051-1106: a weighted mesh the rescaling can use -- a
// cell table, every weight finite and non-negative (by its bits), a
// positive total. The host form reads host weights, the device form device
// ones.
int check_weights_host(const float* w, uint64_t n, const float* cell_w, double total);
// exodus_finder.c:1114-1137 for a survey: seeds on their nodes only (invalid
// otherwise), then check_options.
int survey_check_options(sif_option opt);
// The survey's checks on its meshes (weights, totals, padding) and its inputs
// to R: the observed nodes (from the random mesh), alpha, the random grid and
// mesh. R's grid, data mesh and box are set before.
int survey_prepare(RunSpec& R, const sif_amd_mesh_t& data, const sif_amd_mesh_t& randoms,
  const float* d_randoms, const sif_real* radii, uint32_t n_radii, sif_option opt);

// sif_profiles on device arrays (tracer_profiles.hip). The spec's geometry,
// bin edges and volumes, mean density and flags come from
// tracer_profile_spec() (profiles.c:332-352, 144-150, 408-470); the caller
// sets the arrays. vx NULL: no velocities; dens or vel NULL: not written.
// Rows of voids of non-positive radius are left as they are.
struct TracerProfileSpec {
  const uint64_t* offsets = nullptr;
  const float *mx = nullptr, *my = nullptr, *mz = nullptr, *mw = nullptr;
  const float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  uint32_t mesh_n = 0;
  float box = 0.0f, cell_length = 0.0f;
  int periodic = 1;
  const float *cx = nullptr, *cy = nullptr, *cz = nullptr, *rv = nullptr;
  uint32_t n_voids = 0;
  float ext = 0.0f, bin_width = 0.0f;
  uint32_t n_bins = 0;
  int differential = 0, vel_weighted = 0;
  float mean_dens = 1.0f;
  std::vector<float> edges, vols;  // n_bins + 1 edges; n_bins volumes, in void radii
  const float* bin_vols = nullptr;
  float *dens = nullptr, *vel = nullptr;
};
void tracer_profile_spec(TracerProfileSpec* S, uint32_t mesh_n, float box, double total_weight,
  float ext, uint32_t n_bins, sif_option opt, bool weighted, bool want_vel);
int tracer_profiles(const TracerProfileSpec& S);
int check_weights_device(const float* d_w, uint64_t n, const float* d_cell_w, double total);

}  // namespace sif_amd
