// Device code of the sphere growth (grow.hip): one wavefront per candidate.
// Kept in a header of its own so that the same text also compiles for the
// host under a wavefront emulator (tests/emu/), where it is checked lane by
// lane without a GPU.
#pragma once

#if defined(__HIPCC__)
#define SIF_GROW_HD __host__ __device__ __forceinline__
#else
#define SIF_GROW_HD inline
#endif

// The kernel's LDS is dynamic, sized per launch; the CPU emulator of the
// kernel supplies its own block buffer through this hook.
#ifndef GROW_DYNAMIC_LDS
#define GROW_DYNAMIC_LDS(name) extern __shared__ __align__(16) uint32_t name[]
#endif

namespace sif_amd {

// One wave per candidate, one wave per block: a block's LDS is its wave's,
// so occupancy follows each launch's LDS size wave by wave.
constexpr int GROW_WAVES = 1;
constexpr int MAX_BINS = 2048;
constexpr int REFINE_CAP = 1024;
constexpr int SUB_BINS = 64;

// exodus_rescale.c:793-800: the histogram's bin count, about 16 expected
// particles a bin over the annulus, between 64 and MAX_BINS. Host and device
// evaluate the same float operations (no contraction), so the host sizes the
// launch's LDS for exactly the count every wave computes.
SIF_GROW_HD uint32_t grow_bin_count(float r_search, float rmin, float vol_factor) {
  const float rs3 = r_search * r_search * r_search;
  const float rm3 = rmin * rmin * rmin;
  const float want = vol_factor * (rs3 - rm3) / 16.0f;
  if (!(want >= 64.0f))
    return 64;
  if (want < (float)MAX_BINS)
    return (uint32_t)want;
  return (uint32_t)MAX_BINS;
}

enum { CELL_CORE = 1, CELL_STRADDLES = 2, CELL_EDGE = 3, CELL_SHELL = 4 };

enum {
  R_OK = 0,
  R_EMPTY = 1,
  R_BEYOND = 2,
  R_BELOW_RUNG = 3,
  R_DEGENERATE = 4,
  R_OVERFLOW = 5
};

struct GrowParams {
  const uint64_t* offsets;
  const float* mx;
  const float* my;
  const float* mz;
  int32_t mesh_n;
  float box;
  float inv_l;
  const int4* tcell;    // dx, dy, dz, type
  const float2* tbound; // min_d2, max_d2
  uint32_t tcount;
  float rmin, r_search, threshold, vol_factor;
  const float* cx;
  const float* cy;
  const float* cz;
  uint32_t n_cand;
  // weighted tracers (exodus_rescale.c:1029-1037): per-tracer and per-cell
  // weights, and the mean weight density in place of vol_factor in every
  // density test; NULL and unused on an unweighted mesh
  const float* mw;
  const float* cw;
  float mass_factor;
  float* radius;
  uint32_t* reason;
  uint64_t* n_in;
  // LDS words a wave holds for its histogram (at least the bin count; three
  // words a bin when weighted: the count and the weight in double) and the
  // entries of its annulus cache (at least REFINE_CAP: the window buffer
  // shares it; a word an entry, two when weighted: distance and weight)
  uint32_t lds_bins, lds_cache;
  // The first pass split over `slices` waves a candidate (1: not split):
  // grow_bin_kernel leaves each wave's histogram (nb words) in part_bins and
  // its counts (PART_COUNTS words) in part_counts, slice-major per candidate;
  // grow_kernel then starts from their sums.
  uint32_t slices;
  uint32_t* part_bins;
  uint64_t* part_counts;
  // optional: the candidate's algorithmic bytes -- 8 per template cell whose
  // offsets it reads, 12 per particle of the cells whose particles it must
  // read once (the surface for the early refusal, every non-CORE cell past it)
  uint64_t* algo_bytes;
};

// A wave's LDS, carved from the block's: the histogram (lds_bins words),
// the annulus cache (lds_cache words) -- every annulus distance while they
// fit, so a candidate is resolved from LDS after one pass over its cells; the
// pass-per-window walk of a candidate whose annulus does not fit collects its
// window in the same words -- and the tables of the template chunk in flight.
constexpr uint32_t CACHE_MAX = 2048;
constexpr uint32_t LDS_TAIL_WORDS = 6 * 64 + SUB_BINS;

SIF_GROW_HD uint32_t grow_wave_words(uint32_t lds_bins, uint32_t lds_cache,
  bool weighted = false) {
  return lds_bins + lds_cache * (weighted ? 2u : 1u) + LDS_TAIL_WORDS;
}

// Words of a weighted histogram of nb bins: nb counts, then nb doubles.
SIF_GROW_HD uint32_t grow_weighted_bin_words(uint32_t nb) { return 3u * ((nb + 3u) & ~3u); }

// A launch's LDS words a wave: the rung's histogram, then as much annulus
// cache as keeps 16 waves on a CU (four a SIMD, what the kernel's registers
// allow) within its 160 KiB, and never less than the window buffer.
// cache_words > 0 asks for that cache instead.
SIF_GROW_HD void grow_lds_words(uint32_t nb, uint32_t cache_words, uint32_t* lds_bins,
  uint32_t* lds_cache) {
  *lds_bins = (nb + 3u) & ~3u;
  uint32_t cache = cache_words;
  if (cache == 0) {
    const uint32_t budget = (160u * 1024u / 16u - 256u) / 4u;
    const uint32_t used = *lds_bins + LDS_TAIL_WORDS;
    cache = (budget > used ? budget - used : 0u) & ~63u;
    cache = cache > CACHE_MAX ? CACHE_MAX : cache;
  }
  *lds_cache = cache < (uint32_t)REFINE_CAP ? (uint32_t)REFINE_CAP : cache;
}

// The weighted kernel's: the histogram's counts and weights, and a cache of
// REFINE_CAP (distance, weight) pairs -- what its pair sort holds in
// registers.
SIF_GROW_HD void grow_lds_words_weighted(uint32_t nb, uint32_t* lds_bins, uint32_t* lds_cache) {
  *lds_bins = grow_weighted_bin_words(nb);
  *lds_cache = (uint32_t)REFINE_CAP;
}

struct WaveLDS {
  uint32_t* bins;
  double* wbins;    // weighted: the bins' weights
  float* cache;
  float* cache_w;   // weighted: the weights of the cached distances
  float* buf;
  float* buf_w;
  uint32_t* t_start;
  uint32_t* t_scan;
  float* t_ex;
  float* t_ey;
  float* t_ez;
  uint32_t* t_type;
  uint32_t* sub;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// Every lane ends with the same sum: each butterfly step adds the same two
// values on both lanes of a pair, so the order is fixed by the lane layout.
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_max_i32(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(v, off, 64);
    if (lane >= off)
      v += y;
  }
  return v;
}

// Values every lane of the wave holds equally, moved to scalar registers so
// that the compiler sees them as uniform. Every loop bound and branch below
// that decides the wave's path goes through one of these: a condition the
// compiler cannot prove uniform makes the loop divergent, and it may then
// restructure it so that some lanes re-enter a cross-lane operation (a
// shuffle) without the lane that supplies its value.
__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni(uint64_t v) {
  return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | (uint64_t)uni((uint32_t)v);
}
__device__ __forceinline__ float unif(float v) { return __int_as_float(uni(__float_as_int(v))); }
__device__ __forceinline__ double unid(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  b = uni(b);
  double r;
  __builtin_memcpy(&r, &b, 8);
  return r;
}

// exodus_rescale.c:349-355, no contraction
__device__ __forceinline__ float dist2(float px, float py, float pz, float cx,
  float cy, float cz) {
  const float dx = __fsub_rn(px, cx), dy = __fsub_rn(py, cy), dz = __fsub_rn(pz, cz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// exodus_rescale.c:360-369
__device__ __forceinline__ uint32_t bin_of(float d2, float rc2, float inv_w, uint32_t nb) {
  const int32_t b = (int32_t)__fmul_rn(__fsub_rn(d2, rc2), inv_w);
  if (b < 0)
    return 0;
  if ((uint32_t)b >= nb)
    return nb - 1;
  return (uint32_t)b;
}

struct Query {
  float cx, cy, cz;
  int32_t ix, iy, iz;
};

// exodus_rescale.c:295-342: template entry -> particle range and the image
// of the centre it is measured against.
__device__ __forceinline__ void query_cell(const GrowParams& P, const Query& q,
  const int4 t, uint32_t* start, uint32_t* count, float* ex, float* ey, float* ez,
  uint64_t* cell) {
  const int32_t N = P.mesh_n;
  const float L = P.box;
  int32_t ix = q.ix + t.x;
  float cx = q.cx;
  if (ix < 0) {
    ix += N;
    cx = __fadd_rn(cx, L);
  } else if (ix >= N) {
    ix -= N;
    cx = __fsub_rn(cx, L);
  }
  int32_t iy = q.iy + t.y;
  float cy = q.cy;
  if (iy < 0) {
    iy += N;
    cy = __fadd_rn(cy, L);
  } else if (iy >= N) {
    iy -= N;
    cy = __fsub_rn(cy, L);
  }
  int32_t iz = q.iz + t.z;
  float cz = q.cz;
  if (iz < 0) {
    iz += N;
    cz = __fadd_rn(cz, L);
  } else if (iz >= N) {
    iz -= N;
    cz = __fsub_rn(cz, L);
  }
  const uint64_t flat = ((uint64_t)ix * N + (uint64_t)iy) * N + (uint64_t)iz;
  const uint64_t s = P.offsets[flat], e = P.offsets[flat + 1];
  *start = (uint32_t)s;
  *count = (uint32_t)(e - s);
  *cell = flat;
  *ex = cx;
  *ey = cy;
  *ez = cz;
}

// Steps of 64 particles a wave has in flight: their loads are issued
// together.
constexpr int PARTICLE_STEPS = 8;

// A template entry's fields as lane `l` holds them (l uniform), uniform over
// the wave. The CPU emulator of the kernel supplies its own, one rendezvous
// for the six.
#ifndef GROW_ENTRY_FIELDS
__device__ __forceinline__ void entry_fields(int l, uint32_t cnt, uint32_t start,
  uint32_t type, float ex, float ey, float ez, uint32_t* o_cnt, uint32_t* o_start,
  uint32_t* o_type, float* o_x, float* o_y, float* o_z) {
  *o_cnt = (uint32_t)__builtin_amdgcn_readlane((int)cnt, l);
  *o_start = (uint32_t)__builtin_amdgcn_readlane((int)start, l);
  *o_type = (uint32_t)__builtin_amdgcn_readlane((int)type, l);
  *o_x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ex), l));
  *o_y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ey), l));
  *o_z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ez), l));
}
#endif

// Walks the template 64 entries at a time, from chunk `first`, every
// `stride`-th chunk. `want(type, bound)` says whether an entry is visited;
// `whole(type, count, cell)` consumes a whole cell without reading its
// particles (returns true when it did); `visit(d2, w, type, valid)` sees
// every particle of the remaining cells, with its weight when WT (1
// otherwise). The wave takes one cell at a time, 64 of its particles a step
// (lane k the step's k-th; lanes past the cell's end have valid false), with
// the cell's start, image of the centre and type uniform over the wave.
// visit is called by every lane the same number of times, with `type`
// uniform, so a visitor may use wave-wide operations and branch on type.
template <bool WT, class Want, class Whole, class Visit>
__device__ __forceinline__ void for_particles(const GrowParams& P, const Query& q,
  WaveLDS& S, int lane, Want want, Whole whole, Visit visit, uint32_t first = 0,
  uint32_t stride = 1) {
  (void)S;
  for (uint32_t base = first * 64; base < P.tcount; base += stride * 64) {
    const uint32_t e = base + lane;
    uint32_t cnt = 0, start = 0, type = 0;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (e < P.tcount) {
      const int4 t = P.tcell[e];
      type = (uint32_t)t.w;
      if (want(type, P.tbound[e])) {
        uint64_t cell = 0;
        query_cell(P, q, t, &start, &cnt, &ex, &ey, &ez, &cell);
        if (cnt && whole(type, cnt, cell))
          cnt = 0;
      }
    }
    // the entries with particles to read, lowest lane first; the current
    // one's fields are read from its lane once
    uint64_t live = uni((uint64_t)__ballot(cnt != 0));
    int cur = live ? __builtin_ctzll(live) : 64;  // uniform: the entry read next
    uint32_t k = 0;                               // uniform: particles of it done
    uint32_t c_cnt = 0, c_start = 0, c_type = 0;
    float c_x = 0.0f, c_y = 0.0f, c_z = 0.0f;
    auto take_entry = [&]() {
      if (cur < 64)
        entry_fields(cur, cnt, start, type, ex, ey, ez, &c_cnt, &c_start, &c_type, &c_x, &c_y,
          &c_z);
    };
    take_entry();
    while (cur < 64) {
      uint32_t st_start[PARTICLE_STEPS], st_n[PARTICLE_STEPS], st_type[PARTICLE_STEPS];
      float st_x[PARTICLE_STEPS], st_y[PARTICLE_STEPS], st_z[PARTICLE_STEPS];
      // the next steps: (entry, first particle, particles), all uniform
#pragma unroll
      for (int u = 0; u < PARTICLE_STEPS; u++) {
        st_n[u] = 0;
        st_start[u] = 0;
        st_type[u] = 0;
        st_x[u] = st_y[u] = st_z[u] = 0.0f;
        if (cur < 64) {
          st_start[u] = c_start + k;
          st_n[u] = c_cnt - k < 64u ? c_cnt - k : 64u;
          st_type[u] = c_type;
          st_x[u] = c_x;
          st_y[u] = c_y;
          st_z[u] = c_z;
          k += 64;
          if (k >= c_cnt) {
            live &= live - 1;
            cur = live ? __builtin_ctzll(live) : 64;
            k = 0;
            take_entry();
          }
        }
      }
      float px[PARTICLE_STEPS], py[PARTICLE_STEPS], pz[PARTICLE_STEPS];
      float pw[PARTICLE_STEPS];
#pragma unroll
      for (int u = 0; u < PARTICLE_STEPS; u++) {
        // lanes past the step's end read its last particle again, unused
        const uint32_t off = (uint32_t)lane < st_n[u] ? (uint32_t)lane : st_n[u] - 1u;
        if (st_n[u]) {
          // uniform base, lane offset: one address register a step
          px[u] = (P.mx + st_start[u])[off];
          py[u] = (P.my + st_start[u])[off];
          pz[u] = (P.mz + st_start[u])[off];
          if constexpr (WT)
            pw[u] = (P.mw + st_start[u])[off];
          else
            pw[u] = 1.0f;
        } else {
          px[u] = py[u] = pz[u] = 0.0f;
          pw[u] = 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < PARTICLE_STEPS; u++) {
        if (st_n[u]) {
          const bool valid = (uint32_t)lane < st_n[u];
          const float d2 = valid ? dist2(px[u], py[u], pz[u], st_x[u], st_y[u], st_z[u]) : 0.0f;
          visit(d2, valid ? pw[u] : 0.0f, st_type[u], valid);
        }
      }
    }
  }
}

// Appends `d2` from every lane where `take` holds to buf[count ...] in lane
// order (slots from a ballot, no shared counter); `count` is uniform. Slots
// at or past `cap` are counted but not stored.
__device__ __forceinline__ void wave_append(float* buf, uint32_t cap, uint32_t& count,
  bool take, float d2, int lane) {
  const uint64_t m = __ballot(take);
  if (take) {
    const uint32_t slot = count + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (slot < cap)
      buf[slot] = d2;
  }
  count += (uint32_t)__popcll(m);
}

// Ascending bitonic sort of buf[0..n) in registers: element e = r * 64 + lane
// is register r of the lane; partners less than 64 apart are exchanged by
// shuffles, farther ones within the lane. The next power of two above n (at
// least 64) is R * 64.
template <int R>
__device__ void wave_sort_regs(float* buf, uint32_t n, int lane) {
  float v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    v[r] = e < n ? buf[e] : __int_as_float(0x7f800000);  // +inf sorts last
  }
#pragma unroll
  for (uint32_t k = 2; k <= (uint32_t)R * 64; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int jr = (int)(j / 64);
#pragma unroll
        for (int r = 0; r < R; r++) {
          if ((r & jr) == 0) {
            const bool asc = ((((uint32_t)r * 64 + lane) & k) == 0);
            const float a = v[r], b = v[r | jr];
            const bool swap = (a > b) == asc;
            v[r] = swap ? b : a;
            v[r | jr] = swap ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint32_t e = (uint32_t)r * 64 + lane;
          const float o = __shfl_xor(v[r], (int)j, 64);
          const bool asc = (e & k) == 0, lower = (e & j) == 0;
          const float mn = fminf(v[r], o), mx = fmaxf(v[r], o);
          v[r] = (lower == asc) ? mn : mx;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    if (e < n)
      buf[e] = v[r];
  }
}

// Ascending sort of buf[0..n), n <= 2048, by one wave.
__device__ __forceinline__ void wave_sort(float* buf, uint32_t n, int lane) {
  wave_sync();
  if (n <= 64)
    wave_sort_regs<1>(buf, n, lane);
  else if (n <= 128)
    wave_sort_regs<2>(buf, n, lane);
  else if (n <= 256)
    wave_sort_regs<4>(buf, n, lane);
  else if (n <= 512)
    wave_sort_regs<8>(buf, n, lane);
  else if (n <= 1024)
    wave_sort_regs<16>(buf, n, lane);
  else
    wave_sort_regs<32>(buf, n, lane);
  wave_sync();
}

// The same for n <= CACHE_MAX / 2, the windows of a cache at most half full.
__device__ __forceinline__ void wave_sort_half(float* buf, uint32_t n, int lane) {
  wave_sync();
  if (n <= 64)
    wave_sort_regs<1>(buf, n, lane);
  else if (n <= 128)
    wave_sort_regs<2>(buf, n, lane);
  else if (n <= 256)
    wave_sort_regs<4>(buf, n, lane);
  else if (n <= 512)
    wave_sort_regs<8>(buf, n, lane);
  else
    wave_sort_regs<16>(buf, n, lane);
  wave_sync();
}

// wave_append for (distance, weight) pairs: both go to the same slot.
__device__ __forceinline__ void wave_append2(float* buf, float* buf_w, uint32_t cap,
  uint32_t& count, bool take, float d2, float w, int lane) {
  const uint64_t m = __ballot(take);
  if (take) {
    const uint32_t slot = count + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (slot < cap) {
      buf[slot] = d2;
      buf_w[slot] = w;
    }
  }
  count += (uint32_t)__popcll(m);
}

// Ascending bitonic sort of the pairs (key[e], val[e]), e < n, by key and
// then by value: both are non-negative floats (a -0 weight is taken as +0),
// so their bit patterns, key above value, order as one 64-bit integer.
// Layout as wave_sort_regs.
template <int R>
__device__ void wave_sort_pairs_regs(float* key, float* val, uint32_t n, int lane) {
  uint64_t v[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    v[r] = e < n ? ((uint64_t)__float_as_uint(key[e]) << 32) |
                     (uint64_t)__float_as_uint(__fadd_rn(val[e], 0.0f))
                 : ~0ull;  // sorts last
  }
#pragma unroll
  for (uint32_t k = 2; k <= (uint32_t)R * 64; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int jr = (int)(j / 64);
#pragma unroll
        for (int r = 0; r < R; r++) {
          if ((r & jr) == 0) {
            const bool asc = ((((uint32_t)r * 64 + lane) & k) == 0);
            const uint64_t a = v[r], b = v[r | jr];
            const bool swap = (a > b) == asc;
            v[r] = swap ? b : a;
            v[r | jr] = swap ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint32_t e = (uint32_t)r * 64 + lane;
          const uint64_t o = __shfl_xor(v[r], (int)j, 64);
          const bool asc = (e & k) == 0, lower = (e & j) == 0;
          const uint64_t mn = v[r] < o ? v[r] : o, mx = v[r] < o ? o : v[r];
          v[r] = (lower == asc) ? mn : mx;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t e = (uint32_t)r * 64 + lane;
    if (e < n) {
      key[e] = __uint_as_float((uint32_t)(v[r] >> 32));
      val[e] = __uint_as_float((uint32_t)v[r]);
    }
  }
}

// Ascending sort of the pairs e < n, n <= REFINE_CAP, by one wave.
__device__ __forceinline__ void wave_sort_pairs(float* key, float* val, uint32_t n, int lane) {
  wave_sync();
  if (n <= 64)
    wave_sort_pairs_regs<1>(key, val, n, lane);
  else if (n <= 128)
    wave_sort_pairs_regs<2>(key, val, n, lane);
  else if (n <= 256)
    wave_sort_pairs_regs<4>(key, val, n, lane);
  else if (n <= 512)
    wave_sort_pairs_regs<8>(key, val, n, lane);
  else
    wave_sort_pairs_regs<16>(key, val, n, lane);
  wave_sync();
}

// The same for n <= REFINE_CAP / 2, the windows sorted after the cache.
__device__ __forceinline__ void wave_sort_pairs_half(float* key, float* val, uint32_t n,
  int lane) {
  wave_sync();
  if (n <= 64)
    wave_sort_pairs_regs<1>(key, val, n, lane);
  else if (n <= 128)
    wave_sort_pairs_regs<2>(key, val, n, lane);
  else if (n <= 256)
    wave_sort_pairs_regs<4>(key, val, n, lane);
  else
    wave_sort_pairs_regs<8>(key, val, n, lane);
  wave_sync();
}

// exodus_rescale.c:460-484: the weighted walk, outside in, over the pairs
// key/val[0..cnt) in ascending order, top_w the weight at or below the top
// of the range. The tracer at index k encloses top_w less the weight of
// every tracer from k up, itself included; it passes when that weight
// squared is at most K2 d2^3. The weight above each index comes from a
// suffix scan 64 pairs at a time, from the top; the first block (from the
// top) holding a pass holds the answer, its highest passing index. Returns
// that index (uniform), or -1 with *range_w the weight of the whole range.
__device__ __forceinline__ int weighted_walk(const float* key, const float* val, uint32_t cnt,
  double top_w, float K2, int lane, double* range_w) {
  double carry = 0.0;  // uniform: weight of the blocks above
  int found = -1;
  for (int32_t top = (int32_t)cnt; top > 0; top -= 64) {
    const int32_t k = top - 64 + lane;
    float d2 = 0.0f, w = 0.0f;
    if (k >= 0) {
      d2 = key[k];
      w = val[k];
    }
    double s = (double)w;  // weight from k to the top of the block
    for (int off = 1; off < 64; off <<= 1) {
      const double y = __shfl_down(s, off, 64);
      if (lane + off < 64)
        s += y;
    }
    int f = -1;
    if (k >= 0 && d2 != 0.0f) {
      const float w_in = (float)(top_w - (carry + s));
      const float lhs = __fmul_rn(w_in, w_in);
      const float rhs = __fmul_rn(K2, __fmul_rn(__fmul_rn(d2, d2), d2));
      if (lhs <= rhs)
        f = (int)k;
    }
    f = uni(wave_max_i32(f));
    if (f >= 0) {
      found = f;
      break;
    }
    carry += unid(__shfl(s, 0, 64));
  }
  *range_w = carry;
  return found;
}

enum { PC_WITHIN = 0, PC_SURFACE, PC_SHELL, PC_CORE, PC_ANNULUS, PART_COUNTS };

// exodus_rescale.c:711-865 for one slice of a candidate's template: the
// chunks s, s + slices, ... of candidate c = block / slices, s = block %
// slices. The same counts and histogram as grow_kernel's first pass over
// those chunks (unweighted, no annulus cache), written out for grow_kernel
// to sum: integers, so the sums do not depend on how the chunks are shared.
// LDS: lds_bins words of histogram, then the chunk tables.
__global__ void __launch_bounds__(64, 4) grow_bin_kernel(GrowParams P) {
  GROW_DYNAMIC_LDS(lds_words);
  const int lane = threadIdx.x & 63;
  WaveLDS S;
  S.bins = lds_words;
  S.wbins = nullptr;
  S.cache = S.cache_w = S.buf = S.buf_w = nullptr;
  uint32_t* tail = lds_words + P.lds_bins;
  S.t_start = tail;
  S.t_scan = tail + 64;
  S.t_ex = reinterpret_cast<float*>(tail + 128);
  S.t_ey = reinterpret_cast<float*>(tail + 192);
  S.t_ez = reinterpret_cast<float*>(tail + 256);
  S.t_type = tail + 320;
  S.sub = tail + 384;

  const uint32_t slices = P.slices;
  const uint32_t c = uni((uint32_t)blockIdx.x / slices);
  const uint32_t sl = uni((uint32_t)blockIdx.x - c * slices);
  if (c >= P.n_cand)
    return;
  const float rs2 = __fmul_rn(P.r_search, P.r_search);
  const float rc2 = __fmul_rn(P.rmin, P.rmin);
  const float span = __fsub_rn(rs2, rc2);
  const uint32_t nb = grow_bin_count(P.r_search, P.rmin, P.vol_factor);
  const float inv_w = __fdiv_rn((float)nb, span);
  Query q;
  q.cx = P.cx[c];
  q.cy = P.cy[c];
  q.cz = P.cz[c];
  q.ix = (int32_t)__fmul_rn(q.cx, P.inv_l);
  q.iy = (int32_t)__fmul_rn(q.cy, P.inv_l);
  q.iz = (int32_t)__fmul_rn(q.cz, P.inv_l);
  uint32_t* out_bins = P.part_bins + ((uint64_t)c * slices + sl) * nb;
  uint64_t* out_counts = P.part_counts + ((uint64_t)c * slices + sl) * PART_COUNTS;
  // grow_kernel refuses these before it reads the partial results
  const bool inside = q.cx >= 0.0f && q.cx <= P.box && q.cy >= 0.0f &&
                      q.cy <= P.box && q.cz >= 0.0f && q.cz <= P.box &&
                      q.ix <= P.mesh_n && q.iy <= P.mesh_n && q.iz <= P.mesh_n;
  if (!inside || !(span > 0.0f) || nb > P.lds_bins)
    return;
  for (uint32_t b = lane; b < nb; b += 64)
    S.bins[b] = 0;
  wave_sync();
  uint64_t within = 0, surface = 0, shell = 0, core = 0, annulus_n = 0;
  for_particles<false>(P, q, S, lane,
    [](uint32_t, float2) { return true; },
    [&](uint32_t type, uint32_t cnt, uint64_t) {
      if (type == CELL_CORE) {
        within += cnt;
        core += cnt;
        return true;
      }
      return false;
    },
    [&](float d2, float, uint32_t type, bool valid) {
      if (!valid)
        return;
      bool annulus;
      if (type == CELL_SHELL) {
        within++;
        shell++;
        annulus = true;
      } else {
        within += d2 <= rs2 ? 1u : 0u;
        surface++;
        core += d2 <= rc2 ? 1u : 0u;
        annulus = d2 > rc2 && d2 <= rs2;
      }
      if (annulus) {
        atomicAdd(&S.bins[bin_of(d2, rc2, inv_w, nb)], 1u);
        annulus_n++;
      }
    },
    sl, slices);
  wave_sync();
  within = wave_sum_u64(within);
  surface = wave_sum_u64(surface);
  shell = wave_sum_u64(shell);
  core = wave_sum_u64(core);
  annulus_n = wave_sum_u64(annulus_n);
  for (uint32_t b = lane; b < nb; b += 64)
    out_bins[b] = S.bins[b];
  if (lane == 0) {
    out_counts[PC_WITHIN] = within;
    out_counts[PC_SURFACE] = surface;
    out_counts[PC_SHELL] = shell;
    out_counts[PC_CORE] = core;
    out_counts[PC_ANNULUS] = annulus_n;
  }
}

template <bool WT>
__global__ void __launch_bounds__(64 * GROW_WAVES, 4) grow_kernel(GrowParams P) {
  GROW_DYNAMIC_LDS(lds_words);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  WaveLDS S;
  {
    uint32_t* base =
      lds_words + uni((uint32_t)w * grow_wave_words(P.lds_bins, P.lds_cache, WT));
    S.bins = base;
    // weighted: the doubles follow the counts, 16-byte aligned
    S.wbins = reinterpret_cast<double*>(base + ((P.lds_bins / 3u) & ~3u));
    S.cache = reinterpret_cast<float*>(base + P.lds_bins);
    S.cache_w = S.cache + P.lds_cache;
    S.buf = S.cache;
    S.buf_w = S.cache_w;
    uint32_t* tail = base + P.lds_bins + P.lds_cache * (WT ? 2u : 1u);
    S.t_start = tail;
    S.t_scan = tail + 64;
    S.t_ex = reinterpret_cast<float*>(tail + 128);
    S.t_ey = reinterpret_cast<float*>(tail + 192);
    S.t_ez = reinterpret_cast<float*>(tail + 256);
    S.t_type = tail + 320;
    S.sub = tail + 384;
  }

  const float rs2 = __fmul_rn(P.r_search, P.r_search);
  const float rc2 = __fmul_rn(P.rmin, P.rmin);
  const float span = __fsub_rn(rs2, rc2);
  // the histogram is sized by the tracer count either way (exodus_rescale.c:
  // 631-636); the density tests use the mean weight density when weighted
  const uint32_t nb = grow_bin_count(P.r_search, P.rmin, P.vol_factor);
  const float dens = WT ? P.mass_factor : P.vol_factor;
  const float expected = __fmul_rn(__fmul_rn(__fmul_rn(dens, P.r_search), P.r_search), P.r_search);
  const float K = __fmul_rn(__fadd_rn(P.threshold, 1.0f), dens);
  const float K2 = __fmul_rn(K, K);
  const float inv_w = __fdiv_rn((float)nb, span);
  const float bin_w = __fdiv_rn(span, (float)nb);

  // One candidate per wavefront, indexed from the block: the index, and so
  // every branch on it, is uniform over the wave by construction.
  const uint32_t c = blockIdx.x * GROW_WAVES + (uint32_t)uni(w);
  if (c >= P.n_cand)
    return;
  float radius = -1.0f;
  uint32_t reason = R_OK;
  uint64_t n_in_out = 0;
  uint64_t algo = 0;
  bool cached = false;
  uint64_t n_annulus = 0, n_core = 0;
  double total_w = 0.0;  // weighted: the search sphere's weight
  {
    Query q;
    q.cx = P.cx[c];
    q.cy = P.cy[c];
    q.cz = P.cz[c];
    q.ix = (int32_t)__fmul_rn(q.cx, P.inv_l);
    q.iy = (int32_t)__fmul_rn(q.cy, P.inv_l);
    q.iz = (int32_t)__fmul_rn(q.cz, P.inv_l);

    // A centre outside [0, box] would index past the mesh: refused.
    const bool inside = q.cx >= 0.0f && q.cx <= P.box && q.cy >= 0.0f &&
                        q.cy <= P.box && q.cz >= 0.0f && q.cz <= P.box &&
                        q.ix <= P.mesh_n && q.iy <= P.mesh_n && q.iz <= P.mesh_n;
    if (!inside) {
      reason = R_DEGENERATE;
    } else if (!(span > 0.0f)) {
      reason = R_DEGENERATE;
    } else if ((WT ? grow_weighted_bin_words(nb) : nb) > P.lds_bins ||
               P.lds_cache < (uint32_t)REFINE_CAP ||
               (WT && P.lds_cache != (uint32_t)REFINE_CAP)) {
      reason = R_OVERFLOW;  // a launch sized for fewer bins: never expected
    } else {
      // 1. early refusal, O(surface): exodus_rescale.c:711-782. The pass also
      // reads the SHELL cells (counted whole by the reference), keeps every
      // annulus distance in LDS while they fit, and builds the histogram of
      // the annulus (exodus_rescale.c:793-865, which reads the same
      // particles again) as it goes.
      uint64_t within = 0, surface = 0, shell = 0, core = 0, whole_cells = 0;
      // the first pass done by grow_bin_kernel: its slices' sums
      const bool binned = !WT && P.slices > 1;
      if (binned) {
        const uint64_t first = (uint64_t)c * P.slices;
        for (uint32_t b = lane; b < nb; b += 64) {
          uint32_t v = 0;
          for (uint32_t s = 0; s < P.slices; s++)
            v += P.part_bins[(first + s) * nb + b];
          S.bins[b] = v;
        }
        uint64_t mine = 0;  // lane k < PART_COUNTS: count k over the slices
        if (lane < PART_COUNTS)
          for (uint32_t s = 0; s < P.slices; s++)
            mine += P.part_counts[(first + s) * PART_COUNTS + lane];
        within = uni(__shfl(mine, PC_WITHIN, 64));
        surface = uni(__shfl(mine, PC_SURFACE, 64));
        shell = uni(__shfl(mine, PC_SHELL, 64));
        core = uni(__shfl(mine, PC_CORE, 64));
        n_annulus = uni(__shfl(mine, PC_ANNULUS, 64));
        wave_sync();
      } else {
        for (uint32_t b = lane; b < nb; b += 64) {
          S.bins[b] = 0;
          if constexpr (WT)
            S.wbins[b] = 0.0;
        }
        wave_sync();
      }
      // weighted (exodus_rescale.c:745-865): the early refusal's weight takes
      // whole CORE and SHELL cells from the cell table; the histogram's core
      // and annulus weights are the tracers' own, but for whole CORE cells.
      // Each lane sums its own share in double, in a fixed order.
      double within_w = 0.0, core_w = 0.0, shell_w = 0.0;
      uint32_t appended = 0;  // uniform
      if (!binned) {
        for_particles<WT>(P, q, S, lane,
          [](uint32_t, float2) { return true; },
          [&](uint32_t type, uint32_t cnt, uint64_t cell) {
            if constexpr (WT) {
              if (type == CELL_CORE || type == CELL_SHELL) {
                const double c = (double)P.cw[cell];
                within_w += c;
                if (type == CELL_CORE)
                  core_w += c;
                whole_cells++;
              }
            }
            if (type == CELL_CORE) {
              within += cnt;
              core += cnt;
              return true;
            }
            return false;
          },
          [&](float d2, float wt, uint32_t type, bool valid) {
            bool annulus = false;
            if (valid) {
              if (type == CELL_SHELL) {
                within++;
                shell++;
                annulus = true;
                if constexpr (WT)
                  shell_w += (double)wt;
              } else {
                within += d2 <= rs2 ? 1u : 0u;
                surface++;
                core += d2 <= rc2 ? 1u : 0u;
                annulus = d2 > rc2 && d2 <= rs2;
                if constexpr (WT) {
                  if (d2 <= rs2)
                    within_w += (double)wt;
                  if (d2 <= rc2)
                    core_w += (double)wt;
                  else if (annulus)
                    shell_w += (double)wt;
                }
              }
              if (annulus) {
                const uint32_t b = bin_of(d2, rc2, inv_w, nb);
                atomicAdd(&S.bins[b], 1u);
                // one wave owns these bins, so the order of the additions is
                // the wave's own program and lane order
                if constexpr (WT)
                  atomicAdd(&S.wbins[b], (double)wt);
              }
            }
            if constexpr (WT)
              wave_append2(S.cache, S.cache_w, P.lds_cache, appended, annulus, d2, wt, lane);
            else
              wave_append(S.cache, P.lds_cache, appended, annulus, d2, lane);
          });
        wave_sync();
        within = uni(wave_sum_u64(within));
        surface = uni(wave_sum_u64(surface));
        shell = uni(wave_sum_u64(shell));
        core = uni(wave_sum_u64(core));
        n_annulus = uni(appended);
      }
      if constexpr (WT) {
        within_w = unid(wave_sum_f64(within_w));
        // exodus_rescale.c:864: w_core + w_shell
        total_w = unid(wave_sum_f64(core_w)) + unid(wave_sum_f64(shell_w));
        whole_cells = uni(wave_sum_u64(whole_cells));
        // on a weighted mesh "empty" is "no weight inside"
        if (!(within_w > 0.0)) {
          reason = R_EMPTY;
        } else if (__fsub_rn(__fdiv_rn((float)within_w, expected), 1.0f) <= P.threshold) {
          reason = R_BEYOND;
        }
        algo = 8ull * P.tcount + 4ull * whole_cells +
               16ull * (surface + (reason == R_OK ? shell : 0ull));
      } else {
        if (within == 0) {
          reason = R_EMPTY;
        } else if (__fsub_rn(__fdiv_rn((float)within, expected), 1.0f) <= P.threshold) {
          reason = R_BEYOND;
        }
        algo = 8ull * P.tcount + 12ull * (surface + (reason == R_OK ? shell : 0ull));
      }
      // nothing cached when the first pass was split
      cached = !binned && n_annulus <= P.lds_cache;
      n_core = core;
    }

    if (reason == R_OK) {
      // 2. the histogram is the first pass's; every particle it holds is in
      // the annulus. A cached annulus of more than half the cache is sorted
      // whole now; a smaller one leaves room after it for any window.
      const bool cache_sorted = cached && n_annulus > P.lds_cache / 2;
      if (cache_sorted) {
        if constexpr (WT)
          wave_sort_pairs(S.cache, S.cache_w, (uint32_t)n_annulus, lane);
        else
          wave_sort(S.cache, (uint32_t)n_annulus, lane);
      }
      const uint64_t total = n_core + n_annulus;
      const bool empty = WT ? !(total_w > 0.0) : total == 0;
      const float total_f = WT ? (float)total_w : (float)total;
      if (empty) {
        reason = R_EMPTY;
      } else if (__fsub_rn(__fdiv_rn(total_f, expected), 1.0f) <= P.threshold) {
        reason = R_BEYOND;
      } else {
        // 3. skip flags: lane L owns bins [L*per, (L+1)*per), top down.
        const uint32_t per = (nb + 63) / 64;
        const uint32_t b0 = lane * per;
        const uint32_t b1 = b0 + per < nb ? b0 + per : nb;
        uint64_t mine = 0;
        for (uint32_t b = b0; b < b1; b++)
          mine += S.bins[b];
        // bins above this lane's range: suffix over higher lanes
        uint64_t suffix = mine;
        for (int off = 1; off < 64; off <<= 1) {
          const uint64_t y = __shfl_down(suffix, off, 64);
          if (lane + off < 64)
            suffix += y;
        }
        uint64_t above = suffix - mine;
        if constexpr (WT) {
          // the same in weight (exodus_rescale.c:915-957): a bin is skipped
          // on the weight below it; one with no weight is never opened. A
          // bin the walk must open is flagged by negating its weight.
          double mine_w = 0.0;
          for (uint32_t b = b0; b < b1; b++)
            mine_w += S.wbins[b];
          double suffix_w = mine_w;
          for (int off = 1; off < 64; off <<= 1) {
            const double y = __shfl_down(suffix_w, off, 64);
            if (lane + off < 64)
              suffix_w += y;
          }
          double above_w = __shfl_down(suffix_w, 1, 64);
          if (lane == 63)
            above_w = 0.0;
          for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--) {
            const double cw_b = S.wbins[b];
            if (!(cw_b > 0.0))
              continue;
            float hi = __fadd_rn(rc2, __fmul_rn((float)(b + 1), bin_w));
            if (b == (int32_t)nb - 1)
              hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
            const float n_min = (float)(total_w - above_w - cw_b);
            const float lhs = __fmul_rn(n_min, n_min);
            const float rhs = __fmul_rn(__fmul_rn(__fmul_rn(K2, hi), hi), hi);
            if (!(lhs > rhs))
              S.wbins[b] = -cw_b;
            above_w += cw_b;
          }
        } else {
          for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--) {
            const uint32_t cnt = S.bins[b];
            if (cnt == 0)
              continue;
            float hi = __fadd_rn(rc2, __fmul_rn((float)(b + 1), bin_w));
            if (b == (int32_t)nb - 1)
              hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
            const float n_min = (float)(total - above - cnt);
            const float lhs = __fmul_rn(n_min, n_min);
            const float rhs = __fmul_rn(__fmul_rn(__fmul_rn(K2, hi), hi), hi);
            if (!(lhs > rhs))
              S.bins[b] = cnt | 0x80000000u;  // a bin the walk must open
            above += cnt;
          }
        }
        wave_sync();

        // 4. walk: next flagged bin below `limit`, resolve a window, repeat.
        int32_t limit = (int32_t)nb;
        while (reason == R_OK && radius < 0.0f) {
          int best = -1;
          for (int32_t b = (int32_t)b1 - 1; b >= (int32_t)b0; b--)
            if (b < limit && (WT ? S.wbins[b] < 0.0 : (S.bins[b] & 0x80000000u) != 0u)) {
              best = b;
              break;
            }
          best = uni(wave_max_i32(best));
          if (best < 0) {
            reason = R_BELOW_RUNG;
            break;
          }
          // particles at or below the top of bin `best`
          uint64_t abv = 0;
          for (uint32_t b = (uint32_t)best + 1 + lane; b < nb; b += 64)
            abv += S.bins[b] & 0x7fffffffu;
          abv = uni(wave_sum_u64(abv));
          const uint64_t cur_n = total - abv;
          // weighted: the weight at or below the top of bin `best`
          double cur_w = 0.0;
          if constexpr (WT) {
            double abv_w = 0.0;
            for (uint32_t b = (uint32_t)best + 1 + lane; b < nb; b += 64)
              abv_w += fabs(S.wbins[b]);
            cur_w = total_w - unid(wave_sum_f64(abv_w));
          }
          // window [best - W + 1, best]: as many bins as fit the buffer
          uint32_t wcnt = 0;
          if (best - lane >= 0)
            wcnt = S.bins[best - lane] & 0x7fffffffu;
          const uint32_t wscan = wave_incl_scan_u32(wcnt, lane);
          const uint64_t fits = __ballot(best - lane >= 0 && wscan <= REFINE_CAP);
          int W = uni(fits ? __popcll(fits) : 0);
          const bool split = W == 0;  // the top bin alone overflows
          if (split)
            W = 1;
          const int32_t blo = best - W + 1;
          float hi = __fadd_rn(rc2, __fmul_rn((float)(best + 1), bin_w));
          if (best == (int32_t)nb - 1)
            hi = __fmul_rn(rs2, __fadd_rn(1.0f, 1e-6f));
          const float lo = __fadd_rn(rc2, __fmul_rn((float)blo, bin_w));
          const uint32_t win_total = uni(__shfl(wscan, W - 1, 64));

          if (cached) {
            // the window's distances in ascending order: the run of the
            // cache sorted whole that ends where the bins above `best` begin,
            // or copied out of the cache into its free tail and sorted there.
            // Walked outside in as below.
            uint32_t seg0;
            if (cache_sorted) {
              seg0 = (uint32_t)(n_annulus - abv - win_total);
            } else {
              float* win = S.cache + n_annulus;
              float* win_w = S.cache_w + n_annulus;
              uint32_t got = 0;  // uniform
              for (uint32_t k0 = 0; k0 < (uint32_t)n_annulus; k0 += 64) {
                const uint32_t k = k0 + (uint32_t)lane;
                float d2 = 0.0f, wt = 0.0f;
                bool take = false;
                if (k < (uint32_t)n_annulus) {
                  d2 = S.cache[k];
                  if constexpr (WT)
                    wt = S.cache_w[k];
                  const int32_t b = (int32_t)bin_of(d2, rc2, inv_w, nb);
                  take = b >= blo && b <= best;
                }
                if constexpr (WT)
                  wave_append2(win, win_w, win_total, got, take, d2, wt, lane);
                else
                  wave_append(win, win_total, got, take, d2, lane);
              }
              if (uni(got) != win_total) {
                reason = R_OVERFLOW;  // counts disagree: never expected
                break;
              }
              if constexpr (WT)
                wave_sort_pairs_half(win, win_w, win_total, lane);
              else
                wave_sort_half(win, win_total, lane);
              seg0 = (uint32_t)n_annulus;
            }
            int found = -1;
            if constexpr (WT) {
              double unused;
              found = weighted_walk(S.cache + seg0, S.cache_w + seg0, win_total, cur_w, K2,
                lane, &unused);
            } else {
              for (uint32_t k = lane; k < win_total; k += 64) {
                const float d2 = S.cache[seg0 + k];
                if (d2 == 0.0f)
                  continue;
                const uint64_t n_in = cur_n - win_total + k;
                const float rn = (float)n_in;
                const float lhs = __fmul_rn(rn, rn);
                const float rhs = __fmul_rn(K2, __fmul_rn(__fmul_rn(d2, d2), d2));
                if (lhs <= rhs)
                  found = (int)k;
              }
              found = uni(wave_max_i32(found));
            }
            if (found >= 0) {
              radius = __builtin_sqrtf(unif(S.cache[seg0 + found]));
              n_in_out = cur_n - win_total + (uint64_t)found;
            }
            if (radius < 0.0f)
              limit = blo;
            wave_sync();
            continue;
          }

          // sub-ranges of squared distance: 1 normally, SUB_BINS when split
          const int n_sub = split ? SUB_BINS : 1;
          const float inv_sub = n_sub > 1 ? __fdiv_rn((float)n_sub, __fsub_rn(hi, lo)) : 0.0f;
          uint32_t sub_count_mine = 0;  // lane s holds the count of sub-range s
          if (split) {
            S.sub[lane] = 0;
            wave_sync();
            for_particles<false>(P, q, S, lane,
              [&](uint32_t type, float2 bd) {
                return type != CELL_CORE && !(bd.y < lo) && !(bd.x > hi);
              },
              [](uint32_t, uint32_t, uint64_t) { return false; },
              [&](float d2, float, uint32_t type, bool valid) {
                if (!valid || (type != CELL_SHELL && !(d2 > rc2 && d2 <= rs2)))
                  return;
                const uint32_t b = bin_of(d2, rc2, inv_w, nb);
                if ((int32_t)b < blo || (int32_t)b > best)
                  return;
                int32_t sb = (int32_t)__fmul_rn(__fsub_rn(d2, lo), inv_sub);
                sb = sb < 0 ? 0 : (sb >= SUB_BINS ? SUB_BINS - 1 : sb);
                atomicAdd(&S.sub[sb], 1u);
              });
            wave_sync();
            sub_count_mine = S.sub[lane];
          }

          uint64_t top_n = cur_n;  // particles at or below the top of the range
          double top_w = cur_w;    // their weight, when weighted
          bool exhausted = false;
          for (int s = n_sub - 1; s >= 0 && radius < 0.0f && !exhausted; s--) {
            const uint32_t expect_cnt = uni(split ? __shfl(sub_count_mine, s, 64) : win_total);
            if (expect_cnt == 0)
              continue;
            if (expect_cnt > (uint32_t)REFINE_CAP) {
              reason = R_OVERFLOW;
              break;
            }
            uint32_t collected = 0;  // uniform
            for_particles<WT>(P, q, S, lane,
              [&](uint32_t type, float2 bd) {
                return type != CELL_CORE && !(bd.y < lo) && !(bd.x > hi);
              },
              [](uint32_t, uint32_t, uint64_t) { return false; },
              [&](float d2, float wt, uint32_t type, bool valid) {
                bool take = valid && (type == CELL_SHELL || (d2 > rc2 && d2 <= rs2));
                if (take) {
                  const uint32_t b = bin_of(d2, rc2, inv_w, nb);
                  take = (int32_t)b >= blo && (int32_t)b <= best;
                }
                if (take && split) {
                  int32_t sb = (int32_t)__fmul_rn(__fsub_rn(d2, lo), inv_sub);
                  sb = sb < 0 ? 0 : (sb >= SUB_BINS ? SUB_BINS - 1 : sb);
                  take = sb == s;
                }
                if constexpr (WT)
                  wave_append2(S.buf, S.buf_w, (uint32_t)REFINE_CAP, collected, take, d2, wt,
                    lane);
                else
                  wave_append(S.buf, (uint32_t)REFINE_CAP, collected, take, d2, lane);
              });
            wave_sync();
            const uint32_t got = uni(collected);
            if (got != expect_cnt) {
              reason = R_OVERFLOW;  // counts disagree: never expected
              break;
            }
            int found = -1;
            if constexpr (WT) {
              wave_sort_pairs(S.buf, S.buf_w, got, lane);
              double range_w = 0.0;
              found = weighted_walk(S.buf, S.buf_w, got, top_w, K2, lane, &range_w);
              if (found < 0)
                top_w -= range_w;
            } else {
              wave_sort(S.buf, got, lane);
              // outside in: element k encloses top_n - got + k particles
              for (uint32_t k = lane; k < got; k += 64) {
                const float d2 = S.buf[k];
                if (d2 == 0.0f)
                  continue;
                const uint64_t n_in = top_n - got + k;
                const float rn = (float)n_in;
                const float lhs = __fmul_rn(rn, rn);
                const float rhs = __fmul_rn(K2, __fmul_rn(__fmul_rn(d2, d2), d2));
                if (lhs <= rhs)
                  found = (int)k;
              }
              found = uni(wave_max_i32(found));
            }
            if (found >= 0) {
              const float d2 = unif(S.buf[found]);
              // correctly rounded: HIP's __fsqrt_rn is the 1-ulp native
              // instruction unless OCML rounded operations are enabled
              radius = __builtin_sqrtf(d2);
              n_in_out = top_n - got + (uint64_t)found;
            }
            top_n -= got;
            wave_sync();
          }
          if (radius < 0.0f && reason == R_OK)
            limit = blo;
          wave_sync();
        }
      }
    }
  }
  // Past every loop: only lane 0 writes the wave's result.
  if (lane == 0) {
    P.radius[c] = radius;
    P.reason[c] = reason;
    if (P.n_in)
      P.n_in[c] = n_in_out;
    if (P.algo_bytes)
      P.algo_bytes[c] = algo;
  }
}


}  // namespace sif_amd
