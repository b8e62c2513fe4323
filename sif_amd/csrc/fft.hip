// 3D real FFT for the exodus smoothing: r2c once per run, then per rung a
// filtered c2r, in FP32, written for CDNA4.
//
// Replaces the FFTW plane + pencil plans of src/math/fft.c:378-434, 790-888
// and the filter/window multiply of fft.c:698-788 of sif v0.5.0.
//
// Layout: real field [n][n][n] (z fastest), as FFTW's; spectrum
// [n][n][pitch] complex with the n/2+1 modes of each row first and the row
// pitch rounded up to 16 complex values (spectrum_pitch), so that the column
// passes' 16-column tiles start on a 128-byte line and every line they fetch
// is theirs whole. Each axis is one pass over memory: a workgroup loads
// a tile of 8192 complex values -- TK = 8192/n pencils of length n -- into LDS
// with coalesced reads (consecutive kz columns for the x and y passes, whole
// rows for the z pass), transforms every pencil, and writes the tile back the
// same way. The transform is a Stockham autosort FFT in registers (the
// formulation of Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC 2008): each of the n/8 threads of a
// pencil holds eight elements, does one radix-8 butterfly per stage (a
// radix-2 or radix-4 first stage when log2 n is not a multiple of 3), and
// exchanges through LDS between stages -- three stages for n = 512. Twiddles
// come from a table of exp(-2 pi i m / n) computed in double on the host.
//
// The per-mode smoothing factor (float)(LUT[k^2] * w(kx) w(ky) w(kz)) of
// fft.c:753-780 is applied as the spectrum is first read by the inverse x
// pass, so filtering costs no extra pass over memory. Power-of-two n from 16
// to 1024; other lengths through the mixed-radix kernels below.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"
#include "sif_amd.h"

namespace sif_amd {

constexpr int FFT_THREADS = 256;
constexpr int FFT_TILE = 8192;  // complex values per workgroup tile of the x and y passes
// The z pass's tile: fewer rows a workgroup, so that more workgroups share a
// CU and hide each other's load-transform-store chain (the x and y passes
// keep 16 columns of 128-byte lines; a row is contiguous at any count).
constexpr int ROWS_TILE = 2048;

// LDS layout of a tile: pencil p's element i at p * pitch + i + (i >> SH).
// The Stockham writes of the early stages stride the pencil by R or R NS
// elements, which would put a wavefront's 64 stores on a few of the 32
// banks; one padding slot every 2^SH elements spreads them. The pitch is
// 1 mod 16 slots at the lengths used, so the column phases (16 pencils side
// by side) stay conflict-free. SH = 3 where the tile and its twiddles then
// still fit two workgroups a CU (80 KB each), else 4.
constexpr int pad_shift(int len, int tw_entries, int tile = FFT_TILE) {
  return (tile / len) * (len + len / 8 + 1) * 8 + tw_entries * 8 <= 80 * 1024 ? 3 : 4;
}
template <int N, int SH>
struct Lay {
  static constexpr int pitch = N + (N >> SH) + 1;
  static __device__ __forceinline__ int at(int i) { return i + (i >> SH); }
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
// a * (-i) forward, a * (+i) inverse
template <bool INV>
__device__ __forceinline__ float2 rot(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// In-register DFTs of size 2, 4 and 8, exp(-+2 pi i / R) for INV false/true,
// decimation in frequency with natural-order outputs.
template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
  const float2 a0 = cadd(v[0], v[2]), a2 = csub(v[0], v[2]);
  const float2 a1 = cadd(v[1], v[3]), a3 = rot<INV>(csub(v[1], v[3]));
  v[0] = cadd(a0, a1);
  v[2] = csub(a0, a1);
  v[1] = cadd(a2, a3);
  v[3] = csub(a2, a3);
}
template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  const float c = 0.70710678118654752440f;
  const float2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
  const float2 a1 = cadd(v[1], v[5]), d5 = csub(v[1], v[5]);
  const float2 a2 = cadd(v[2], v[6]), d6 = csub(v[2], v[6]);
  const float2 a3 = cadd(v[3], v[7]), d7 = csub(v[3], v[7]);
  // the odd half times w8^1, w8^2, w8^3
  const float2 a5 = INV ? make_float2(c * (d5.x - d5.y), c * (d5.x + d5.y))
                        : make_float2(c * (d5.x + d5.y), c * (d5.y - d5.x));
  const float2 a6 = rot<INV>(d6);
  const float2 a7 = INV ? make_float2(-c * (d7.x + d7.y), c * (d7.x - d7.y))
                        : make_float2(c * (d7.y - d7.x), -c * (d7.x + d7.y));
  const float2 b0 = cadd(a0, a2), b2 = csub(a0, a2);
  const float2 b1 = cadd(a1, a3), b3 = rot<INV>(csub(a1, a3));
  const float2 b4 = cadd(a4, a6), b6 = csub(a4, a6);
  const float2 b5 = cadd(a5, a7), b7 = rot<INV>(csub(a5, a7));
  v[0] = cadd(b0, b1);
  v[4] = csub(b0, b1);
  v[2] = cadd(b2, b3);
  v[6] = csub(b2, b3);
  v[1] = cadd(b4, b5);
  v[5] = csub(b4, b5);
  v[3] = cadd(b6, b7);
  v[7] = csub(b6, b7);
}
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if (R == 2)
    dft2<INV>(v);
  else if (R == 4)
    dft4<INV>(v);
  else
    dft8<INV>(v);
}

// One Stockham stage of radix R over a pencil X (in LDS) of length N, after
// stages whose radices multiply to NS. Thread t of the pencil's N/8 threads
// does 8/R butterflies j = t + b N/8: inputs X[j + r N/R], twiddles
// w^(r (j mod NS) N / (NS R)), outputs Y[(j / NS) NS R + j mod NS + r NS].
template <int N, int SH, int R, int NS, bool INV>
__device__ __forceinline__ void stage_read(const float2* X, int t, float2* v) {
  constexpr int T = N / 8;
#pragma unroll
  for (int b = 0; b < 8 / R; b++)
#pragma unroll
    for (int r = 0; r < R; r++)
      v[b * R + r] = X[Lay<N, SH>::at(t + b * T + r * (N / R))];
}
template <int N, int SH, int R, int NS, bool INV>
__device__ __forceinline__ void stage_write(float2* Y, int t, float2* v, const float2* tw) {
  constexpr int T = N / 8;
#pragma unroll
  for (int b = 0; b < 8 / R; b++) {
    const int j = t + b * T;
    const int jm = j & (NS - 1);
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        float2 w = tw[jm * r * (N / (NS * R))];
        if (INV)
          w.y = -w.y;
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
    dft<R, INV>(&v[b * R]);
    const int d = (j / NS) * NS * R + jm;
#pragma unroll
    for (int r = 0; r < R; r++)
      Y[Lay<N, SH>::at(d + r * NS)] = v[b * R + r];
  }
}

// The radices of a length-N transform: a first stage of 2, 4 or 8 so that
// the rest are radix 8.
template <int N>
struct Plan {
  static constexpr int L = N == 8 ? 3 : N == 16 ? 4 : N == 32 ? 5 : N == 64 ? 6 :
                           N == 128 ? 7 : N == 256 ? 8 : N == 512 ? 9 : 10;
  static constexpr int R0 = (L % 3 == 0) ? 8 : (1 << (L % 3));
};

// Every pencil of the tile (TILE / N pencils, layout Lay<N, SH>) through
// all stages. The 256 threads serve the pencils in groups; the tile is
// transformed in place, a workgroup barrier between the reads and the writes
// of a stage.
template <int N, int SH, int R, int NS, bool INV, int TILE>
__device__ __forceinline__ void tile_stage(float2* L, const float2* tw) {
  constexpr int T = N / 8, PG = FFT_THREADS / T, PP = TILE / N / PG;
  static_assert(PP >= 1, "a tile holds at least one pencil a thread group");
  const int t = threadIdx.x % T, g = threadIdx.x / T;
  float2 v[PP][8];
#pragma unroll
  for (int p = 0; p < PP; p++)
    stage_read<N, SH, R, NS, INV>(L + (p * PG + g) * Lay<N, SH>::pitch, t, v[p]);
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PP; p++)
    stage_write<N, SH, R, NS, INV>(L + (p * PG + g) * Lay<N, SH>::pitch, t, v[p], tw);
  __syncthreads();
}

template <int N, int SH, bool INV, int NS, int TILE>
__device__ __forceinline__ void tile_rest(float2* L, const float2* tw) {
  if constexpr (NS < N) {
    tile_stage<N, SH, 8, NS, INV, TILE>(L, tw);
    tile_rest<N, SH, INV, NS * 8, TILE>(L, tw);
  }
}

template <int N, int SH, bool INV, int TILE = FFT_TILE>
__device__ __forceinline__ void tile_fft(float2* L, const float2* tw) {
  constexpr int R0 = Plan<N>::R0;
  tile_stage<N, SH, R0, 1, INV, TILE>(L, tw);
  tile_rest<N, SH, INV, R0, TILE>(L, tw);
}

template <int N>
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ g) {
  for (int i = threadIdx.x; i < N; i += FFT_THREADS)
    tw[i] = g[i];
}

// --- x and y passes: tiles of TK consecutive kz columns ---
//
// Element i of pencil (a, kz) sits at a * stride_a + i * stride_i + kz. With
// `lut`, the element read is multiplied by the smoothing factor of its mode;
// the pass index i is then the x index and a the y index.
template <int N, bool INV>
__global__ void __launch_bounds__(FFT_THREADS) cols_kernel(const float2* __restrict__ in,
  float2* __restrict__ out, int nz, uint64_t stride_a, uint64_t stride_i,
  const float2* __restrict__ tw_g, const float* __restrict__ lut,
  const double* __restrict__ axis) {
  constexpr int TK = FFT_TILE / N, SH = pad_shift(N, N);
  using LY = Lay<N, SH>;
  __shared__ float2 L[TK * LY::pitch];
  __shared__ float2 tw[N];
  const int tiles = (nz + TK - 1) / TK;
  const int a = blockIdx.x / tiles;
  const int kz0 = (blockIdx.x % tiles) * TK;
  const int cols = nz - kz0 < TK ? nz - kz0 : TK;
  load_twiddles<N>(tw, tw_g);
  const uint64_t base = (uint64_t)a * stride_a + kz0;
  constexpr int half = N / 2;
  // i = x index, a = y index, kz0 + c = z index (fft.c:753-780). The filter
  // factor depends on the x index through kx^2 and the window axis, which is
  // even (axis[i] = axis[N - i] bit for bit), so the tile's factors are built
  // once per |kx| (half the table gathers) into the LDS the values go to
  // next. Their table reads are issued first, so that their latency hides
  // under the tile's loads.
  constexpr int FT = (TK * (half + 1) + FFT_THREADS - 1) / FFT_THREADS;
  const int ky = a > half ? a - N : a;
  float fs[FT];
  if (lut) {
#pragma unroll
    for (int t = 0; t < FT; t++) {
      const int f = threadIdx.x + t * FFT_THREADS;
      const int c = f / (half + 1), kx = f % (half + 1);
      const int kz = kz0 + c < nz ? kz0 + c : nz - 1;
      const uint32_t k2 = (uint32_t)(kx * kx) + (uint32_t)(ky * ky) + (uint32_t)(kz * kz);
      fs[t] = f < TK * (half + 1) ? lut[k2] : 0.0f;
    }
  }
  // every load of the thread issued before the first is used
  constexpr int PER = N * TK / FFT_THREADS;
  float2 v[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    const int i = e / TK, c = e % TK;
    v[q] = c < cols ? in[base + (uint64_t)i * stride_i + c] : make_float2(0.0f, 0.0f);
  }
  if (lut) {
    float* F = reinterpret_cast<float*>(L);
    if (axis) {
      // the window's factors of the tile's |kx| (times its y factor) and kz
      // columns, staged in the LDS past the table: two coalesced reads a
      // thread instead of two gathers a mode
      constexpr int FB = (TK * (half + 1) + 1) / 2;  // float2 slots the table takes
      double* cxy = reinterpret_cast<double*>(L + FB);
      double* az = cxy + (half + 1);
      static_assert(FB * 8 + (half + 1 + TK) * 8 <= TK * LY::pitch * 8, "factors fit the tile");
      const double ay = axis[a];
      for (int kx = threadIdx.x; kx <= half; kx += FFT_THREADS)
        cxy[kx] = axis[kx] * ay;
      if (threadIdx.x < TK) {
        const int kz = kz0 + (int)threadIdx.x < nz ? kz0 + (int)threadIdx.x : nz - 1;
        az[threadIdx.x] = axis[kz];
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < FT; t++) {
        const int f = threadIdx.x + t * FFT_THREADS;
        if (f < TK * (half + 1)) {
          const int c = f / (half + 1), kx = f % (half + 1);
          F[f] = (float)((double)fs[t] * cxy[kx] * az[c]);
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < FT; t++) {
        const int f = threadIdx.x + t * FFT_THREADS;
        if (f < TK * (half + 1))
          F[f] = fs[t];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const int i = e / TK, c = e % TK;
      const float fs = F[c * (half + 1) + (i > half ? N - i : i)];
      v[q] = make_float2(v[q].x * fs, v[q].y * fs);
    }
    __syncthreads();  // the table's LDS takes the tile next
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    L[(e % TK) * LY::pitch + LY::at(e / TK)] = v[q];
  }
  __syncthreads();
  tile_fft<N, SH, INV>(L, tw);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    v[q] = L[(e % TK) * LY::pitch + LY::at(e / TK)];
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    const int i = e / TK, c = e % TK;
    if (c < cols)
      out[base + (uint64_t)i * stride_i + c] = v[q];
  }
}

// --- z pass: whole rows, as complex transforms of half the length ---
//
// A real row x of length N is transformed as the complex sequence
// z[m] = x[2m] + i x[2m+1] of length M = N/2, which halves the work of the
// pass. With Z the length-M DFT of z and W = exp(-2 pi i / N):
//
//   forward   X[k] = E[k] + W^k (-i) O[k],  k = 0 .. M,
//             E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / 2,
//             Z[M] taken as Z[0];
//   inverse   Z[k] = A[k] + i B[k],  k = 0 .. M-1,
//             A[k] = X[k] + conj X[M-k],  B[k] = (X[k] - conj X[M-k]) W^-k,
//             and the length-M inverse of Z is x[2m] + i x[2m+1]
//
// (X[k + M] = conj X[M - k] for a real row). Rows are TK2 = ROWS_TILE / M to
// a workgroup, laid out as Lay<M, SH> in LDS (element M of a row, X[M], fits
// in its padding); the twiddles of the length-M transform (W^2m) and the
// split's (W^k, k <= M) sit side by side in LDS.

template <int N>
__device__ __forceinline__ void load_row_twiddles(float2* tw, const float2* __restrict__ g) {
  constexpr int M = N / 2;
  for (int i = threadIdx.x; i < M; i += FFT_THREADS)
    tw[i] = g[2 * i];
  for (int i = threadIdx.x; i <= M; i += FFT_THREADS)
    tw[M + i] = g[i];
}

template <int N>
__global__ void __launch_bounds__(FFT_THREADS) r2c_rows_kernel(const float* __restrict__ in,
  float2* __restrict__ spec, uint64_t rows, uint32_t pitch, uint64_t plane,
  const float2* __restrict__ tw_g) {
  constexpr int M = N / 2, TK2 = ROWS_TILE / M, nz = M + 1;
  constexpr int SH = pad_shift(M, 2 * M + 1, ROWS_TILE);
  using LY = Lay<M, SH>;
  __shared__ float2 L[TK2 * LY::pitch];
  __shared__ float2 tw[2 * M + 1];
  load_row_twiddles<N>(tw, tw_g);
  const uint64_t row0 = (uint64_t)blockIdx.x * TK2;
  constexpr int PER = M * TK2 / FFT_THREADS;
  constexpr int PER_H = (nz * TK2 + FFT_THREADS - 1) / FFT_THREADS;
  const float2* in2 = reinterpret_cast<const float2*>(in);
  {
    float2 z[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const uint64_t row = row0 + e / M;
      z[q] = row < rows ? in2[row * M + e % M] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      L[(e / M) * LY::pitch + LY::at(e % M)] = z[q];
    }
  }
  __syncthreads();
  tile_fft<M, SH, false, ROWS_TILE>(L, tw);
  float2 v[PER_H];
#pragma unroll
  for (int q = 0; q < PER_H; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    v[q] = make_float2(0.0f, 0.0f);
    if (e < nz * TK2) {
      const int r = e / nz, k = e % nz;
      const float2 zk = L[r * LY::pitch + LY::at(k == M ? 0 : k)];
      const float2 zm = L[r * LY::pitch + LY::at(k == 0 ? 0 : M - k)];
      const float2 ev = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 od = make_float2(0.5f * (zk.x - zm.x), 0.5f * (zk.y + zm.y));
      const float2 w = tw[M + k];
      v[q] = cadd(ev, cmul(w, make_float2(od.y, -od.x)));
    }
  }
#pragma unroll
  for (int q = 0; q < PER_H; q++) {
    const int e = threadIdx.x + q * FFT_THREADS;
    const uint64_t row = row0 + e / nz;
    if (e < nz * TK2 && row < rows)
      spec[(row / N) * plane + (row % N) * pitch + e % nz] = v[q];
  }
}

// The Hermitian half row (n/2 + 1 modes) to the real row, scaled by `norm`
// (fft.c:832). With LIST the real rows are not written: the block's cells at
// or below t (TK2 rows of N, LIST_BLOCK cells) go to its list in cell order
// (CellLists), the candidate scan's first half done where the values are
// made.
template <int N, bool LIST>
__global__ void __launch_bounds__(FFT_THREADS) c2r_rows_kernel(const float2* __restrict__ spec,
  float* __restrict__ out, uint64_t rows, uint32_t pitch, uint64_t plane, float norm,
  const float2* __restrict__ tw_g, float t, uint32_t* __restrict__ lkeys,
  uint16_t* __restrict__ llocal, uint32_t* __restrict__ lcount) {
  constexpr int M = N / 2, TK2 = ROWS_TILE / M, nz = M + 1;
  constexpr int SH = pad_shift(M, 2 * M + 1, ROWS_TILE);
  using LY = Lay<M, SH>;
  __shared__ float2 L[TK2 * LY::pitch];
  __shared__ float2 tw[2 * M + 1];
  load_row_twiddles<N>(tw, tw_g);
  const uint64_t row0 = (uint64_t)blockIdx.x * TK2;
  constexpr int PER = M * TK2 / FFT_THREADS;
  constexpr int PER_H = (nz * TK2 + FFT_THREADS - 1) / FFT_THREADS;
  {
    float2 v[PER_H];
#pragma unroll
    for (int q = 0; q < PER_H; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const uint64_t row = row0 + e / nz;
      v[q] = e < nz * TK2 && row < rows ? spec[(row / N) * plane + (row % N) * pitch + e % nz]
                                        : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int q = 0; q < PER_H; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      if (e < nz * TK2)
        L[(e / nz) * LY::pitch + LY::at(e % nz)] = v[q];
    }
  }
  __syncthreads();
  // Z[k] and Z[M - k] from X[k] and X[M - k], each pair by one thread, in
  // place (X[M] sits past the last Z and is read by the pair k = 0 only)
  constexpr int PAIRS = M / 2 + 1;
  for (int e = threadIdx.x; e < PAIRS * TK2; e += FFT_THREADS) {
    const int r = e / PAIRS, k = e % PAIRS;
    float2* row = L + r * LY::pitch;
    const float2 xk = row[LY::at(k)], xm = row[LY::at(M - k)];
    // W^-k and W^-(M-k) = -W^k conjugated: the forward table conjugated
    const float2 wk = make_float2(tw[M + k].x, -tw[M + k].y);
    const float2 a = make_float2(xk.x + xm.x, xk.y - xm.y);
    const float2 b = cmul(make_float2(xk.x - xm.x, xk.y + xm.y), wk);
    const float2 zk = make_float2(a.x - b.y, a.y + b.x);
    if (k > 0 && k < M - k) {
      const float2 wm = make_float2(tw[2 * M - k].x, -tw[2 * M - k].y);
      const float2 a2 = make_float2(xm.x + xk.x, xm.y - xk.y);
      const float2 b2 = cmul(make_float2(xm.x - xk.x, xm.y + xk.y), wm);
      row[LY::at(M - k)] = make_float2(a2.x - b2.y, a2.y + b2.x);
    }
    row[LY::at(k)] = zk;
  }
  __syncthreads();
  tile_fft<M, SH, true, ROWS_TILE>(L, tw);
  if constexpr (!LIST) {
    float2 r[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const float2 z = L[(e / M) * LY::pitch + LY::at(e % M)];
      r[q] = make_float2(z.x * norm, z.y * norm);
    }
    float2* out2 = reinterpret_cast<float2*>(out);
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const uint64_t row = row0 + e / M;
      if (row < rows)
        out2[row * M + e % M] = r[q];
    }
  } else {
    // Thread element e holds cells 2e and 2e + 1 of the block, so the
    // block's cell order is (q, wave, lane, pair half). The values come out
    // of LDS into registers once. Each wave counts its cells at or below t
    // for every q by ballot (lane q keeps the count of step q), one wave
    // turns the (q, wave) counts into offsets, and each cell's place is its
    // run's offset plus the cells before it in the run. Every block is whole
    // when N^2 rows are a multiple of its rows (N >= 32), and then no cell
    // is tested for being in the grid.
    static_assert(PER <= 64 && PER * (FFT_THREADS / 64) <= 128, "one offset pair a lane");
    constexpr int WAVES = FFT_THREADS / 64;
    constexpr bool WHOLE = ((uint64_t)N * N) % TK2 == 0;
    __shared__ uint32_t run_off[PER * WAVES + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    float2 z[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const float2 w = L[(e / M) * LY::pitch + LY::at(e % M)];
      z[q] = make_float2(w.x * norm, w.y * norm);
    }
    auto inside = [&](int q) -> bool {
      return WHOLE || row0 + (threadIdx.x + q * FFT_THREADS) / M < rows;
    };
    uint32_t my_count = 0;  // lane q: the wave's cells of step q
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const bool in = inside(q);
      const uint64_t m0 = __ballot(in && z[q].x <= t), m1 = __ballot(in && z[q].y <= t);
      const uint32_t c = (uint32_t)(__popcll(m0) + __popcll(m1));
      my_count = lane == q ? c : my_count;
    }
    if (lane < PER)
      run_off[lane * WAVES + wave] = my_count;
    __syncthreads();
    if (wave == 0) {  // exclusive scan of the PER * WAVES run counts
      const int i0 = 2 * lane, i1 = 2 * lane + 1;
      const uint32_t c0 = i0 < PER * WAVES ? run_off[i0] : 0u;
      const uint32_t c1 = i1 < PER * WAVES ? run_off[i1] : 0u;
      uint32_t x = c0 + c1;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
          x += y;
      }
      const uint32_t ex = x - c0 - c1;
      if (i0 < PER * WAVES)
        run_off[i0] = ex;
      if (i1 < PER * WAVES)
        run_off[i1] = ex + c0;
      if (lane == 63)
        run_off[PER * WAVES] = x;
    }
    __syncthreads();
    const uint32_t my_off = lane < PER ? run_off[lane * WAVES + wave] : 0u;
    const uint64_t base = (uint64_t)blockIdx.x * LIST_BLOCK;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * FFT_THREADS;
      const bool in = inside(q);
      const bool b0 = in && z[q].x <= t, b1 = in && z[q].y <= t;
      const uint64_t m0 = __ballot(b0), m1 = __ballot(b1);
      uint32_t pos = (uint32_t)__builtin_amdgcn_readlane((int)my_off, q) +
                     (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below));
      if (b0) {
        lkeys[base + pos] = float_key(z[q].x);
        llocal[base + pos] = (uint16_t)(2 * e);
        pos++;
      }
      if (b1) {
        lkeys[base + pos] = float_key(z[q].y);
        llocal[base + pos] = (uint16_t)(2 * e + 1);
      }
    }
    if (threadIdx.x == 0)
      lcount[blockIdx.x] = run_off[PER * WAVES];
  }
}

// --- any length: mixed-radix Stockham with a plan chosen at run time ---
//
// Lengths n = 2^a 3^b 5^c ... whose prime factors are at most 31 (every
// multiple of 8 from 64 to 256, 100, 384, 1536 ...). The radices are 8, 4
// and 2 for the power of two, then each odd prime; the stage formulation is
// the one above (any order of radices gives the natural order). A tile of
// TK pencils (pitch n + 1) is transformed in two LDS buffers in turn, one
// butterfly per thread and step; radix 2, 4 and 8 use the butterflies above,
// the odd primes a direct DFT whose roots come from the length's twiddle
// table. Twiddles are read from global memory: tw[k * tws] is
// exp(-2 pi i k / L) for a transform of length L over a table of L * tws.

constexpr int GEN_MAX_STAGES = 24;
constexpr int GEN_MAX_PRIME = 31;
constexpr int GEN_LDS_BYTES = 64 * 1024;

struct GenPlan {
  int len;  // transform length
  int stages;
  int radix[GEN_MAX_STAGES];
};

// One stage of radix R over the tile's `pencils` pencils (length L, pitch
// `pitch`): X -> Y. Radix 2, 4 and 8 transform in registers; an odd prime's
// outputs out[m] = sum_r v[r] w^(r m) (w = exp(-+2 pi i / R), the table's
// entry (L / R) * tws) are formed and stored one at a time, in a loop kept
// rolled so that the kernels stay small.
template <int R, bool INV>
__device__ __forceinline__ void gen_stage(const float2* X, float2* Y, int L, int NS, int pitch,
  int pencils, const float2* __restrict__ tw, int tws) {
  const int nb = L / R;
  const int twmul = L / (NS * R);
  for (int b = threadIdx.x; b < pencils * nb; b += FFT_THREADS) {
    const int p = b / nb, j = b - p * nb;
    const float2* x = X + p * pitch;
    float2* y = Y + p * pitch;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++)
      v[r] = x[j + r * nb];
    const int jm = j % NS;
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        float2 w = tw[(jm * r * twmul) * tws];
        if (INV)
          w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    const int d = (j / NS) * NS * R + jm;
    if constexpr (R == 2 || R == 4 || R == 8) {
      dft<R, INV>(v);
#pragma unroll
      for (int r = 0; r < R; r++)
        y[d + r * NS] = v[r];
    } else {
      const int step = (L / R) * tws;
#pragma unroll 1
      for (int m = 0; m < R; m++) {
        float2 acc = v[0];
        int e = 0;  // (r m) mod R, stepped
#pragma unroll
        for (int r = 1; r < R; r++) {
          e += m;
          e = e >= R ? e - R : e;
          float2 w = tw[e * step];
          if (INV)
            w.y = -w.y;
          acc = cadd(acc, cmul(v[r], w));
        }
        y[d + m * NS] = acc;
      }
    }
  }
}

// The plan's stages over the tile; returns the buffer holding the result.
// Inlined into each kernel: no call (and no return address) in the stage
// loop's barriers' path.
template <bool INV>
__device__ __forceinline__ float2* gen_fft(float2* X, float2* Y, const GenPlan& P, int pitch,
  int pencils, const float2* __restrict__ tw, int tws) {
  int ns = 1;
  for (int s = 0; s < P.stages; s++) {
    const int R = P.radix[s];
    switch (R) {
    case 2: gen_stage<2, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 3: gen_stage<3, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 4: gen_stage<4, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 5: gen_stage<5, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 7: gen_stage<7, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 8: gen_stage<8, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 11: gen_stage<11, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 13: gen_stage<13, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 17: gen_stage<17, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 19: gen_stage<19, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 23: gen_stage<23, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    case 29: gen_stage<29, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    default: gen_stage<31, INV>(X, Y, P.len, ns, pitch, pencils, tw, tws); break;
    }
    __syncthreads();
    float2* t = X;
    X = Y;
    Y = t;
    ns *= R;
  }
  return X;
}

// x and y passes for any n: TK columns a tile (as cols_kernel), the filter
// and window fused into the first inverse pass.
template <bool INV>
__global__ void __launch_bounds__(FFT_THREADS) gen_cols_kernel(const float2* __restrict__ in,
  float2* __restrict__ out, int nz, int TK, uint64_t stride_a, uint64_t stride_i, GenPlan P,
  const float2* __restrict__ tw, const float* __restrict__ lut,
  const double* __restrict__ axis) {
  extern __shared__ float2 sm[];
  const int N = P.len, pitch = N + 1, half = N / 2;
  float2* X = sm;
  float2* Y = sm + TK * pitch;
  const int tiles = (nz + TK - 1) / TK;
  const int a = blockIdx.x / tiles;
  const int kz0 = (blockIdx.x % tiles) * TK;
  const int cols = nz - kz0 < TK ? nz - kz0 : TK;
  const uint64_t base = (uint64_t)a * stride_a + kz0;
  const int ky = a > half ? a - N : a;
  for (int e = threadIdx.x; e < N * TK; e += FFT_THREADS) {
    const int i = e / TK, c = e % TK;
    float2 v = c < cols ? in[base + (uint64_t)i * stride_i + c] : make_float2(0.0f, 0.0f);
    if (lut && c < cols) {  // fft.c:753-780
      const int kx = i > half ? i - N : i;
      const int kz = kz0 + c;
      float f = lut[(uint32_t)(kx * kx) + (uint32_t)(ky * ky) + (uint32_t)(kz * kz)];
      if (axis)
        f = (float)((double)f * (axis[i] * axis[a]) * axis[kz]);
      v = make_float2(v.x * f, v.y * f);
    }
    X[c * pitch + i] = v;
  }
  __syncthreads();
  const float2* R = gen_fft<INV>(X, Y, P, pitch, TK, tw, 1);
  for (int e = threadIdx.x; e < N * TK; e += FFT_THREADS) {
    const int i = e / TK, c = e % TK;
    if (c < cols)
      out[base + (uint64_t)i * stride_i + c] = R[c * pitch + i];
  }
}

// z pass for any n (n even: the half-length transform of the rows above,
// over plan P of length n / 2; n odd: the whole row, plan P of length n),
// TK2 rows a workgroup.
__global__ void __launch_bounds__(FFT_THREADS) gen_r2c_rows_kernel(const float* __restrict__ in,
  float2* __restrict__ spec, uint64_t rows, int n, int TK2, uint32_t pitch_s, uint64_t plane,
  GenPlan P, const float2* __restrict__ tw) {
  extern __shared__ float2 sm[];
  const bool even = (n & 1) == 0;
  const int L = P.len, pitch = L + 1, nz = n / 2 + 1;
  float2* X = sm;
  float2* Y = sm + TK2 * pitch;
  const uint64_t row0 = (uint64_t)blockIdx.x * TK2;
  for (int e = threadIdx.x; e < L * TK2; e += FFT_THREADS) {
    const uint64_t row = row0 + e / L;
    const int m = e % L;
    float2 z = make_float2(0.0f, 0.0f);
    if (row < rows)
      z = even ? reinterpret_cast<const float2*>(in + row * n)[m]
               : make_float2(in[row * n + m], 0.0f);
    X[(e / L) * pitch + m] = z;
  }
  __syncthreads();
  const float2* Z = gen_fft<false>(X, Y, P, pitch, TK2, tw, even ? 2 : 1);
  for (int e = threadIdx.x; e < nz * TK2; e += FFT_THREADS) {
    const int r = e / nz, k = e % nz;
    const uint64_t row = row0 + r;
    if (row >= rows)
      continue;
    float2 v;
    if (even) {  // the split above, W^k = tw[k]
      const float2 zk = Z[r * pitch + (k == L ? 0 : k)];
      const float2 zm = Z[r * pitch + (k == 0 ? 0 : L - k)];
      const float2 ev = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 od = make_float2(0.5f * (zk.x - zm.x), 0.5f * (zk.y + zm.y));
      v = cadd(ev, cmul(tw[k], make_float2(od.y, -od.x)));
    } else {
      v = Z[r * pitch + k];
    }
    spec[(row / n) * plane + (row % n) * pitch_s + k] = v;
  }
}

__global__ void __launch_bounds__(FFT_THREADS) gen_c2r_rows_kernel(const float2* __restrict__ spec,
  float* __restrict__ out, uint64_t rows, int n, int TK2, uint32_t pitch_s, uint64_t plane,
  float norm, GenPlan P, const float2* __restrict__ tw) {
  extern __shared__ float2 sm[];
  const bool even = (n & 1) == 0;
  const int L = P.len, pitch = L + 1, nz = n / 2 + 1;
  float2* X = sm;
  float2* Y = sm + TK2 * pitch;
  const uint64_t row0 = (uint64_t)blockIdx.x * TK2;
  // the half row into Y (n even: X[M] at index M, in the pitch's spare slot)
  for (int e = threadIdx.x; e < nz * TK2; e += FFT_THREADS) {
    const int r = e / nz, k = e % nz;
    const uint64_t row = row0 + r;
    Y[r * pitch + k] = row < rows ? spec[(row / n) * plane + (row % n) * pitch_s + k]
                                  : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < L * TK2; e += FFT_THREADS) {
    const int r = e / L, k = e % L;
    const float2* h = Y + r * pitch;
    float2 z;
    if (even) {  // Z[k] = A + i B, as in c2r_rows_kernel
      const float2 xk = h[k], xm = h[L - k];
      const float2 wk = make_float2(tw[k].x, -tw[k].y);
      const float2 A = make_float2(xk.x + xm.x, xk.y - xm.y);
      const float2 B = cmul(make_float2(xk.x - xm.x, xk.y + xm.y), wk);
      z = make_float2(A.x - B.y, A.y + B.x);
    } else {  // the Hermitian row whole: v(n - k) = conj v(k)
      z = k < nz ? h[k] : make_float2(h[n - k].x, -h[n - k].y);
    }
    X[r * pitch + k] = z;
  }
  __syncthreads();
  const float2* R = gen_fft<true>(X, Y, P, pitch, TK2, tw, even ? 2 : 1);
  for (int e = threadIdx.x; e < L * TK2; e += FFT_THREADS) {
    const int r = e / L, m = e % L;
    const uint64_t row = row0 + r;
    if (row >= rows)
      continue;
    const float2 z = R[r * pitch + m];
    if (even)
      reinterpret_cast<float2*>(out + row * n)[m] = make_float2(z.x * norm, z.y * norm);
    else
      out[row * n + m] = z.x * norm;
  }
}

// --- host side ---

namespace {

struct Twiddles {
  int n = 0;
  DeviceBuffer fwd;  // exp(-2 pi i m / n), m < n
};
thread_local Twiddles g_tw;
thread_local DeviceBuffer g_lut, g_axis;
thread_local std::vector<float> g_lut_host;
thread_local std::vector<double> g_axis_host;

int log2_exact(uint32_t n) {
  int l = 0;
  while ((1u << l) < n)
    l++;
  return (1u << l) == n ? l : -1;
}

int twiddles(uint32_t n) {
  if (g_tw.n == (int)n)
    return SIF_OK;
  std::vector<float2> f(n);
  for (uint32_t m = 0; m < n; m++) {
    const double ang = 2.0 * 3.14159265358979323846 * (double)m / (double)n;
    f[m] = make_float2((float)cos(ang), (float)-sin(ang));
  }
  int st = g_tw.fwd.reserve(n * sizeof(float2));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpy(g_tw.fwd.ptr, f.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  g_tw.n = (int)n;
  return SIF_OK;
}

// The three passes for one n, dispatched to the kernels built for it.
template <int N>
int pass_cols(bool inv, const float2* in, float2* out, uint64_t stride_a, uint64_t stride_i,
  const float* lut, const double* axis, hipStream_t s) {
  constexpr int TK = FFT_TILE / N, nz = N / 2 + 1;
  const int tiles = (nz + TK - 1) / TK;
  if (inv)
    cols_kernel<N, true><<<N * tiles, FFT_THREADS, 0, s>>>(in, out, nz, stride_a,
      stride_i, g_tw.fwd.as<float2>(), lut, axis);
  else
    cols_kernel<N, false><<<N * tiles, FFT_THREADS, 0, s>>>(in, out, nz, stride_a,
      stride_i, g_tw.fwd.as<float2>(), lut, axis);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

template <int N>
int pass_r2c_rows(const float* in, float2* spec) {
  constexpr int TK2 = ROWS_TILE / (N / 2);
  const uint64_t rows = (uint64_t)N * N;
  r2c_rows_kernel<N><<<(unsigned)((rows + TK2 - 1) / TK2), FFT_THREADS, 0, stream()>>>(in,
    spec, rows, spectrum_pitch(N), spectrum_plane(N), g_tw.fwd.as<float2>());
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

template <int N>
int pass_c2r_rows(const float2* spec, float* out, float norm, hipStream_t s,
  const CellLists* lists, float t) {
  constexpr int TK2 = ROWS_TILE / (N / 2);
  static_assert((uint64_t)TK2 * N == LIST_BLOCK, "a rows block is one list block");
  const uint64_t rows = (uint64_t)N * N;
  const unsigned blocks = (unsigned)((rows + TK2 - 1) / TK2);
  if (lists)
    c2r_rows_kernel<N, true><<<blocks, FFT_THREADS, 0, s>>>(spec, nullptr, rows,
      spectrum_pitch(N), spectrum_plane(N), norm, g_tw.fwd.as<float2>(), t, lists->keys,
      lists->local, lists->count);
  else
    c2r_rows_kernel<N, false><<<blocks, FFT_THREADS, 0, s>>>(spec, out, rows,
      spectrum_pitch(N), spectrum_plane(N), norm, g_tw.fwd.as<float2>(), 0.0f, nullptr,
      nullptr, nullptr);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

#define SIF_FFT_DISPATCH(n, call)  \
  switch (n) {                     \
  case 16: { constexpr int N = 16; return call; }     \
  case 32: { constexpr int N = 32; return call; }     \
  case 64: { constexpr int N = 64; return call; }     \
  case 128: { constexpr int N = 128; return call; }   \
  case 256: { constexpr int N = 256; return call; }   \
  case 512: { constexpr int N = 512; return call; }   \
  default: { constexpr int N = 1024; return call; }   \
  }

int gen_run_cols(uint32_t n, bool inv, const float2* in, float2* out, uint64_t stride_a,
  uint64_t stride_i, const float* lut, const double* axis, hipStream_t s);
int gen_rows(uint32_t n, bool inv, const float* rin, const float2* cin, float2* cout,
  float* rout, float norm, hipStream_t s);
bool native_pow2(uint32_t n);

int run_cols(uint32_t n, bool inv, const float2* in, float2* out, uint64_t stride_a,
  uint64_t stride_i, const float* lut, const double* axis, hipStream_t s) {
  if (!native_pow2(n))
    return gen_run_cols(n, inv, in, out, stride_a, stride_i, lut, axis, s);
  SIF_FFT_DISPATCH(n, pass_cols<N>(inv, in, out, stride_a, stride_i, lut, axis, s))
}
int run_r2c_rows(uint32_t n, const float* in, float2* spec) {
  if (!native_pow2(n))
    return gen_rows(n, false, in, nullptr, spec, nullptr, 0.0f, stream());
  SIF_FFT_DISPATCH(n, pass_r2c_rows<N>(in, spec))
}
int run_c2r_rows(uint32_t n, const float2* spec, float* out, float norm, hipStream_t s,
  const CellLists* lists = nullptr, float t = 0.0f) {
  if (!native_pow2(n))
    return gen_rows(n, true, nullptr, spec, nullptr, out, norm, s);
  SIF_FFT_DISPATCH(n, pass_c2r_rows<N>(spec, out, norm, s, lists, t))
}

// The radices of a length: 8s, a 4, a 2, then each odd prime up to 31.
bool gen_plan(int len, GenPlan* P) {
  P->len = len;
  P->stages = 0;
  int m = len;
  auto push = [&](int r) {
    if (P->stages < GEN_MAX_STAGES)
      P->radix[P->stages] = r;
    P->stages++;
  };
  while (m % 8 == 0) {
    push(8);
    m /= 8;
  }
  if (m % 4 == 0) {
    push(4);
    m /= 4;
  }
  if (m % 2 == 0) {
    push(2);
    m /= 2;
  }
  for (int q = 3; q <= GEN_MAX_PRIME; q += 2)
    while (m % q == 0) {
      push(q);
      m /= q;
    }
  return m == 1 && P->stages <= GEN_MAX_STAGES;
}

// Pencils a tile for the general kernels: up to `want`, two buffers within
// GEN_LDS_BYTES.
int gen_tile(int len, int want) {
  int t = want;
  while (t > 1 && 2 * t * (len + 1) * (int)sizeof(float2) > GEN_LDS_BYTES)
    t--;
  return 2 * t * (len + 1) * (int)sizeof(float2) <= GEN_LDS_BYTES ? t : 0;
}

bool native_pow2(uint32_t n) {
  return n >= 16 && n <= 1024 && (n & (n - 1)) == 0;
}

int check_n(uint32_t n, int* log2n) {
  *log2n = log2_exact(n);
  if (native_pow2(n))
    return SIF_OK;
  GenPlan P;
  if (n < 2 || !gen_plan((int)n, &P) || (n % 2 == 0 && !gen_plan((int)n / 2, &P)) ||
      gen_tile((int)n, 1) == 0) {
    error_set(SIF_ERR_UNSUPPORTED,
      "fft: n = %u cells per side (this build transforms n >= 2 whose prime factors are at "
      "most %d, n below 4096)", n, GEN_MAX_PRIME);
    return SIF_ERR_UNSUPPORTED;
  }
  return SIF_OK;
}

// The three passes of the general kernels.
int gen_run_cols(uint32_t n, bool inv, const float2* in, float2* out, uint64_t stride_a,
  uint64_t stride_i, const float* lut, const double* axis, hipStream_t s) {
  GenPlan P;
  gen_plan((int)n, &P);
  const int nz = (int)n / 2 + 1;
  const int tk = gen_tile((int)n, 16);
  const int tiles = (nz + tk - 1) / tk;
  const size_t lds = 2 * (size_t)tk * (n + 1) * sizeof(float2);
  if (inv)
    gen_cols_kernel<true><<<n * tiles, FFT_THREADS, lds, s>>>(in, out, nz, tk, stride_a,
      stride_i, P, g_tw.fwd.as<float2>(), lut, axis);
  else
    gen_cols_kernel<false><<<n * tiles, FFT_THREADS, lds, s>>>(in, out, nz, tk, stride_a,
      stride_i, P, g_tw.fwd.as<float2>(), lut, axis);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int gen_rows(uint32_t n, bool inv, const float* rin, const float2* cin, float2* cout,
  float* rout, float norm, hipStream_t s) {
  GenPlan P;
  const int len = n % 2 == 0 ? (int)n / 2 : (int)n;
  gen_plan(len, &P);
  int want = 8192 / (len + 1);
  want = want < 1 ? 1 : (want > 256 ? 256 : want);
  const int tk2 = gen_tile(len, want);
  const uint64_t rows = (uint64_t)n * n;
  const unsigned blocks = (unsigned)((rows + tk2 - 1) / tk2);
  const size_t lds = 2 * (size_t)tk2 * (len + 1) * sizeof(float2);
  if (inv)
    gen_c2r_rows_kernel<<<blocks, FFT_THREADS, lds, s>>>(cin, rout, rows, (int)n, tk2,
      spectrum_pitch(n), spectrum_plane(n), norm, P, g_tw.fwd.as<float2>());
  else
    gen_r2c_rows_kernel<<<blocks, FFT_THREADS, lds, s>>>(rin, cout, rows, (int)n, tk2,
      spectrum_pitch(n), spectrum_plane(n), P, g_tw.fwd.as<float2>());
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // namespace

// src/math/fft.c:521-568, in float, on the host (the same libm as the
// reference's own table), indexed by the integer |k|^2: entries k2_begin to
// k2_end - 1 of the table into lut[k2].
void build_filter_lut_range(int filter, float r, float sigma, float box, uint32_t k2_begin,
  uint32_t k2_end, float* lut) {
  const float factor = (float)(r * 2.0 * SIF_PI / box);
  const float factor2 = factor * factor;
  const float gfactor = (float)(sigma * 2.0 * SIF_PI / box);
  const float gfactor2 = gfactor * gfactor;
  for (uint32_t k2 = k2_begin; k2 < k2_end; k2++) {
    float v;
    if (k2 == 0) {
      v = 1.0f;
    } else if (filter == SIF_AMD_FILTER_TOP_HAT) {
      const float kr = factor * sqrtf((float)k2);
      v = kr > 1e-4f ? 3.0f * (sinf(kr) - kr * cosf(kr)) / (kr * kr * kr) : 1.0f;
    } else if (filter == SIF_AMD_FILTER_GAUSSIAN) {
      v = expf(-0.5f * factor2 * (float)k2);
    } else {
      v = 1.0f;
    }
    if (sigma > 0.0f && k2 > 0)
      v *= expf(-0.5f * gfactor2 * (float)k2);
    lut[k2] = v;
  }
}

void build_filter_lut(int filter, float r, float sigma, float box, uint32_t n,
  std::vector<float>& lut) {
  lut.resize(filter_lut_size(n));
  build_filter_lut_range(filter, r, sigma, box, 0, (uint32_t)lut.size(), lut.data());
}

// fft.c:890-975: per-axis window correction, 1/W (full) or W/S (aliasing).
void build_window_axis(uint32_t n, int order, int mode, std::vector<double>& axis) {
  axis.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t h = n >> 1;
    const int32_t k = i > h ? (int32_t)i - (int32_t)n : (int32_t)i;
    const double x = (double)SIF_PI * (double)k / (double)n;
    const double s = (x < 1e-9 && x > -1e-9) ? 1.0 : sin(x) / x;
    double w = 1.0;
    for (int j = 0; j < order; j++)
      w *= s;
    double S = 1.0;
    if (mode) {
      const double s2 = sin(x) * sin(x);
      switch (order) {
      case 1:
        S = 1.0;
        break;
      case 2:
        S = 1.0 - 2.0 * s2 / 3.0;
        break;
      case 3:
        S = 1.0 - s2 + 2.0 * s2 * s2 / 15.0;
        break;
      default:
        S = 1.0 - 4.0 * s2 / 3.0 + 2.0 * s2 * s2 / 5.0 - 4.0 * s2 * s2 * s2 / 315.0;
      }
    }
    axis[i] = order == 0 ? 1.0 : (mode ? w / S : 1.0 / w);
  }
}

int fft_r2c(const float* values, uint32_t n, float2* spec, float2* work) {
  int log2n;
  int st = check_n(n, &log2n);
  if (st == SIF_OK)
    st = twiddles(n);
  if (st != SIF_OK)
    return st;
  const uint64_t nz = n / 2 + 1, pitch = spectrum_pitch(n), plane = spectrum_plane(n);
  const uint64_t rows = (uint64_t)n * n;
  timing_begin(SIF_AMD_STAGE_FFT_FORWARD);
  st = run_r2c_rows(n, values, work);
  // y pass: pencil a = x, i = y
  if (st == SIF_OK)
    st = run_cols(n, false, work, work, plane, pitch, nullptr, nullptr, stream());
  // x pass: pencil a = y, i = x
  if (st == SIF_OK)
    st = run_cols(n, false, work, spec, pitch, plane, nullptr, nullptr, stream());
  if (st != SIF_OK)
    return st;
  timing_end(SIF_AMD_STAGE_FFT_FORWARD);
  if (timing_active())  // the grid once, the spectrum once
    timing_bytes(SIF_AMD_STAGE_FFT_FORWARD, 4.0 * (double)rows * n + 8.0 * (double)rows * nz);
  return SIF_OK;
}

// The three inverse passes with a device-resident filter LUT (indexed by
// |k|^2; null for none) and window axis (null for none); the rows pass
// writes the field (out) or, with lists, the lists of its cells <= t.
static int c2r_lut_impl(const float2* spec, uint32_t n, const float* d_lut, const double* d_axis,
  float* out, const CellLists* lists, float t, float2* work, hipStream_t s) {
  int log2n;
  int st = check_n(n, &log2n);
  if (st == SIF_OK)
    st = twiddles(n);
  if (st != SIF_OK)
    return st;
  if (lists && !native_pow2(n)) {
    error_set(SIF_ERR_UNSUPPORTED, "fft: cell lists for n = %u (powers of two 16 to 1024)", n);
    return SIF_ERR_UNSUPPORTED;
  }
  const uint64_t nz = n / 2 + 1, pitch = spectrum_pitch(n), plane = spectrum_plane(n);
  const uint64_t rows = (uint64_t)n * n;
  const float norm = 1.0f / (float)((uint64_t)n * n * n);
  timing_begin(SIF_AMD_STAGE_SMOOTH, s);
  // x pass with the filter fused: pencil a = y, i = x
  st = run_cols(n, true, spec, work, pitch, plane, d_lut, d_axis, s);
  // y pass: pencil a = x, i = y
  if (st == SIF_OK)
    st = run_cols(n, true, work, work, plane, pitch, nullptr, nullptr, s);
  if (st == SIF_OK)
    st = run_c2r_rows(n, work, out, norm, s, lists, t);
  if (st != SIF_OK)
    return st;
  timing_end(SIF_AMD_STAGE_SMOOTH, s);
  // the spectrum once, and the field once (without lists; the lists' bytes
  // are known only on the device)
  if (timing_active())
    timing_bytes(SIF_AMD_STAGE_SMOOTH, 8.0 * (double)rows * nz + (lists ? 0.0 : 4.0 * (double)rows * n));
  return SIF_OK;
}

int fft_c2r_lut(const float2* spec, uint32_t n, const float* d_lut, const double* d_axis,
  float* out, float2* work, hipStream_t s) {
  return c2r_lut_impl(spec, n, d_lut, d_axis, out, nullptr, 0.0f, work, s);
}

int fft_c2r_lut_lists(const float2* spec, uint32_t n, const float* d_lut, const double* d_axis,
  float t, const CellLists& lists, float2* work, hipStream_t s) {
  return c2r_lut_impl(spec, n, d_lut, d_axis, nullptr, &lists, t, work, s);
}

bool fft_lists_supported(uint32_t n) {
  return native_pow2(n);
}

uint64_t list_blocks(uint32_t n) {
  return ((uint64_t)n * n * n + LIST_BLOCK - 1) / LIST_BLOCK;
}

// n rows of the pitch and three 128-byte lines more: successive planes then
// start on different memory channels, and the x pass, whose pencils stride
// by a plane, streams faster (512^3: that access pattern alone 281 -> 250 us).
uint64_t spectrum_plane(uint32_t n) {
  return (uint64_t)n * spectrum_pitch(n) + 48;
}

uint64_t filter_lut_size(uint32_t n) {
  const uint64_t h = n >> 1;
  return 3 * h * h + 1;
}

// The backward stage of fft.c:703-888 for one filter: spectrum (unchanged)
// -> smoothed real field, normalised by (1.0f / n^3) as fft.c:832 does.
int fft_filter_c2r(const float2* spec, uint32_t n, float box, int filter,
  float r, float sigma, int window_order, int window_mode, float* out,
  float2* work) {
  int log2n;
  int st = check_n(n, &log2n);
  if (st != SIF_OK)
    return st;
  const bool filtered = filter != SIF_AMD_FILTER_NONE;
  const bool windowed = window_order > 0 && filtered;
  if (filtered) {
    build_filter_lut(filter, r, sigma, box, n, g_lut_host);
    st = g_lut.reserve(g_lut_host.size() * sizeof(float));
    if (st != SIF_OK)
      return st;
    SIF_HIP(hipMemcpyAsync(g_lut.ptr, g_lut_host.data(), g_lut_host.size() * sizeof(float),
      hipMemcpyHostToDevice, stream()));
  }
  if (windowed) {
    build_window_axis(n, window_order, window_mode, g_axis_host);
    st = g_axis.reserve(n * sizeof(double));
    if (st != SIF_OK)
      return st;
    SIF_HIP(hipMemcpyAsync(g_axis.ptr, g_axis_host.data(), n * sizeof(double),
      hipMemcpyHostToDevice, stream()));
  }
  return fft_c2r_lut(spec, n, filtered ? g_lut.as<float>() : nullptr,
    windowed ? g_axis.as<double>() : nullptr, out, work, stream());
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

uint64_t sif_amd_spectrum_bytes(uint32_t n) {
  return (uint64_t)n * spectrum_plane(n) * 2 * sizeof(float);
}

uint32_t sif_amd_spectrum_pitch(uint32_t n) {
  return spectrum_pitch(n);
}

int sif_amd_fft_r2c(const float* values, uint32_t n, float* spectrum, void* work) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  st = fft_r2c(values, n, (float2*)spectrum, (float2*)work);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_fft_filter_c2r(const float* spectrum, uint32_t n, float box,
  int filter, float r, float sigma, int window_order, int window_mode,
  float* out, void* work) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  st = fft_filter_c2r((const float2*)spectrum, n, box, filter, r, sigma,
    window_order, window_mode, out, (float2*)work);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
