/* Single-precision 3D real FFT for the oracle pipeline's CPU baseline.
 *
 * Test infrastructure (timed as the CPU baseline, never called by the
 * product). The reference smooths through FFTW's single-precision plans: the
 * forward r2c of the grid once, then per rung a copy of the spectrum times the
 * mode factors and a c2r, plane plans over x-planes and pencil plans along x
 * (src/math/fft.c:378-434, 703-888). FFTW is absent here, so this is a
 * transform of its own with the same structure:
 *
 *  - a pencil pass along x (the filter applied as the spectrum is read), then
 *    per x-plane the pencils along y and the real rows along z;
 *  - each 1D transform is an iterative radix-2^2 Cooley-Tukey (bit reversal,
 *    then the butterfly stages two at a time) run on a block of FB pencils at once,
 *    held column-wise with real and imaginary parts apart, so that every
 *    butterfly is a loop over FB contiguous floats the compiler vectorises;
 *  - twiddles exp(-2 pi i k / n) are tabulated once per length in double;
 *  - a real row of length n is one complex transform of length n/2
 *    (x[2m] + i x[2m+1]) and a split, both ways.
 *
 * Pinned by tests/test_oracle.py against numpy's pocketfft (round trip and
 * forward and inverse values). Power-of-two n >= 4 only. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define FB 16 /* pencils a block */

typedef struct {
  uint32_t n;
  float* wr; /* cos(2 pi k / n), k <= n/2 */
  float* wi; /* -sin(2 pi k / n) */
  uint32_t* rev;
} fplan_t;

static int fplan_make(fplan_t* p, uint32_t n) {
  p->n = n;
  p->wr = malloc(sizeof(float) * (n / 2 + 1));
  p->wi = malloc(sizeof(float) * (n / 2 + 1));
  p->rev = malloc(sizeof(uint32_t) * n);
  if (!p->wr || !p->wi || !p->rev)
    return -1;
  for (uint32_t k = 0; k <= n / 2; k++) {
    const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)n;
    p->wr[k] = (float)cos(a);
    p->wi[k] = (float)-sin(a);
  }
  uint32_t bits = 0;
  while ((1u << bits) < n)
    bits++;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < bits; b++)
      r |= ((i >> b) & 1u) << (bits - 1 - b);
    p->rev[i] = r;
  }
  return 0;
}

static void fplan_free(fplan_t* p) {
  free(p->wr);
  free(p->wi);
  free(p->rev);
}

/* FB transforms of length p->n in place: element i of pencil c at
 * re[i FB + c], im[i FB + c]. sign -1 forward, +1 inverse (unnormalised). */
__attribute__((target_clones("avx2", "default")))
static void block_fft(float* restrict re, float* restrict im, const fplan_t* p, int sign) {
  const uint32_t n = p->n;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t j = p->rev[i];
    if (i < j)
      for (int c = 0; c < FB; c++) {
        float t = re[i * FB + c];
        re[i * FB + c] = re[j * FB + c];
        re[j * FB + c] = t;
        t = im[i * FB + c];
        im[i * FB + c] = im[j * FB + c];
        im[j * FB + c] = t;
      }
  }
  /* the stages two at a time (lengths len and 2 len over the same four
   * elements: radix 2^2, half the passes over the block), the last alone
   * when log2 n is odd */
  uint32_t len = 2;
  for (; 2 * len <= n; len *= 4) {
    const uint32_t half = len >> 1, s1 = n / len, s2 = n / (2 * len);
    for (uint32_t i = 0; i < n; i += 2 * len)
      for (uint32_t k = 0; k < half; k++) {
        const float w1r = p->wr[k * s1], w1i = sign < 0 ? p->wi[k * s1] : -p->wi[k * s1];
        const float w2r = p->wr[k * s2], w2i = sign < 0 ? p->wi[k * s2] : -p->wi[k * s2];
        const float w3r = p->wr[(k + half) * s2],
                    w3i = sign < 0 ? p->wi[(k + half) * s2] : -p->wi[(k + half) * s2];
        float* restrict r0 = re + (i + k) * FB;
        float* restrict i0 = im + (i + k) * FB;
        float* restrict r1 = re + (i + k + half) * FB;
        float* restrict i1 = im + (i + k + half) * FB;
        float* restrict r2 = re + (i + k + len) * FB;
        float* restrict i2 = im + (i + k + len) * FB;
        float* restrict r3 = re + (i + k + len + half) * FB;
        float* restrict i3 = im + (i + k + len + half) * FB;
        for (int c = 0; c < FB; c++) {
          /* stage len: (0, 1) and (2, 3) by w1 */
          float xr = r1[c] * w1r - i1[c] * w1i, xi = r1[c] * w1i + i1[c] * w1r;
          const float a0r = r0[c] + xr, a0i = i0[c] + xi;
          const float a1r = r0[c] - xr, a1i = i0[c] - xi;
          xr = r3[c] * w1r - i3[c] * w1i;
          xi = r3[c] * w1i + i3[c] * w1r;
          const float a2r = r2[c] + xr, a2i = i2[c] + xi;
          const float a3r = r2[c] - xr, a3i = i2[c] - xi;
          /* stage 2 len: (0, 2) by w2, (1, 3) by w3 */
          xr = a2r * w2r - a2i * w2i;
          xi = a2r * w2i + a2i * w2r;
          r0[c] = a0r + xr;
          i0[c] = a0i + xi;
          r2[c] = a0r - xr;
          i2[c] = a0i - xi;
          xr = a3r * w3r - a3i * w3i;
          xi = a3r * w3i + a3i * w3r;
          r1[c] = a1r + xr;
          i1[c] = a1i + xi;
          r3[c] = a1r - xr;
          i3[c] = a1i - xi;
        }
      }
  }
  if (len <= n) {
    const uint32_t half = len >> 1, step = n / len;
    for (uint32_t i = 0; i < n; i += len)
      for (uint32_t k = 0; k < half; k++) {
        const float wr = p->wr[k * step];
        const float wi = sign < 0 ? p->wi[k * step] : -p->wi[k * step];
        float* restrict ar = re + (i + k) * FB;
        float* restrict ai = im + (i + k) * FB;
        float* restrict br = re + (i + k + half) * FB;
        float* restrict bi = im + (i + k + half) * FB;
        for (int c = 0; c < FB; c++) {
          const float xr = br[c] * wr - bi[c] * wi;
          const float xi = br[c] * wi + bi[c] * wr;
          br[c] = ar[c] - xr;
          bi[c] = ai[c] - xi;
          ar[c] = ar[c] + xr;
          ai[c] = ai[c] + xi;
        }
      }
  }
}

/* Pencils along one axis of the half spectrum [n][n][nz] (interleaved
 * complex): for a = 0 .. n-1 (the other axis) and kz, element i of the pencil
 * at (i si + a sa) nz + kz. With lut, the elements are multiplied as they
 * are read by the mode factor of (x, y, kz) = (i, a, kz) (the c2r's x pass).
 * Runs the pencils a_lo <= a < a_hi on the calling thread, in the block
 * buffers re and im (n FB floats each); with in and out the base of one
 * x-plane, si = 1 and sa = 0 it runs that plane's pencils along y. */
static void pencils(const float* in, float* out, uint32_t n, int sign, uint64_t si,
  uint64_t sa, const float* lut, const double* axis, const fplan_t* p, float* re, float* im,
  uint32_t a_lo, uint32_t a_hi) {
  const uint32_t nz = n / 2 + 1;
  for (uint32_t a = a_lo; a < a_hi; a++)
    for (uint32_t kz0 = 0; kz0 < nz; kz0 += FB) {
      const uint32_t w = nz - kz0 < FB ? nz - kz0 : FB;
      for (uint32_t i = 0; i < n; i++) {
        const uint64_t m0 = ((uint64_t)i * si + (uint64_t)a * sa) * nz + kz0;
        for (uint32_t c = 0; c < FB; c++) {
          float vr = 0.0f, vi = 0.0f;
          if (c < w) {
            vr = in[2 * (m0 + c)];
            vi = in[2 * (m0 + c) + 1];
            if (lut) {
              const float s = orc_mode_factor(lut, axis, n, i, a, kz0 + c);
              vr *= s;
              vi *= s;
            }
          }
          re[i * FB + c] = vr;
          im[i * FB + c] = vi;
        }
      }
      block_fft(re, im, p, sign);
      for (uint32_t i = 0; i < n; i++) {
        const uint64_t m0 = ((uint64_t)i * si + (uint64_t)a * sa) * nz + kz0;
        for (uint32_t c = 0; c < w; c++) {
          out[2 * (m0 + c)] = re[i * FB + c];
          out[2 * (m0 + c) + 1] = im[i * FB + c];
        }
      }
    }
}

/* n rows of one x-plane, half spectrum (rows of nz complex) to real rows
 * times norm. Per block of FB rows: Z[k] = A[k] + i B[k], k < M = n/2, with
 * A[k] = X[k] + conj X[M-k], B[k] = (X[k] - conj X[M-k]) exp(2 pi i k / n);
 * the inverse length-M transform of Z is x[2m] + i x[2m+1]. The imaginary
 * parts of X[0] and X[M] are taken as zero, as FFTW's c2r takes them. */
static void rows_c2r(const float* spec, float* out, uint32_t n, float norm, const fplan_t* pn,
  const fplan_t* pm, float* re, float* im) {
  const uint32_t M = n / 2, nz = M + 1;
  for (uint32_t r0 = 0; r0 < n; r0 += FB) {
    const uint32_t rows = n - r0 < FB ? n - r0 : FB;
    for (uint32_t c = 0; c < FB; c++) {
      if (c >= rows) {
        for (uint32_t k = 0; k < M; k++)
          re[k * FB + c] = im[k * FB + c] = 0.0f;
        continue;
      }
      const float* X = spec + 2 * (uint64_t)(r0 + c) * nz;
      for (uint32_t k = 0; k < M; k++) {
        const float xkr = X[2 * k], xki = k == 0 ? 0.0f : X[2 * k + 1];
        const float xmr = X[2 * (M - k)], xmi = k == 0 ? 0.0f : X[2 * (M - k) + 1];
        const float ar = xkr + xmr, ai = xki - xmi;
        const float dr = xkr - xmr, di = xki + xmi;
        /* exp(+2 pi i k / n) = (wr, -wi) of the length-n table */
        const float er = pn->wr[k], ei = -pn->wi[k];
        const float br = dr * er - di * ei, bi = dr * ei + di * er;
        re[k * FB + c] = ar - bi;
        im[k * FB + c] = ai + br;
      }
    }
    block_fft(re, im, pm, +1);
    for (uint32_t c = 0; c < rows; c++) {
      float* x = out + (uint64_t)(r0 + c) * n;
      for (uint32_t m = 0; m < M; m++) {
        x[2 * m] = re[m * FB + c] * norm;
        x[2 * m + 1] = im[m * FB + c] * norm;
      }
    }
  }
}

/* n real rows of one x-plane to their half spectra: z[m] = x[2m] + i x[2m+1],
 * Z its forward length-M transform, then for k = 0 .. M (Z[M] = Z[0])
 * X[k] = E[k] + exp(-2 pi i k / n) (-i) O[k], E = (Z[k] + conj Z[M-k]) / 2,
 * O = (Z[k] - conj Z[M-k]) / 2. */
static void rows_r2c(const float* in, float* spec, uint32_t n, const fplan_t* pn,
  const fplan_t* pm, float* re, float* im) {
  const uint32_t M = n / 2, nz = M + 1;
  for (uint32_t r0 = 0; r0 < n; r0 += FB) {
    const uint32_t rows = n - r0 < FB ? n - r0 : FB;
    for (uint32_t c = 0; c < FB; c++) {
      const float* x = in + (uint64_t)(r0 + c) * n;
      for (uint32_t m = 0; m < M; m++) {
        re[m * FB + c] = c < rows ? x[2 * m] : 0.0f;
        im[m * FB + c] = c < rows ? x[2 * m + 1] : 0.0f;
      }
    }
    block_fft(re, im, pm, -1);
    for (uint32_t c = 0; c < rows; c++) {
      float* X = spec + 2 * (uint64_t)(r0 + c) * nz;
      for (uint32_t k = 0; k <= M; k++) {
        const uint32_t a = k == M ? 0 : k, b = k == 0 ? 0 : M - k;
        const float zkr = re[a * FB + c], zki = im[a * FB + c];
        const float zmr = re[b * FB + c], zmi = im[b * FB + c];
        const float er = 0.5f * (zkr + zmr), ei = 0.5f * (zki - zmi);
        const float or_ = 0.5f * (zkr - zmr), oi = 0.5f * (zki + zmi);
        /* (-i) O = (oi, -or) */
        const float tr = oi, ti = -or_;
        const float wr = pn->wr[k], wi = pn->wi[k]; /* exp(-2 pi i k / n) */
        X[2 * k] = er + (wr * tr - wi * ti);
        X[2 * k + 1] = ei + (wr * ti + wi * tr);
      }
    }
  }
}

typedef struct {
  uint32_t n;
  fplan_t pn, pm;
} fft3f_t;

static int fft3f_make(fft3f_t* f, uint32_t n) {
  if (n < 4 || (n & (n - 1)))
    return -1;
  f->n = n;
  if (fplan_make(&f->pn, n) || fplan_make(&f->pm, n / 2))
    return -2;
  return 0;
}

static void fft3f_free(fft3f_t* f) {
  fplan_free(&f->pn);
  fplan_free(&f->pm);
}

/* Forward: real field [n][n][n] -> half spectrum [n][n][nz] (interleaved
 * complex float), FFTW's sign and order, unnormalised. */
int orc_fftf_r2c(const float* in, uint32_t n, float* spec) {
  fft3f_t f;
  if (fft3f_make(&f, n))
    return -1;
  const uint32_t nz = n / 2 + 1;
  const uint64_t plane = (uint64_t)n * nz;
#pragma omp parallel
  {
    float* re = malloc(sizeof(float) * n * FB);
    float* im = malloc(sizeof(float) * n * FB);
    /* per x-plane: rows along z, then pencils along y */
#pragma omp for schedule(static)
    for (uint32_t x = 0; x < n; x++) {
      rows_r2c(in + (uint64_t)x * n * n, spec + 2 * x * plane, n, &f.pn, &f.pm, re, im);
      pencils(spec + 2 * x * plane, spec + 2 * x * plane, n, -1, 1, 0, NULL, NULL, &f.pn, re,
        im, 0, 1);
    }
    /* pencils along x */
#pragma omp for schedule(static)
    for (uint32_t y = 0; y < n; y++)
      pencils(spec, spec, n, -1, n, 1, NULL, NULL, &f.pn, re, im, y, y + 1);
    free(re);
    free(im);
  }
  fft3f_free(&f);
  return 0;
}

/* One rung's smoothing: the spectrum (unchanged) times the mode factors of
 * lut and axis (orc_mode_factor; lut null for none), inverse transformed and
 * scaled by norm into out [n][n][n]. work holds the intermediate spectrum. */
int orc_fftf_filter_c2r(const float* spec, uint32_t n, const float* lut, const double* axis,
  float norm, float* work, float* out) {
  fft3f_t f;
  if (fft3f_make(&f, n))
    return -1;
  const uint32_t nz = n / 2 + 1;
  const uint64_t plane = (uint64_t)n * nz;
#pragma omp parallel
  {
    float* re = malloc(sizeof(float) * n * FB);
    float* im = malloc(sizeof(float) * n * FB);
    /* pencils along x, the factors applied as the spectrum is read */
#pragma omp for schedule(static)
    for (uint32_t y = 0; y < n; y++)
      pencils(spec, work, n, +1, n, 1, lut, axis, &f.pn, re, im, y, y + 1);
    /* per x-plane: pencils along y, then the real rows */
#pragma omp for schedule(static)
    for (uint32_t x = 0; x < n; x++) {
      float* pl = work + 2 * x * plane;
      pencils(pl, pl, n, +1, 1, 0, NULL, NULL, &f.pn, re, im, 0, 1);
      rows_c2r(pl, out + (uint64_t)x * n * n, n, norm, &f.pn, &f.pm, re, im);
    }
    free(re);
    free(im);
  }
  fft3f_free(&f);
  return 0;
}
