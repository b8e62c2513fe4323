"""Seeded Zel'dovich particle sets in a periodic box (the synthetic inputs of
the exodus benchmarks). The reference ships no generator, so this one is the
project's own:

* lattice q_i = (i + 1/2) L / N on each axis, particle order = lattice order
  (x slowest, z fastest);
* a Gaussian random field: unit white noise per lattice cell (numpy PCG64,
  seeded), transformed and shaped by sqrt(N^3 P(k) / V);
* P(k) = A k^n_s T^2(k) with the BBKS transfer function of shape
  Gamma = Omega_m h (q = k / Gamma), normalised to sigma_8 by top-hat
  quadrature at R = 8 Mpc/h;
* displacement Psi(k) = i k / k^2 delta(k) (Nyquist modes dropped), x = q + Psi
  folded into [0, L), stored float32.

The transforms use scipy.fft with all host threads; results are deterministic
for a seed. crc32() of the three coordinate arrays identifies a set.
"""
import zlib

import numpy as np
import scipy.fft as sfft
from scipy import integrate

OMEGA_M, H, N_S, SIGMA_8 = 0.31, 0.68, 0.965, 0.81


def bbks_transfer(k, gamma=OMEGA_M * H):
    q = np.maximum(k / gamma, 1e-30)
    return (np.log(1.0 + 2.34 * q) / (2.34 * q) *
            (1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25)


def _tophat(x):
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    m = x > 1e-4
    xm = x[m]
    out[m] = 3.0 * (np.sin(xm) - xm * np.cos(xm)) / xm ** 3
    return out


def amplitude():
    """A such that sigma(R = 8 Mpc/h) = SIGMA_8 for P = A k^n_s T^2."""
    def integrand(lnk):
        k = np.exp(lnk)
        return k ** 3 * k ** N_S * bbks_transfer(k) ** 2 * _tophat(8.0 * k) ** 2
    s2, _ = integrate.quad(integrand, np.log(1e-5), np.log(1e3), limit=400)
    s2 /= 2.0 * np.pi ** 2
    return SIGMA_8 ** 2 / s2


def power(k):
    return amplitude() * k ** N_S * bbks_transfer(k) ** 2


def generate(n_side, box, seed, workers=-1, progress=None):
    """(x, y, z) float32 arrays of n_side^3 particles in [0, box).
    progress(message), when given, is told as each transform is done."""
    say = progress or (lambda msg: None)
    n = int(n_side)
    L = float(box)
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    white = rng.standard_normal((n, n, n), dtype=np.float32)
    wk = sfft.rfftn(white, workers=workers)
    del white
    say(f"white noise transformed ({n}^3)")
    kf = 2.0 * np.pi / L
    kx = (sfft.fftfreq(n, 1.0 / n) * kf).astype(np.float32)
    kz = (sfft.rfftfreq(n, 1.0 / n) * kf).astype(np.float32)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k = np.sqrt(k2)
    A = amplitude()
    with np.errstate(divide="ignore", invalid="ignore"):
        pk = A * np.power(k.astype(np.float64), N_S) * bbks_transfer(k.astype(np.float64)) ** 2
    pk[0, 0, 0] = 0.0
    amp = np.sqrt(float(n) ** 3 * pk / L ** 3).astype(np.float32)
    del pk
    nyq = n // 2
    amp[nyq, :, :] = 0.0
    amp[:, nyq, :] = 0.0
    amp[:, :, nyq] = 0.0
    delta_k = wk * amp
    del wk, amp
    say("density modes shaped")
    with np.errstate(divide="ignore", invalid="ignore"):
        inv_k2 = np.where(k2 > 0, 1.0 / k2, 0.0).astype(np.float32)
    del k, k2
    q = ((np.arange(n, dtype=np.float64) + 0.5) * (L / n))
    out = []
    for axis in range(3):
        if axis == 0:
            kk = kx[:, None, None]
        elif axis == 1:
            kk = kx[None, :, None]
        else:
            kk = kz[None, None, :]
        psi = sfft.irfftn(1j * kk * inv_k2 * delta_k, s=(n, n, n), workers=workers)
        shape = [1, 1, 1]
        shape[axis] = n
        pos = (q.reshape(shape) + psi.astype(np.float64)).reshape(-1)
        del psi
        pos = np.mod(pos, L).astype(np.float32)
        pos[pos >= np.float32(L)] = 0.0
        out.append(np.ascontiguousarray(pos))
        say(f"displacement {'xyz'[axis]} done")
    return tuple(out)


def crc32(x, y, z):
    c = 0
    for a in (x, y, z):
        c = zlib.crc32(np.ascontiguousarray(a, dtype=np.float32).tobytes(), c)
    return c & 0xFFFFFFFF


def ladder(lo, hi, step=0.05):
    """Geometric radius ladder from lo to hi in steps of (1 + step), descending."""
    r = [float(lo)]
    while r[-1] * (1.0 + step) <= hi * (1.0 + 1e-9):
        r.append(r[-1] * (1.0 + step))
    return np.array(r[::-1], dtype=np.float32)
