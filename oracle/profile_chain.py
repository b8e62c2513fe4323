"""TEST INFRASTRUCTURE ONLY -- an FP64 numpy restatement of the differentiable
profile chain: particle positions -> CIC density -> density contrast ->
Gaussian smoothing -> cumulative spherically averaged density profiles of
the smoothed grid -> sum of squared residuals, and its gradient with respect
to the positions. Used by tests/ and by bench.py's checks only; the product
(sif_amd/) never imports it.

What each step follows:

- CIC: grid_compute_cic (src/structures/grid.c:585-650, 679-746): u =
  fold(x) n / L, base node (int)u with u >= n -> 0, weights (1 - t, t) per
  axis, density per unit volume (x 1/l^3).
- contrast: sif_grid_to_density_contrast (grid.c:1177-1244), delta = rho /
  mean - 1.
- smoothing: the Gaussian filter table of fft_build_filter_lut
  (src/math/fft.c:521-568), W(k) = exp(-0.5 (R 2 pi / L)^2 |k|^2) with |k|^2
  in integer mode numbers, applied to the real-to-complex transform and
  inverted (fft.c:703-888), no window correction.
- profiles: sif_profiles (src/measure/profiles.c:354-673), density part,
  cumulative (profiles.c:332-350, 619-635), with the grid's nodes (g l) as
  the mass elements (1 + delta) rho_bar l^3, binned as profiles.c:202-223
  bins a tracer. The bins are taken in float32 with the device kernel's
  operations, so that both sides bin every node alike; sums are float64.

Pinned by tests/test_profile_chain.py: the profiles against a literal
per-node loop over the whole grid, the smoothing against the closed form of
a single Fourier mode, the CIC against the C oracle's deposit, the adjoint
against the forward (dot-product identity) and the gradient against FP64
central differences.
"""
import numpy as np


def fold(v, box):
    """field_internal.h:58-64."""
    v = np.where(v >= box, v - box, np.where(v < 0, v + box, v))
    return np.where(v >= box, 0.0, v)


def _cic_index(x, y, z, n, box):
    idx = n / box
    us = [fold(np.asarray(c, np.float64), box) * idx for c in (x, y, z)]
    base, frac = [], []
    for u in us:
        r = np.floor(u).astype(np.int64)
        base.append(np.where(r >= n, 0, r))
        frac.append(u - r)
    return base, frac


def cic(x, y, z, n, box):
    """Density per unit volume, float64 [n, n, n]."""
    (ix, iy, iz), (tx, ty, tz) = _cic_index(x, y, z, n, box)
    l = box / n
    rho = np.zeros(n * n * n)
    for a in (0, 1):
        wa = tx if a else 1.0 - tx
        ga = (ix + a) % n
        for b in (0, 1):
            wb = ty if b else 1.0 - ty
            gb = (iy + b) % n
            for c in (0, 1):
                wc = tz if c else 1.0 - tz
                gc = (iz + c) % n
                np.add.at(rho, (ga * n + gb) * n + gc, wa * wb * wc)
    return (rho / l ** 3).reshape(n, n, n)


def cic_adjoint(x, y, z, n, box, g):
    """d(sum g * rho) / d(x, y, z) for rho = cic(x, y, z)."""
    (ix, iy, iz), (tx, ty, tz) = _cic_index(x, y, z, n, box)
    l = box / n
    scale = (n / box) / l ** 3
    gf = np.asarray(g, np.float64).reshape(-1)
    wx = (1.0 - tx, tx)
    wy = (1.0 - ty, ty)
    wz = (1.0 - tz, tz)
    gx = np.zeros(len(ix))
    gy = np.zeros(len(ix))
    gz = np.zeros(len(ix))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                v = gf[(((ix + a) % n) * n + (iy + b) % n) * n + (iz + c) % n]
                gx += (v if a else -v) * wy[b] * wz[c]
                gy += (v if b else -v) * wx[a] * wz[c]
                gz += (v if c else -v) * wx[a] * wy[b]
    return gx * scale, gy * scale, gz * scale


def gaussian_table(n, box, r):
    """W over the real-to-complex half spectrum [n, n, n/2 + 1]."""
    k = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    f = r * 2.0 * np.pi / box
    return np.exp(-0.5 * f * f * k2)


def smooth(delta, box, r):
    n = delta.shape[0]
    return np.fft.irfftn(np.fft.rfftn(delta) * gaussian_table(n, box, r), s=delta.shape,
                         axes=(0, 1, 2))


def _bin_edges_volume(j, bw):
    r_outer = np.float32(j + 1) * np.float32(bw)
    return np.float32(4.0 / 3.0) * np.float32(np.pi) * (r_outer * r_outer * r_outer)


def node_bins(n, box, cx, cy, cz, rv, ext, n_bins):
    """Per centre: (flat node indices, bins) of the nodes of its profile, with
    the device kernel's float32 operations (cube, nearest image, r^2, sqrt,
    bin)."""
    f = np.float32
    l = f(box) / f(n)
    box32, half = f(box), f(box) * f(0.5)
    bw = f(ext) / f(n_bins)
    out = []
    for c in range(len(cx)):
        r_v = f(rv[c])
        if not (r_v > 0):
            out.append((np.zeros(0, np.int64), np.zeros(0, np.int64)))
            continue
        r_max = r_v * f(ext)
        r_max_sq = r_max * r_max
        inv_bin = f(1.0) / (r_v * bw)
        axes = []
        for cc in (f(cx[c]), f(cy[c]), f(cz[c])):
            lo = int(np.floor((cc - r_max) / l))
            hi = int(np.ceil((cc + r_max) / l))
            span = min(hi - lo + 1, n)
            g = (np.arange(lo, lo + span) % n).astype(np.int64)
            d = g.astype(f) * l - cc
            d = d - (box32 * (d > half).astype(f) - box32 * (d < -half).astype(f))
            axes.append((g, (d * d).astype(f)))
        (gx, dx2), (gy, dy2), (gz, dz2) = axes
        dxy = (dx2[:, None] + dy2[None, :]).astype(f)
        r2 = (dxy[:, :, None] + dz2[None, None, :]).astype(f)
        keep = r2 <= r_max_sq
        b = (np.sqrt(r2[keep]) * inv_bin).astype(np.uint32).astype(np.int64)
        flat = ((gx[:, None, None] * n + gy[None, :, None]) * n + gz[None, None, :])
        flat = np.broadcast_to(flat, r2.shape)[keep]
        ok = b < n_bins
        out.append((flat[ok], b[ok]))
    return out


def profiles(delta_s, bins, box, rv, ext, n_bins):
    """Cumulative density profiles [n_centres, n_bins], float64 sums."""
    n = delta_s.shape[0]
    l = np.float64(np.float32(box) / np.float32(n))
    bw = np.float32(ext) / np.float32(n_bins)
    d = np.asarray(delta_s, np.float64).reshape(-1)
    vols = np.array([_bin_edges_volume(j, bw) for j in range(n_bins)], np.float64)
    out = np.zeros((len(bins), n_bins))
    for c, (flat, b) in enumerate(bins):
        if not (np.float32(rv[c]) > 0):
            continue
        cnt = np.bincount(b, minlength=n_bins).astype(np.float64)
        s = np.bincount(b, weights=d[flat], minlength=n_bins)
        rv3 = np.float64(np.float32(rv[c])) ** 3
        out[c] = l ** 3 * np.cumsum(cnt + s) / (vols * rv3) - 1.0
    return out


def profiles_adjoint(grad_p, bins, n, box, rv, ext, n_bins):
    """The transpose of profiles() in delta_s: a float64 grid."""
    l = np.float64(np.float32(box) / np.float32(n))
    bw = np.float32(ext) / np.float32(n_bins)
    vols = np.array([_bin_edges_volume(j, bw) for j in range(n_bins)], np.float64)
    g = np.zeros(n * n * n)
    for c, (flat, b) in enumerate(bins):
        if not (np.float32(rv[c]) > 0) or len(flat) == 0:
            continue
        rv3 = np.float64(np.float32(rv[c])) ** 3
        per_bin = np.asarray(grad_p[c], np.float64) * l ** 3 / (vols * rv3)
        w = np.cumsum(per_bin[::-1])[::-1]  # w(b) = sum_{j >= b}
        np.add.at(g, flat, w[b])
    return g.reshape(n, n, n)


class Chain:
    """The chain at fixed centres, radii and target; positions vary."""

    def __init__(self, n, box, r_smooth, cx, cy, cz, rv, target=None, ext=3.0, n_bins=16):
        self.n, self.box, self.r = n, box, r_smooth
        self.rv, self.ext, self.n_bins = np.asarray(rv), ext, n_bins
        self.bins = node_bins(n, box, cx, cy, cz, rv, ext, n_bins)
        self.target = target

    def forward(self, x, y, z):
        rho = cic(x, y, z, self.n, self.box)
        delta = rho / rho.mean() - 1.0
        ds = smooth(delta, self.box, self.r)
        return profiles(ds, self.bins, self.box, self.rv, self.ext, self.n_bins)

    def loss(self, x, y, z):
        p = self.forward(x, y, z)
        return float(np.sum((p - self.target) ** 2))

    def grad(self, x, y, z):
        rho = cic(x, y, z, self.n, self.box)
        mean = rho.mean()
        ds = smooth(rho / mean - 1.0, self.box, self.r)
        p = profiles(ds, self.bins, self.box, self.rv, self.ext, self.n_bins)
        gp = 2.0 * (p - self.target)
        g_ds = profiles_adjoint(gp, self.bins, self.n, self.box, self.rv, self.ext, self.n_bins)
        g_delta = smooth(g_ds, self.box, self.r)  # a real, even filter: S^T = S
        return cic_adjoint(x, y, z, self.n, self.box, g_delta / mean)


def brute_profiles(delta_s, box, cx, cy, cz, rv, ext, n_bins):
    """profiles.c:202-223 and 619-635 literally, over every node of the grid,
    in float32 per node (tiny grids only)."""
    f = np.float32
    n = delta_s.shape[0]
    l = f(box) / f(n)
    half = f(box) * f(0.5)
    bw = f(ext) / f(n_bins)
    out = np.zeros((len(cx), n_bins))
    for c in range(len(cx)):
        r_v = f(rv[c])
        if not (r_v > 0):
            continue
        r_max = r_v * f(ext)
        inv_bin = f(1.0) / (r_v * bw)
        mass = np.zeros(n_bins)
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    d = []
                    for g, cc in ((i, cx[c]), (j, cy[c]), (k, cz[c])):
                        v = f(g) * l - f(cc)
                        v = v - (f(box) * f(v > half) - f(box) * f(v < -half))
                        d.append(v)
                    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
                    if r2 <= r_max * r_max:
                        b = int(np.sqrt(r2) * inv_bin)
                        if b < n_bins:
                            mass[b] += 1.0 + float(delta_s[i, j, k])
        cum = np.cumsum(mass)
        for b in range(n_bins):
            vol = float(_bin_edges_volume(b, bw)) * float(r_v) ** 3
            out[c, b] = float(l) ** 3 * cum[b] / vol - 1.0
    return out
