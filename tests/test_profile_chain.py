"""The differentiable profile chain: positions -> CIC -> contrast -> Gaussian
smoothing -> cumulative sphere-averaged density profiles -> sum of squared
residuals, and its gradient in the positions.

CPU (the FP64 restatement in oracle/profile_chain.py, pinned):
- profiles against a literal loop over every node of the grid, in the
  convention of profiles.c:202-223, 332-350, 619-635;
- the smoothing of one Fourier mode against its closed form, and the table
  against the C oracle's Gaussian LUT (fft.c:521-568);
- the deposit against the C oracle's textbook double deposit;
- the profiles' transpose against the profiles (dot-product identity);
- the gradient against FP64 central differences.

GPU (through the C ABI and sif_amd.autograd):
- sif_amd_grid_profiles, sif_amd_grid_profiles_adjoint and
  sif_amd_cic_adjoint against the restatement on the same inputs;
- the dot-product identity of the device profiles and their transpose;
- the full chain's gradient (torch.autograd) against FP64 central
  differences with step 1e-4 on a 64^3 case, and against the FP64 gradient
  over every particle: relative error <= 1e-4 (2-norm over the coordinates
  compared).
"""
import numpy as np
import pytest

import oracle
from oracle import profile_chain as pc


def centres_for(rng, k, box, rlo, rhi):
    c = rng.uniform(0, box, (k, 3)).astype(np.float32)
    r = rng.uniform(rlo, rhi, k).astype(np.float32)
    return c[:, 0], c[:, 1], c[:, 2], r


def test_profiles_match_literal_loop():
    rng = np.random.default_rng(3)
    n, box = 12, 30.0
    ds = rng.normal(0, 0.5, (n, n, n))
    cx, cy, cz, rv = centres_for(rng, 6, box, 1.0, 4.5)
    cx[0], cy[0], cz[0] = 0.0, 29.9, 15.0  # on and near the faces
    rv[1] = 0.0                            # no radius: a zero row
    bins = pc.node_bins(n, box, cx, cy, cz, rv, 3.0, 16)
    got = pc.profiles(ds, bins, box, rv, 3.0, 16)
    want = pc.brute_profiles(ds, box, cx, cy, cz, rv, 3.0, 16)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    assert np.all(got[1] == 0)


def test_smoothing_single_mode_and_table():
    n, box, r = 32, 100.0, 3.0
    x = np.arange(n) * box / n
    m = 3
    delta = np.cos(2 * np.pi * m * x / box)[:, None, None] * np.ones((n, n, n))
    got = pc.smooth(delta, box, r)
    want = np.exp(-0.5 * (r * 2 * np.pi * m / box) ** 2) * delta
    np.testing.assert_allclose(got, want, atol=1e-12)
    lut = oracle.filter_lut(2, r, 0.0, box, n)  # float, indexed by |k|^2
    tab = pc.gaussian_table(n, box, r)
    k = np.fft.fftfreq(n, 1.0 / n)
    k2 = (k[:, None, None] ** 2 + k[None, :, None] ** 2
          + np.arange(n // 2 + 1)[None, None, :] ** 2).astype(np.int64)
    np.testing.assert_allclose(tab, lut[k2], rtol=2e-6, atol=1e-30)


def test_deposit_matches_c_oracle():
    rng = np.random.default_rng(5)
    n, box = 16, 50.0
    x, y, z = rng.uniform(0, box, (3, 5000)).astype(np.float32)
    got = pc.cic(x, y, z, n, box).reshape(-1)
    want = oracle.cic_brute(x, y, z, n, box)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    np.testing.assert_allclose(got.sum() * (box / n) ** 3, len(x), rtol=1e-12)


def test_profiles_transpose_identity():
    rng = np.random.default_rng(7)
    n, box = 16, 40.0
    cx, cy, cz, rv = centres_for(rng, 10, box, 1.5, 6.0)
    bins = pc.node_bins(n, box, cx, cy, cz, rv, 3.0, 16)
    v = rng.normal(size=(n, n, n))
    g = rng.normal(size=(10, 16))
    av = pc.profiles(v, bins, box, rv, 3.0, 16) - pc.profiles(np.zeros_like(v), bins, box, rv,
                                                              3.0, 16)
    atg = pc.profiles_adjoint(g, bins, n, box, rv, 3.0, 16)
    lhs, rhs = np.sum(av * g), np.sum(v * atg)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def small_chain(seed, n_side, n, box, r_smooth, k, rlo, rhi):
    from sif_amd import zeldovich
    x, y, z = zeldovich.generate(n_side, box, seed)
    rng = np.random.default_rng(seed + 100)
    cx, cy, cz, rv = centres_for(rng, k, box, rlo, rhi)
    ch = pc.Chain(n, box, r_smooth, cx, cy, cz, rv)
    p0 = ch.forward(x, y, z)
    ch.target = p0 + 0.1 * np.random.default_rng(seed + 200).normal(size=p0.shape)
    return ch, (x, y, z), (cx, cy, cz, rv)


def fd_coordinates(x, y, z, n, box, g, count, h, rng):
    """Coordinates with the largest gradients whose step stays inside one
    CIC cell (the deposit is linear there)."""
    pos = np.stack([x, y, z]).astype(np.float64)
    u = pos * (n / box)
    inner = (u - np.floor(u) > 2 * h * n / box) & (np.ceil(u) - u > 2 * h * n / box)
    mag = np.where(inner, np.abs(np.stack(g)), 0.0)
    top = np.argsort(mag.reshape(-1))[::-1][:count * 4]
    pick = rng.choice(top, count, replace=False)
    return [divmod(int(i), len(x)) for i in pick]


def central_differences(ch, xyz, coords, h):
    out = []
    base = [np.asarray(a, np.float64) for a in xyz]
    for axis, i in coords:
        plus = [a.copy() for a in base]
        minus = [a.copy() for a in base]
        plus[axis][i] += h
        minus[axis][i] -= h
        out.append((ch.loss(*plus) - ch.loss(*minus)) / (2 * h))
    return np.array(out)


def test_fp64_gradient_matches_central_differences():
    ch, (x, y, z), _ = small_chain(21, 16, 16, 40.0, 2.0, 8, 2.0, 5.0)
    g = ch.grad(x, y, z)
    coords = fd_coordinates(x, y, z, 16, 40.0, g, 12, 1e-4, np.random.default_rng(1))
    fd = central_differences(ch, (x, y, z), coords, 1e-4)
    an = np.array([g[a][i] for a, i in coords])
    assert np.linalg.norm(an - fd) <= 1e-7 * np.linalg.norm(fd)


# ---- GPU --------------------------------------------------------------------

@pytest.fixture(scope="module")
def torch_gpu(native):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("n_bins", [16, 24])
def test_gpu_profiles_and_transpose_match_oracle(torch_gpu, native, n_bins):
    torch = torch_gpu
    import ctypes as C
    rng = np.random.default_rng(11 + n_bins)
    n, box = 64, 125.0
    field = rng.normal(0, 0.3, (n, n, n)).astype(np.float32)
    cx, cy, cz, rv = centres_for(rng, 200, box, 3.0, 20.0)
    rv[5] = 0.0
    print(f"[profiles] n={n} centres=200 n_bins={n_bins}", flush=True)
    t = [dev(torch, a) for a in (field.reshape(-1), cx, cy, cz, rv)]
    out = torch.empty((200, n_bins), dtype=torch.float32, device="cuda")
    P = lambda a: C.c_void_p(a.data_ptr())
    native.check(native.lib.sif_amd_grid_profiles(P(t[0]), n, box, P(t[1]), P(t[2]), P(t[3]),
                                                  P(t[4]), 200, 3.0, n_bins, P(out)))
    native.check(native.lib.sif_amd_synchronize())
    bins = pc.node_bins(n, box, cx, cy, cz, rv, 3.0, n_bins)
    want = pc.profiles(field.astype(np.float64), bins, box, rv, 3.0, n_bins)
    got = out.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6)
    assert np.all(got[5] == 0)
    # the transpose
    g = rng.normal(size=(200, n_bins)).astype(np.float32)
    tg = dev(torch, g)
    gd = torch.empty(n ** 3, dtype=torch.float32, device="cuda")
    native.check(native.lib.sif_amd_grid_profiles_adjoint(P(tg), n, box, P(t[1]), P(t[2]),
                                                          P(t[3]), P(t[4]), 200, 3.0, n_bins,
                                                          P(gd)))
    native.check(native.lib.sif_amd_synchronize())
    got_t = gd.cpu().numpy().astype(np.float64)
    want_t = pc.profiles_adjoint(g.astype(np.float64), bins, n, box, rv, 3.0, n_bins).reshape(-1)
    assert np.linalg.norm(got_t - want_t) <= 1e-6 * np.linalg.norm(want_t)
    # dot-product identity between the device's two operators, in FP64; v is
    # large enough that the profiles' affine part (and its float rounding)
    # is small beside their linear part
    v = (100.0 * rng.normal(size=n ** 3)).astype(np.float32)
    zero = np.zeros(n ** 3, np.float32)
    outs = []
    for f in (v, zero):
        tf = dev(torch, f)
        native.check(native.lib.sif_amd_grid_profiles(P(tf), n, box, P(t[1]), P(t[2]), P(t[3]),
                                                      P(t[4]), 200, 3.0, n_bins, P(out)))
        native.check(native.lib.sif_amd_synchronize())
        outs.append(out.cpu().numpy().astype(np.float64))
    terms = (outs[0] - outs[1]) * g
    lhs = np.sum(terms)
    rhs = np.sum(v.astype(np.float64) * got_t)
    assert abs(lhs - rhs) <= 1e-5 * np.sum(np.abs(terms))


@pytest.mark.gpu
def test_gpu_cic_adjoint_matches_oracle(torch_gpu, native):
    torch = torch_gpu
    import ctypes as C
    rng = np.random.default_rng(13)
    n, box = 32, 80.0
    x, y, z = rng.uniform(0, box, (3, 20000)).astype(np.float32)
    x[:4] = [0.0, box, 79.99999, 40.0]  # faces and a node
    G = rng.normal(size=n ** 3).astype(np.float32)
    print("[cic adjoint] n=32 particles=20000", flush=True)
    tx, ty, tz, tG = (dev(torch, a) for a in (x, y, z, G))
    gx, gy, gz = (torch.empty_like(tx) for _ in range(3))
    P = lambda a: C.c_void_p(a.data_ptr())
    native.check(native.lib.sif_amd_cic_adjoint(P(tx), P(ty), P(tz), len(x), n, box, P(tG),
                                                P(gx), P(gy), P(gz)))
    native.check(native.lib.sif_amd_synchronize())
    want = pc.cic_adjoint(x, y, z, n, box, G.astype(np.float64))
    for got, w in zip((gx, gy, gz), want):
        got = got.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - w) <= 1e-5 * np.linalg.norm(w)


@pytest.mark.gpu
def test_gpu_chain_gradient_matches_fp64(torch_gpu, native):
    """64^3 Zel'dovich particles in a 125 Mpc/h box (the density of the
    256^3 / 500 Mpc/h chain), 64^3 grid, Gaussian R = 3, 64 centres of radius
    5-15, 16 bins to 3R, target the forward profile plus seeded noise."""
    torch = torch_gpu
    from sif_amd.autograd import ProfileChain, void_profiles
    n, box, r_s = 64, 125.0, 3.0
    print("[chain gradient] 64^3 particles, grid 64^3, 64 centres", flush=True)
    ch, (x, y, z), (cx, cy, cz, rv) = small_chain(31, 64, n, box, r_s, 64, 5.0, 15.0)
    chain = ProfileChain(n, box, r_s, *(dev(torch, a) for a in (cx, cy, cz, rv)))
    tx, ty, tz = (dev(torch, a).requires_grad_(True) for a in (x, y, z))
    p = void_profiles(tx, ty, tz, chain)
    p64 = ch.forward(x, y, z)
    np.testing.assert_allclose(p.detach().cpu().numpy(), p64, rtol=1e-4, atol=1e-5)
    loss = ((p - dev(torch, ch.target)) ** 2).sum()
    loss.backward()
    got = [t.grad.cpu().numpy().astype(np.float64) for t in (tx, ty, tz)]
    g64 = ch.grad(x, y, z)
    num = np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(got, g64)))
    den = np.sqrt(sum(np.sum(b ** 2) for b in g64))
    assert num <= 1e-4 * den, (num, den)
    coords = fd_coordinates(x, y, z, n, box, g64, 12, 1e-4, np.random.default_rng(2))
    fd = central_differences(ch, (x, y, z), coords, 1e-4)
    gg = np.array([got[a][i] for a, i in coords])
    assert np.linalg.norm(gg - fd) <= 1e-4 * np.linalg.norm(fd), (gg, fd)
