"""SIF_FINDER_SEEDS_MINIMA: the candidates are the local minima of the rung's
field, on their nodes (src/finder/utils.c:228-262, 406-552 with minima_only
and no refinement; exodus_finder.c:1740-1760).

CPU: the oracle's scan against an independent numpy formulation (np.roll of
the 26 neighbours) on random fields with planted ties, masks, NaNs and the
smallest periodic grids, and the reference's own known answers
(tests/test_finders.c:916-1040): one minimum per anisotropic well, on the
node nearest its centre, across the box face too; one minimum, the first
node, on a flat field.

GPU: the device scan (sif_amd_scan_minima) bit-equal to the oracle. A whole
exodus run under SEEDS_MINIMA against the oracle's run is in
test_gpu_parity.py.
"""
import numpy as np
import pytest

import oracle


def numpy_minima(d, n, threshold, mask_bits=None):
    f = d.reshape(n, n, n)
    idx = np.arange(n ** 3).reshape(n, n, n)
    ok = np.ones(f.shape, bool)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                if a == b == c == 0:
                    continue
                u = np.roll(f, (-a, -b, -c), (0, 1, 2))
                j = np.roll(idx, (-a, -b, -c), (0, 1, 2))
                with np.errstate(invalid="ignore"):
                    ok &= np.where(j < idx, f < u, f <= u)
    with np.errstate(invalid="ignore"):
        sel = (ok & (f <= threshold)).ravel()
    if mask_bits is not None:
        sel &= ~mask_bits
    fl = np.flatnonzero(sel)
    order = np.lexsort((fl, d[fl]))
    return d[fl][order], fl[order].astype(np.uint32)


def pack_mask(bits):
    n = len(bits)
    words = np.zeros((n + 63) // 64, np.uint64)
    for i in np.flatnonzero(bits):
        words[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    return words


def random_field(n, seed, ties=True, nan=False):
    rng = np.random.default_rng(seed)
    d = rng.normal(0, 1, n ** 3).astype(np.float32)
    if ties:  # plateaus of equal values, some straddling the box face
        for _ in range(max(1, n ** 3 // 200)):
            i = rng.integers(n ** 3)
            v = np.float32(rng.normal(-2, 0.5))
            d[i] = v
            d[(i + 1) % n ** 3] = v
            d[(i + n) % n ** 3] = v
        d = np.round(d * 4) / 4  # coarse values: many equal neighbours
        d = d.astype(np.float32)
    if nan:
        d[rng.integers(n ** 3, size=5)] = np.nan
    return d


@pytest.mark.parametrize("n,seed,nan", [(1, 0, False), (2, 1, False), (3, 2, False),
                                        (8, 3, False), (16, 4, True), (20, 5, False)])
def test_oracle_minima_match_numpy(n, seed, nan):
    d = random_field(n, seed, nan=nan)
    for threshold in (-0.5, 0.0, 10.0):
        got = oracle.scan_minima(d, n, threshold)
        want = numpy_minima(d, n, threshold)
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[0], want[0])
    rng = np.random.default_rng(seed + 100)
    bits = rng.random(n ** 3) < 0.3
    got = oracle.scan_minima(d, n, 0.0, mask=pack_mask(bits))
    np.testing.assert_array_equal(got[1], numpy_minima(d, n, 0.0, bits)[1])


def wrap_delta(a, b, L):
    d = np.fmod(a - b, L)
    d = np.where(d < -0.5 * L, d + L, d)
    return np.where(d >= 0.5 * L, d - L, d)


def well_field(centre, n=32):
    """tests/test_finders.c:936-964: -0.95 exp(-q/2), q on a turned
    anisotropic quadratic form, on a box of 32 cells of one."""
    a, b = 0.6, 0.9
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    R = np.array([[ca, -sa, 0.0], [sa * cb, ca * cb, -sb], [sa * sb, ca * sb, cb]])
    Q = R @ np.diag([1 / 9, 1 / 16, 1 / 25]) @ R.T
    i = np.arange(n, dtype=np.float64)
    d = [wrap_delta(i, centre[k], float(n)) for k in range(3)]
    D = np.stack(np.meshgrid(*d, indexing="ij"), -1)
    q = np.einsum("...u,uv,...v->...", D, Q, D)
    return (-0.95 * np.exp(-0.5 * q)).astype(np.float32).ravel()


WELLS = [(10.3, 20.7, 5.5), (0.2, 31.9, 15.0), (16.49, 16.51, 16.0)]


@pytest.mark.parametrize("centre", WELLS)
def test_oracle_one_minimum_per_well(centre):
    n = 32
    d = well_field(centre, n)
    depths, flat = oracle.scan_minima(d, n, -0.5)
    assert len(flat) == 1 and len(oracle.scan_candidates(d, -0.5)[1]) > 1
    assert flat[0] == int(np.argmin(d))  # the deepest node
    node = np.array([flat[0] // (n * n), (flat[0] // n) % n, flat[0] % n], np.float64)
    assert np.all(np.abs(wrap_delta(node, np.array(centre), float(n))) <= 1.0)


def test_oracle_flat_field_has_one_minimum():
    n = 32
    d = np.full(n ** 3, -0.9, np.float32)
    depths, flat = oracle.scan_minima(d, n, -0.5)
    assert flat.tolist() == [0] and depths.tolist() == [np.float32(-0.9)]
    assert len(oracle.scan_candidates(d, -0.5)[1]) == n ** 3


# ---- GPU --------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(2, 1), (8, 3), (16, 4), (33, 6), (64, 7)])
def test_device_minima_scan_matches_oracle(native, n, seed):
    from sif_amd import device as D
    d = random_field(n, seed, nan=(n == 16))
    dd = D.to_device(d)
    rng = np.random.default_rng(seed)
    bits = rng.random(n ** 3) < 0.2
    words = pack_mask(bits)
    for mask in (None, words):
        dmask = D.to_device(mask, np.uint64) if mask is not None else None
        for threshold in (-0.5, 0.0, 10.0):
            got = D.scan_minima(dd, n, threshold, dmask)
            want = oracle.scan_minima(d, n, threshold, mask=mask, tie_rule=1)
            np.testing.assert_array_equal(got[1], want[1])
            np.testing.assert_array_equal(got[0].view(np.uint32), want[0].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("centre", WELLS)
def test_device_one_minimum_per_well(native, centre):
    from sif_amd import device as D
    d = well_field(centre)
    depths, flat = D.scan_minima(D.to_device(d), 32, -0.5)
    assert flat.tolist() == [int(np.argmin(d))]
