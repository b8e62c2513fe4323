"""The sphere-growth kernel's own text (sif_amd/csrc/grow_device.h), run on
the CPU under a 64-lane wavefront emulator (tests/emu/), against the oracle's
restatement of exodus_rescale.c:642-1012 and the brute-force radius of
tests/test_finders.c:433-465. Radii, reasons and enclosed counts bit-exact.

Weighted tracers (exodus_rescale.c:1029-1037): the kernel's weighted form
against the oracle's, on weights whose double sums are exact in any order
(floats in [0.5, 2) over fewer than 2^28 tracers: every partial sum is a
multiple of 2^-24 below 2^29), so the two agree bit for bit whatever order
each adds in; with a fifth of the weights zero (weightless bins, empty
spheres); against an FP64 brute force over every tracer; and all-ones
weights against the unweighted kernel (exodus_finder.c:1634-1638).
"""
import ctypes as C

import numpy as np
import pytest

import emu_lib
import oracle
from sif_amd import zeldovich


BOX, MESH = 100.0, 10


@pytest.fixture(scope="module")
def emu():
    lib = emu_lib.load("grow")
    lib.emu_find_radius.restype = C.c_int
    return lib


@pytest.fixture(scope="module")
def field():
    x, y, z = zeldovich.generate(32, BOX, 11)
    m = oracle.chain_mesh(x, y, z, MESH, BOX)
    return x, y, z, m


def run_emu(lib, m, tpl, centres, rung, rs, t, vol, cache_words=0, mass=None, slices=1):
    k = len(centres)
    r = np.zeros(k, np.float32)
    why = np.zeros(k, np.uint32)
    nin = np.zeros(k, np.uint64)
    cx, cy, cz = (np.ascontiguousarray(centres[:, i]) for i in range(3))
    V = C.c_void_p
    st = lib.emu_find_radius(
        V(m["offsets"].ctypes.data), V(m["x"].ctypes.data), V(m["y"].ctypes.data),
        V(m["z"].ctypes.data), C.c_int32(MESH), C.c_float(BOX), V(tpl["dx"].ctypes.data),
        V(tpl["dy"].ctypes.data), V(tpl["dz"].ctypes.data), V(tpl["type"].ctypes.data),
        V(tpl["min_d2"].ctypes.data), V(tpl["max_d2"].ctypes.data),
        C.c_uint32(len(tpl["dx"])), C.c_float(rung), C.c_float(rs), C.c_float(t),
        C.c_float(vol), V(cx.ctypes.data), V(cy.ctypes.data), V(cz.ctypes.data),
        C.c_uint32(k), V(r.ctypes.data), V(why.ctypes.data), V(nin.ctypes.data),
        C.c_uint32(cache_words), V(m["w"].ctypes.data if mass is not None else None),
        V(m["cell_weights"].ctypes.data if mass is not None else None),
        C.c_float(mass if mass is not None else 0.0), C.c_uint32(slices))
    assert st == 0
    return r, why, nin


@pytest.mark.parametrize("slices", [2, 3])
def test_emulated_split_first_pass_matches_oracle(emu, field, slices):
    """The first pass shared by several waves a candidate (grow_bin_kernel,
    every slices-th template chunk each; with 3 slices of a 2-chunk template
    one wave has none), the walk from their summed histogram with no annulus
    cache: the oracle's radii, reasons and counts."""
    x, y, z, m = field
    rung, factor, t = 9.0, 2.0, np.float32(-0.8)
    rs = rung + max((factor - 1.0) * rung, 3.0 * BOX / 32)
    tpl = oracle.OracleTemplate(MESH, np.float32(BOX) / np.float32(MESH), rung, rs).arrays()
    assert 64 < len(tpl["dx"]) <= 128
    vol = (np.float32(4) / np.float32(3)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    centres = np.random.default_rng(17).uniform(0, BOX, (24, 3)).astype(np.float32)
    r, why, nin = run_emu(emu, m, tpl, centres, rung, rs, t, vol, slices=slices)
    om = oracle.OracleMesh(m, MESH, BOX)
    orr, owhy, onin = oracle.find_radius(om, centres, rung, rs, t, vol)
    np.testing.assert_array_equal(why, owhy)
    np.testing.assert_array_equal(r.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(nin[r > 0], onin[orr > 0])
    assert (r > 0).sum() >= 1 and len(set(why.tolist())) >= 2


def annulus_sizes(x, y, z, centres, rung, rs):
    """Particles with rung^2 < d^2 <= rs^2 (periodic), about what a wave's
    first pass caches."""
    p = np.c_[x, y, z].astype(np.float64)
    out = []
    for c in centres.astype(np.float64):
        d = np.abs(p - c)
        d = np.minimum(d, BOX - d)
        d2 = (d * d).sum(1)
        out.append(int(((d2 > rung * rung) & (d2 <= rs * rs)).sum()))
    return np.array(out)


# cache_words 0: the LDS the library gives the launch (a cache of 1920 here);
# 1024, the least: the annulus is then sorted in windows after it (at most
# 512), sorted whole (512..1024), or too large to cache (walked in passes over
# the mesh), and the case covers all three.
@pytest.mark.parametrize("rung,factor,t,cache_words", [(6.0, 1.5, -0.8, 0),
                                                       (9.0, 2.0, -0.8, 1024),
                                                       (4.0, 1.25, -0.6, 0)])
def test_emulated_kernel_matches_oracle(emu, field, rung, factor, t, cache_words):
    x, y, z, m = field
    rs = rung + max((factor - 1.0) * rung, 3.0 * BOX / 32)
    tpl = oracle.OracleTemplate(MESH, np.float32(BOX) / np.float32(MESH), rung, rs).arrays()
    vol = (np.float32(4) / np.float32(3)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    centres = np.random.default_rng(int(rung * 10)).uniform(0, BOX, (64, 3)).astype(np.float32)
    if cache_words:
        ann = annulus_sizes(x, y, z, centres, rung, rs)
        half = cache_words // 2
        assert (ann <= half - 16).sum() >= 5
        assert ((ann > half + 16) & (ann <= cache_words - 16)).sum() >= 5
        assert (ann > cache_words + 16).sum() >= 5
    r, why, nin = run_emu(emu, m, tpl, centres, rung, rs, np.float32(t), vol, cache_words)
    om = oracle.OracleMesh(m, MESH, BOX)
    orr, owhy, onin = oracle.find_radius(om, centres, rung, rs, np.float32(t), vol)
    np.testing.assert_array_equal(why, owhy)
    np.testing.assert_array_equal(r.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(nin[r > 0], onin[orr > 0])
    for i in np.nonzero(r > 0)[0][:10]:
        want = oracle.brute_force_radius(x, y, z, BOX, centres[i], rung, rs, t)
        assert abs(r[i] - want) <= 1e-5 * want


def mass_factor(total_weight):
    """exodus_finder.c:1639-1642, in float as the reference rounds it."""
    f = np.float32
    return (f(4) / f(3)) * f(3.14159265358979323846) * (f(total_weight) / (f(BOX) ** 3))


def brute_force_weighted(x, y, z, w, centre, rung, rs, t, mass):
    """FP64 over every tracer: the outermost tracer distance d in (rung, rs]
    where the weight strictly inside d, over mass * d^3, is at most 1 + t."""
    p = np.c_[x, y, z].astype(np.float64)
    d = np.abs(p - centre.astype(np.float64))
    d = np.minimum(d, BOX - d)
    d = np.sqrt((d * d).sum(1))
    order = np.argsort(d, kind="stable")
    d, ww = d[order], w[order].astype(np.float64)
    inside = np.concatenate([[0.0], np.cumsum(ww)[:-1]])
    ok = (d > rung) & (d <= rs) & (inside <= (1.0 + t) * float(mass) * d ** 3)
    # equal distances: the weight strictly inside is the group's smallest
    return d[np.nonzero(ok)[0][-1]] if ok.any() else -1.0


@pytest.fixture(scope="module")
def weighted_field(field):
    x, y, z, _ = field
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 2.0, len(x)).astype(np.float32)
    w[rng.random(len(x)) < 0.2] = 0.0
    # a ball of weightless tracers: a search sphere inside it is empty
    d = np.c_[x, y, z].astype(np.float64) - EMPTY_AT
    w[(d * d).sum(1) < 12.0 ** 2] = 0.0
    m = oracle.chain_mesh(x, y, z, MESH, BOX, w=w)
    return x, y, z, w, m


def emptiest_cells(x, y, z, k, n=16):
    """Centres of the k cells of an n^3 count grid holding the fewest tracers."""
    idx = np.floor(np.c_[x, y, z] / (BOX / n)).astype(np.int64) % n
    counts = np.bincount((idx[:, 0] * n + idx[:, 1]) * n + idx[:, 2], minlength=n ** 3)
    flat = np.argsort(counts, kind="stable")[:k]
    c = np.c_[flat // (n * n), (flat // n) % n, flat % n] + 0.5
    return (c * (BOX / n)).astype(np.float32)


EMPTY_AT = np.array([50.0, 50.0, 50.0])


# t = -0.4 at the smaller rung also refuses spheres underdense out to the
# search radius; the weighted cache (1024 pairs) is, across the two, sorted
# in windows after it, sorted whole and exceeded.
@pytest.mark.parametrize("rung,factor,t", [(6.0, 1.5, -0.4), (9.0, 2.0, -0.7)])
def test_emulated_weighted_kernel_matches_oracle(emu, weighted_field, rung, factor, t):
    x, y, z, w, m = weighted_field
    rs = rung + max((factor - 1.0) * rung, 3.0 * BOX / 32)
    tpl = oracle.OracleTemplate(MESH, np.float32(BOX) / np.float32(MESH), rung, rs).arrays()
    vol = (np.float32(4) / np.float32(3)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    mass = mass_factor(m["total_weight"])
    centres = np.r_[EMPTY_AT[None], emptiest_cells(x, y, z, 23),
                    np.random.default_rng(int(rung * 10) + 1).uniform(0, BOX, (24, 3))]
    centres = centres.astype(np.float32)
    r, why, nin = run_emu(emu, m, tpl, centres, rung, rs, np.float32(t), vol, mass=mass)
    om = oracle.OracleMesh(m, MESH, BOX)
    orr, owhy, onin = oracle.find_radius(om, centres, rung, rs, np.float32(t), vol,
                                         mass_factor=mass)
    np.testing.assert_array_equal(why, owhy)
    np.testing.assert_array_equal(r.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(nin[r > 0], onin[orr > 0])
    assert (r > 0).sum() >= 10 and set(why.tolist()) >= {0, 3}
    assert why[0] == 1 if rs < 12.0 else True  # the weightless ball
    for i in np.nonzero(r > 0)[0][:10]:
        want = brute_force_weighted(x, y, z, w, centres[i], rung, rs, t, mass)
        assert abs(r[i] - want) <= 1e-5 * want


def test_emulated_unit_weights_reproduce_unweighted(emu, field):
    """exodus_finder.c:1634-1638: a mesh whose weights are all 1 gives the
    unweighted run bit for bit."""
    x, y, z, m = field
    ones = oracle.chain_mesh(x, y, z, MESH, BOX, w=np.ones(len(x), np.float32))
    rung, rs, t = 6.0, 9.375, np.float32(-0.8)
    tpl = oracle.OracleTemplate(MESH, np.float32(BOX) / np.float32(MESH), rung, rs).arrays()
    vol = (np.float32(4) / np.float32(3)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    assert mass_factor(ones["total_weight"]) == vol
    centres = np.random.default_rng(3).uniform(0, BOX, (64, 3)).astype(np.float32)
    r0, why0, nin0 = run_emu(emu, m, tpl, centres, rung, rs, t, vol)
    r1, why1, nin1 = run_emu(emu, ones, tpl, centres, rung, rs, t, vol, mass=vol)
    np.testing.assert_array_equal(why0, why1)
    np.testing.assert_array_equal(r0.view(np.uint32), r1.view(np.uint32))
    np.testing.assert_array_equal(nin0, nin1)
