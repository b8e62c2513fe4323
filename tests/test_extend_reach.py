"""SIF_FINDER_EXTEND_REACH (exodus_finder.c:234-244, 1670-1702, 1824-1832,
2060-2072): a growth that ends still underdense at its rung's reach is
measured again with the reach extended, level by level, up to the run's
largest search radius, instead of being refused.

The case is built so that the extension decides it. The particles are uniform
with an empty sphere of 13 Mpc/h (3% of its tracers kept) at node p. The grid
the finder scans is a separate field, zero but for one deep node at p, so the
top-hat is below the threshold around p at the 6 Mpc/h rung and not at the
12 Mpc/h one. At 6 Mpc/h the reach is 12 (the gap guard), short of the
crossing near 16; the levels reach 18.

Pinned: without the option the run refuses every candidate there and finds
nothing; with it, one void, at a radius equal to the brute-force crossing
(every tracer's distance, tests/test_finders.c:433-465) and past the rung's
reach; the level table against values worked by hand from the reference's
expressions. GPU: the device catalogue equals the oracle's on the device's own
fields, with and without the option, and on a Zel'dovich field where the
option changes nothing.
"""
import numpy as np
import pytest

import oracle

BOX, N, MESH = 100.0, 64, 10
P_NODE = (20, 30, 40)
HOLE_R = 13.0
RADII = np.array([12.0, 6.0], np.float32)
THRESH = -0.4
EXTEND = 1 << 21


def hole_particles(n=150000, seed=4):
    rng = np.random.default_rng(seed)
    p = np.array(P_NODE, np.float64) * (BOX / N)
    pts = rng.uniform(0, BOX, (int(n * 1.2), 3))
    d = pts - p
    d -= BOX * np.round(d / BOX)
    inside = (d ** 2).sum(1) < HOLE_R ** 2
    keep = ~inside | (rng.random(len(pts)) < 0.03)
    pts = pts[keep][:n].astype(np.float32)
    pts = np.minimum(pts, np.nextafter(np.float32(BOX), np.float32(0)))  # inside [0, BOX)
    return [np.ascontiguousarray(pts[:, k]) for k in range(3)]


def dip_grid():
    g = np.zeros(N ** 3, np.float32)
    g[(P_NODE[0] * N + P_NODE[1]) * N + P_NODE[2]] = -300.0
    return g


def oracle_fields(delta):
    spec = oracle.fft_r2c(delta, N)
    axis = oracle.window_axis(N, 2, 0)
    out = []
    for r in RADII:
        lut = oracle.filter_lut(1, float(r), 0.0, BOX, N)
        fac = oracle.mode_factors(lut, axis, N)
        out.append((oracle.fft_c2r(spec * fac, N) / N ** 3).astype(np.float32))
    return out


def levels(radii, l, sf=1.5):
    """The level table as exodus_finder.c:1342-1353, 1680-1696 state it."""
    r_max = max(radii)
    r_ext = np.float32(r_max + max((sf - 1) * r_max, 3 * l))
    r_low = min(radii)
    n = 0
    while n < 32 and (n == 0 or r_ext / np.float32(1.5) ** (n - 1) > r_low):
        n += 1
    reach = [np.float32(r_ext / np.float32(1.5) ** (n - 1 - k)) for k in range(n)]
    rmin = [(reach[k - 1] if k else reach[0] / np.float32(1.5)) / np.float32(1.2)
            for k in range(n)]
    return reach, rmin


def test_level_table_by_hand():
    # radii 12 and 6 on cells of 1.5625: r_ext = 12 + max(6, 4.6875) = 18;
    # 18, 12, 8 > 6 and 5.33 is not: four levels
    reach, rmin = levels([12.0, 6.0], BOX / N)
    np.testing.assert_allclose(reach, [18 / 3.375, 8.0, 12.0, 18.0], rtol=1e-6)
    np.testing.assert_allclose(rmin, [18 / 3.375 / 1.5 / 1.2, 18 / 3.375 / 1.2, 8 / 1.2,
                                      12 / 1.2], rtol=1e-6)


@pytest.fixture(scope="module")
def case():
    x, y, z = hole_particles()
    mc = MESH
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mc, BOX), mc, BOX)
    return x, y, z, om, oracle_fields(dip_grid())


def _run(case, opt, f=0.0):
    x, y, z, om, fields = case
    return oracle.exodus_run(fields, N, BOX, om, len(x), RADII, THRESH, overlap=f,
                             search_opt=opt, tie_rule=1)


def test_the_grid_selects_the_bottom_rung_only(case):
    fields = case[4]
    i = (P_NODE[0] * N + P_NODE[1]) * N + P_NODE[2]
    assert fields[0][i] > THRESH and fields[1][i] <= THRESH


@pytest.mark.parametrize("f", [0.0, 1.0])
def test_extension_finds_the_void_past_the_reach(case, f):
    x, y, z = case[:3]
    plain = _run(case, 0, f)
    assert len(plain[0]) == 0
    assert plain[3][1]["rescale_fail"] > 0
    cat, nin, rung, stats = _run(case, EXTEND, f)
    assert len(cat) == 1 and rung[0] == 1
    cx, cy, cz, r = cat[0]
    assert 12.0 < r <= 18.0
    want = oracle.brute_force_radius(x, y, z, BOX, (cx, cy, cz), 6.0, 18.0, THRESH)
    assert want > 0 and abs(r - want) <= 1e-5 * want
    # the witness is a shortcut: the catalogue is the same without it
    alone = oracle.exodus_run(case[4], N, BOX, case[3], len(x), RADII, THRESH, overlap=f,
                              search_opt=EXTEND, tie_rule=1, witness=False)
    np.testing.assert_array_equal(alone[0], cat)


# ---- GPU --------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("opt", [0, EXTEND, EXTEND | (1 << 9)])
def test_device_extension_matches_oracle(native, case, opt):
    from sif_amd import device as D
    x, y, z, om, _ = case
    g = D.to_device(dip_grid())
    fields = D.exodus_fields(g, N, BOX, RADII, opt)
    m = D.Mesh(D.to_device(x), D.to_device(y), D.to_device(z), len(x), MESH, BOX)
    cat, nin, stats = D.exodus(g, N, BOX, m, RADII, THRESH, options=opt)
    ocat, onin, orung, _ = oracle.exodus_run(list(fields), N, BOX, om, len(x), RADII, THRESH,
                                             search_opt=opt, tie_rule=1)
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)
    assert len(cat) == (1 if opt & EXTEND else 0)
