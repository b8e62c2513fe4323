"""SIF_FINDER_DISPLACE (exodus_finder.c:684-1032, 1871-1880, 2210-2275), and
with it the excursion-peak variant whole (SEEDS_MINIMA | EXTEND_REACH |
DISPLACE at overlap fraction 1, exodus_finder.c:1203-1212): a void is refused
only by a larger one and displaces the smaller ones it collides with; their
nodes leave the mask, and the nodes of the voids that still stand over them
are set again.

CPU, the oracle:
- a constructed displacement: a hole found at the 12 Mpc/h rung, then a
  larger neighbour found at 6 Mpc/h through the extended reach, which takes
  its place; what remains is two voids, neither colliding;
- on Zel'dovich fields, the invariant the rule implies: no two voids of the
  catalogue collide by find_conflicts' own test, at every overlap fraction;
  and each radius is the brute-force crossing;
- the four holes of the reference's tests (tests/test_finders.c:50-164) under
  displacement and under the excursion-peak options.
GPU: the device catalogue equals the oracle's on the same fields, and the
pysif excursion-peak variant runs through the C API.
"""
import numpy as np
import pytest

import oracle
from test_extend_reach import BOX, MESH, N, THRESH, oracle_fields
from test_oracle import (BOX4, N_GRID4, RADII4, check_four_holes, holes)  # noqa: F401

DISPLACE, EXTEND, MINIMA = 1 << 22, 1 << 21, 1 << 9
PEAK = MINIMA | EXTEND | DISPLACE
P_NODE, Q_NODE = (20, 30, 40), (34, 30, 40)  # 14 nodes, 21.9 Mpc/h apart
RADII = np.array([12.0, 6.0], np.float32)


def two_hole_particles(n=150000, seed=4):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, BOX, (int(n * 1.4), 3))
    keep = np.ones(len(pts), bool)
    for c, R in ((P_NODE, 13.0), (Q_NODE, 10.0)):
        d = pts - np.array(c) * (BOX / N)
        d -= BOX * np.round(d / BOX)
        keep &= ((d ** 2).sum(1) >= R * R) | (rng.random(len(pts)) < 0.03)
    pts = pts[keep][:n].astype(np.float32)
    pts = np.minimum(pts, np.nextafter(np.float32(BOX), np.float32(0)))
    return [np.ascontiguousarray(pts[:, k]) for k in range(3)]


def two_dip_grid():
    """One node deep enough for the 12 Mpc/h top-hat at q, one for the 6
    Mpc/h one alone at p."""
    g = np.zeros(N ** 3, np.float32)
    g[(P_NODE[0] * N + P_NODE[1]) * N + P_NODE[2]] = -300.0
    g[(Q_NODE[0] * N + Q_NODE[1]) * N + Q_NODE[2]] = -1200.0
    return g


def colliding_pairs(cat, box, f, scale=1.0):
    """Pairs of voids that collide by find_conflicts' float test
    (exodus_finder.c:728-740)."""
    c = cat.astype(np.float32)
    out = []
    half = np.float32(0.5) * np.float32(box)
    for i in range(len(c)):
        d = np.abs(c[:i, :3] - c[i, :3])
        d = np.where(d > half, np.float32(box) - d, d)
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        ri, rj = c[i, 3], c[:i, 3]
        tot = np.float32(scale) * (ri + rj - np.float32(f) * np.minimum(ri, rj))
        out += [(j, i) for j in np.flatnonzero(d2 < tot * tot)]
    return out


@pytest.fixture(scope="module")
def constructed():
    x, y, z = two_hole_particles()
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, MESH, BOX), MESH, BOX)
    return x, y, z, om, oracle_fields(two_dip_grid())


def _run(case, opt, f=0.0):
    x, y, z, om, fields = case
    return oracle.exodus_run(fields, N, BOX, om, len(x), RADII, THRESH, overlap=f,
                             search_opt=opt, tie_rule=1)


def test_a_larger_void_displaces_a_smaller_one(constructed):
    x, y, z = constructed[:3]
    plain, _, prung, _ = _run(constructed, EXTEND)
    assert len(plain) == 1 and prung.tolist() == [0]
    cat, nin, rung, stats = _run(constructed, EXTEND | DISPLACE)
    assert stats[1]["displaced"] == 1 and stats[1]["accepted"] >= 1
    assert not any(np.array_equal(plain[0], v) for v in cat)  # it is gone
    assert rung.tolist().count(1) == len(cat)                  # all from the low rung
    assert colliding_pairs(cat, BOX, 0.0) == []
    big = cat[np.argmax(cat[:, 3])]
    assert big[3] > plain[0][3] > 0 and big[3] > 12.0
    for cx, cy, cz, r in cat:
        want = oracle.brute_force_radius(x, y, z, BOX, (cx, cy, cz), 6.0, 18.0, THRESH)
        assert abs(r - want) <= 1e-5 * want


def test_without_extension_nothing_is_displaced(constructed):
    cat, _, _, stats = _run(constructed, DISPLACE)
    assert sum(s["displaced"] for s in stats) == 0
    plain = _run(constructed, 0)[0]
    np.testing.assert_array_equal(cat, plain)


@pytest.fixture(scope="module")
def zeldovich_case():
    from sif_amd import zeldovich
    from test_oracle import smoothed_fields
    n_side, box, n, mc = 32, 100.0, 32, 10
    x, y, z = zeldovich.generate(n_side, box, 11)
    radii = np.array([14.0, 11.0, 9.0, 7.5, 6.25], np.float32)
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mc, box), mc, box)
    return x, y, z, om, smoothed_fields(x, y, z, n, box, radii), radii, box, n


@pytest.mark.parametrize("f,opt", [(0.0, DISPLACE), (0.5, DISPLACE), (1.0, DISPLACE),
                                   (1.0, PEAK), (0.0, PEAK)])
def test_no_two_voids_collide(zeldovich_case, f, opt):
    x, y, z, om, fields, radii, box, n = zeldovich_case
    cat, nin, rung, stats = oracle.exodus_run(fields, n, box, om, len(x), radii, -0.8,
                                              overlap=f, search_opt=opt, tie_rule=1)
    assert len(cat) > 0
    assert colliding_pairs(cat, box, f) == []
    top = float(radii[0]) + max(0.5 * float(radii[0]), 3.0 * box / n)
    for (cx, cy, cz, r), k in list(zip(cat, rung))[:25]:
        want = oracle.brute_force_radius(x, y, z, box, (cx, cy, cz), float(radii[k]), top, -0.8)
        assert abs(r - want) <= 1e-5 * want


@pytest.mark.parametrize("opt,f", [(DISPLACE, 0.0), (PEAK, 1.0)])
def test_four_holes_under_displacement(holes, opt, f):  # noqa: F811
    x, y, z, mc, m, fields = holes
    om = oracle.OracleMesh(m, mc, BOX4)
    cat, _, _, _ = oracle.exodus_run(fields, N_GRID4, BOX4, om, len(x), RADII4, -0.7,
                                     overlap=f, search_opt=opt)
    check_four_holes(cat, 0.9, 1.5)


# ---- GPU --------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("opt", [EXTEND | DISPLACE, DISPLACE, PEAK])
def test_device_constructed_displacement(native, constructed, opt):
    from sif_amd import device as D
    x, y, z, om, _ = constructed
    g = D.to_device(two_dip_grid())
    fields = D.exodus_fields(g, N, BOX, RADII, opt)
    m = D.Mesh(D.to_device(x), D.to_device(y), D.to_device(z), len(x), MESH, BOX)
    cat, nin, stats = D.exodus(g, N, BOX, m, RADII, THRESH, options=opt)
    ocat, onin, _, _ = oracle.exodus_run(list(fields), N, BOX, om, len(x), RADII, THRESH,
                                         search_opt=opt, tie_rule=1)
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)


@pytest.mark.gpu
@pytest.mark.parametrize("f,opt", [(0.0, DISPLACE), (1.0, PEAK), (0.5, PEAK)])
def test_device_zeldovich_displacement(native, f, opt):
    from sif_amd import device as D
    from sif_amd import zeldovich
    n_side, box, n, mc = 48, 150.0, 64, 12
    x, y, z = zeldovich.generate(n_side, box, 21)
    radii = np.array([18.0, 14.0, 11.0, 9.0, 7.5, 6.25, 5.0], np.float32)
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    g = D.cic(dx, dy, dz, len(x), n, box)
    D.density_contrast(g, n ** 3)
    fields = D.exodus_fields(g, n, box, radii, opt)
    m = D.Mesh(dx, dy, dz, len(x), mc, box)
    cat, nin, stats = D.exodus(g, n, box, m, radii, -0.7, overlap=f, options=opt)
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mc, box), mc, box)
    ocat, onin, _, _ = oracle.exodus_run(list(fields), n, box, om, len(x), radii, -0.7,
                                         overlap=f, search_opt=opt, tie_rule=1)
    assert len(cat) > 0
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)
    assert colliding_pairs(cat, box, f) == []
    # growth windows of 16 and 100 candidates: many windows a rung, and
    # displacements that free candidates of a window already grown
    try:
        for floor in (16, 100):
            D.set_batch_window(floor, floor)
            g2 = D.cic(dx, dy, dz, len(x), n, box)
            D.density_contrast(g2, n ** 3)
            cat2, nin2, _ = D.exodus(g2, n, box, m, radii, -0.7, overlap=f, options=opt)
            np.testing.assert_array_equal(cat2.view(np.uint32), ocat.view(np.uint32))
            np.testing.assert_array_equal(nin2, onin)
    finally:
        D.set_batch_window(0, 0)


@pytest.mark.gpu
def test_pysif_excursion_peak_variant(native):
    import sif_amd.pysif as pysif
    from sif_amd.pysif import finders
    from sif_amd import zeldovich
    pysif.system_init()
    x, y, z = zeldovich.generate(32, 100.0, 11)
    field = pysif.field_from_numpy(x, y, z)
    grid = pysif.Grid(32, 100.0)
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(10, 100.0, field)
    radii = np.array([14.0, 11.0, 9.0, 7.5, 6.25], np.float32)
    cat = finders.exodus(grid, mesh, radii, -0.8, variant="excursion-peak", checkpoint=False)
    assert cat.n_voids > 0
    assert colliding_pairs(np.column_stack([cat.centres, cat.radii]), 100.0, 1.0) == []
    # the variant is its options (exodus_finder.c:1203-1212), taken one by one,
    # on a grid made afresh (the run restores its grid to float precision only)
    grid2 = pysif.Grid(32, 100.0)
    grid2.assign_cic(field)
    grid2.to_density_contrast()
    raw = finders.exodus_raw(grid2, mesh, radii, -0.8, 1.0, seeds="minima", extend_reach=True,
                             displace=True, checkpoint=False)
    np.testing.assert_array_equal(raw.centres, cat.centres)
    np.testing.assert_array_equal(raw.radii, cat.radii)
