"""The oracle's survey restatement (oracle/oracle_survey.c), pinned on the CPU:

- lambda against its closed form m = lambda / (1 - exp(-lambda));
- the closing against scipy's binary morphology with a box structuring
  element (nodes past a face unset), and the window rule k;
- the marked nodes against a numpy nearest-node count;
- the contrast on constructed fields (ratios, unobserved and empty nodes);
- the footprint against a full mask, a half-space (spherical cap volumes)
  and a void cut by the box face;
- the survey radius against a brute force over every tracer (D and R
  enclosed, R interpolated within its bin), unweighted and weighted;
- the whole run on the reference's own survey scenario (tests/test_survey.c:
  holes recovered, footprints), and its weights check (data weighing 2 and
  randoms 0.5 give the unweighted catalogue bit for bit).
"""
import numpy as np
import pytest

import oracle
from survey_common import (BOX, N_DATA, N_GRID, N_RANDOMS, N_SPARSE, RADII, check_holes,
                           cpu_survey, draw)


def test_randoms_per_node_closed_form():
    for lam in (0.01, 0.3, 1.0, 3.0, 20.0, 150.0):
        m = lam / (1.0 - np.exp(-lam))
        assert abs(oracle.survey_randoms_per_node(m) - lam) <= 1e-9 * max(lam, 1.0)
    assert oracle.survey_randoms_per_node(1.0) == 0.0
    assert oracle.survey_randoms_per_node(0.5) == 0.0


@pytest.mark.parametrize("density", [0.3, 0.8])
@pytest.mark.parametrize("k", [1, 2, 4])
def test_closing_matches_scipy(density, k):
    from scipy import ndimage
    n = 20
    rng = np.random.default_rng(int(density * 10) + k)
    occ = (rng.random((n, n, n)) < density).astype(np.uint8)
    occ[:3] = 0  # an empty slab, so faces and gaps both show
    box = np.ones((2 * k + 1,) * 3, bool)
    want = ndimage.binary_erosion(ndimage.binary_dilation(occ.astype(bool), box, border_value=0),
                                  box, border_value=0)
    got = oracle.survey_close(occ.ravel(), n, k).reshape(n, n, n)
    np.testing.assert_array_equal(got.astype(bool), want)
    # a plausible slip -- faces counted as set in the erosion -- differs
    wrong = ndimage.binary_erosion(ndimage.binary_dilation(occ.astype(bool), box, border_value=0),
                                   box, border_value=1)
    assert (wrong != want).any()


def test_occupancy_marks_and_window_rule():
    n, box = 32, 100.0
    rng = np.random.default_rng(3)
    for n_r in (3000, 40000, 400000):
        p = (rng.random((n_r, 3)) * np.float32(box * 0.999)).astype(np.float32)
        occ, k, lam, marked, observed = oracle.survey_occupancy(p[:, 0], p[:, 1], p[:, 2], n, box)
        inv_l = np.float32(1.0) / (np.float32(box) / np.float32(n))
        idx = np.minimum((p * inv_l + np.float32(0.5)).astype(np.uint32), n - 1).astype(np.int64)
        flat = (idx[:, 0] * n + idx[:, 1]) * n + idx[:, 2]
        assert marked == len(np.unique(flat))
        lam_want = oracle.survey_randoms_per_node(n_r / marked)
        assert lam == lam_want
        k_want = 1
        while k_want < 4 and lam * (2 * k_want + 1) ** 3 < 128.0:
            k_want += 1
        assert k == k_want
        base = np.zeros(n ** 3, np.uint8)
        base[flat] = 1
        np.testing.assert_array_equal(occ, oracle.survey_close(base, n, k))
        assert observed == int(occ.sum())


def test_contrast_constructed():
    n = 8
    rng = np.random.default_rng(4)
    rand = rng.uniform(0.5, 2.0, n ** 3).astype(np.float32)
    occ = np.ones(n ** 3, np.uint8)
    occ[:50] = 0
    rand[60:70] = 0.0
    alpha = np.float32(0.25)
    data = (np.float32(3.0) * alpha * rand).astype(np.float32)  # contrast 2 up to rounding
    got = oracle.survey_contrast(data, rand, occ, alpha)
    assert (got[:50] == np.float32(1e30)).all() and (got[60:70] == np.float32(1e30)).all()
    live = np.ones(n ** 3, bool)
    live[:50] = False
    live[60:70] = False
    np.testing.assert_allclose(got[live], 2.0, rtol=1e-6)


def test_footprint_geometry():
    n, box = 64, 400.0
    l = box / n
    full = np.ones(n ** 3, np.uint8)
    fin, fsh = oracle.survey_footprint(full, n, box, [[200.0, 200.0, 200.0, 30.0]])
    assert fin[0] == 1.0 and fsh[0] == 1.0
    # a half-space: nodes with x below x0 observed; observed fraction of the
    # sphere and of the shell from the spherical caps
    top = 33  # nodes 0..top observed: points below x0 = (top + 1/2) l
    x0 = (top + 0.5) * l
    half = np.zeros((n, n, n), np.uint8)
    half[:top + 1] = 1
    def below(R, a):  # volume of a ball of radius R below x = a (centre 0)
        a = np.clip(a, -R, R)
        return np.pi * (R ** 2 * (a + R) - (a ** 3 + R ** 3) / 3.0)

    # the lattice resolves the plane to a few percent for one centre; over
    # centres spread across a lattice period the errors average out
    for r in (30.0, 22.0):
        cxs = 200.0 + np.linspace(-6.0, 6.0, 17)
        v = np.array([[c, 200.0, 200.0, r] for c in cxs])
        fin, fsh = oracle.survey_footprint(half.ravel(), n, box, v)
        a = x0 - cxs
        want_in = below(r, a) / (4.0 / 3.0 * np.pi * r ** 3)
        want_sh = (below(2 * r, a) - below(r, a)) / (4.0 / 3.0 * np.pi * (8 * r ** 3 - r ** 3))
        assert np.abs(fin - want_in).max() < 0.05 and np.abs(fsh - want_sh).max() < 0.05
        assert abs((fin - want_in).mean()) < 0.01 and abs((fsh - want_sh).mean()) < 0.01
    # cut by the box face: everything past it unobserved
    fin, fsh = oracle.survey_footprint(full, n, box, [[10.0, 200.0, 200.0, 30.0]])
    want = below(30.0, -10.0) / (4.0 / 3.0 * np.pi * 30.0 ** 3)  # the part past x = 0
    assert abs((1.0 - fin[0]) - want) < 0.03


def _brute_survey_radius(dp, dw, rp, rw, c, rmin, r_search, K):
    """Outermost data tracer d in the annulus with D(<d) <= K R(<d): D the
    data weight strictly inside it (itself not counted; ties below it in the
    (d2, w) order count), R the random weight of the core plus the bins below
    its own and its bin's weight times where it sits in the bin, in d^2."""
    f = np.float32
    rs2 = f(f(r_search) * f(r_search))
    rc2 = f(f(rmin) * f(rmin))
    span = f(rs2 - rc2)

    def d2_of(p):
        d = (p - c.astype(np.float32)).astype(np.float32)
        return ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)

    dd2, rd2 = d2_of(dp), d2_of(rp)
    nd_within = int((dd2 <= rs2).sum())
    d_within = float(dw[dd2 <= rs2].astype(np.float64).sum())
    r_within = float(rw[rd2 <= rs2].astype(np.float64).sum())
    if not r_within > 0.0:
        return -1.0, 6
    if f(d_within) <= f(f(K) * f(r_within)):
        return -1.0, 2
    cf = f(f(f(rmin) * f(rmin)) * f(rmin)) / f(f(f(r_search) * f(r_search)) * f(r_search))
    want = f(f(f(nd_within) * f(f(1.0) - f(cf))) / f(16.0))
    nb = 64 if not want >= 64.0 else (int(want) if want < 2048.0 else 2048)
    inv_w = f(f(nb) / span)
    bin_w = f(span / f(nb))

    def bin_of(d2):
        b = ((d2 - rc2) * inv_w).astype(np.float32).astype(np.int64)
        return np.clip(b, 0, nb - 1)

    ra = (rd2 > rc2) & (rd2 <= rs2)
    rb = np.zeros(nb)
    np.add.at(rb, bin_of(rd2[ra]), rw[ra].astype(np.float64))
    r_core = float(rw[rd2 <= rc2].astype(np.float64).sum())
    rpre = r_core + np.concatenate([[0.0], np.cumsum(rb)[:-1]])
    ann = (dd2 > rc2) & (dd2 <= rs2)
    d_total = float(dw[dd2 <= rs2].astype(np.float64).sum())
    order = np.lexsort((dw[ann], dd2[ann]))
    a2, aw = dd2[ann][order], dw[ann][order].astype(np.float64)
    above = np.cumsum(aw[::-1])[::-1]  # weight from k to the top, inclusive
    for k in range(len(a2) - 1, -1, -1):
        if a2[k] == 0.0:
            continue
        b = int(bin_of(np.array([a2[k]]))[0])
        frac = f(f(a2[k] - f(rc2 + f(f(b) * bin_w))) * inv_w)
        frac = min(max(frac, f(0.0)), f(1.0))
        r_in = f(rpre[b] + rb[b] * float(frac))
        d_in = f(d_total - above[k])
        if r_in > 0.0 and d_in <= f(f(K) * r_in):
            return float(np.sqrt(a2[k])), 0
    return -1.0, 3


@pytest.mark.parametrize("weighted", [False, True])
def test_survey_radius_against_brute_force(weighted):
    box, mc = 200.0, 10
    rng = np.random.default_rng(11 + weighted)
    nd, nr = 6000, 40000
    dp = (rng.random((nd, 3)) * box).astype(np.float32)
    holes = np.array([[70.0, 70.0, 70.0], [130.0, 120.0, 100.0], [100.0, 140.0, 60.0]])
    for h in holes:  # underdense pockets, so radii are found
        d = np.sqrt(((dp - h) ** 2).sum(axis=1))
        dp = dp[(d > 22.0) | (rng.random(len(dp)) < 0.1)]
    rp = (rng.random((nr, 3)) * box).astype(np.float32)
    dw = rng.uniform(0.5, 2.0, len(dp)).astype(np.float32) if weighted else None
    rw = rng.uniform(0.2, 1.0, len(rp)).astype(np.float32) if weighted else None
    dm = oracle.chain_mesh(dp[:, 0], dp[:, 1], dp[:, 2], mc, box, w=dw)
    rm = oracle.chain_mesh(rp[:, 0], rp[:, 1], rp[:, 2], 14, box, w=rw)
    omd, omr = oracle.OracleMesh(dm, mc, box), oracle.OracleMesh(rm, 14, box)
    alpha = np.float32(dm["total_weight"] / rm["total_weight"])
    K = float(np.float32(np.float32(1.0 - 0.6) * alpha))
    # centres well inside the box: no periodic image reaches them
    cen = np.vstack([holes + rng.normal(0, 2.0, holes.shape),
                     60.0 + rng.random((40, 3)) * 80.0]).astype(np.float32)
    got, reason, _ = oracle.find_radius_survey(omd, omr, cen, 6.0, 36.0, K, window=4)
    dwt = dw if weighted else np.ones(len(dp), np.float32)
    rwt = rw if weighted else np.ones(len(rp), np.float32)
    found = 0
    for i, c in enumerate(cen):
        want, wr = _brute_survey_radius(dp, dwt, rp, rwt, c, 6.0, 36.0, K)
        assert got[i] == np.float32(want), (i, got[i], want, reason[i], wr)
        assert reason[i] == wr
        found += want > 0
    assert found >= 3


@pytest.mark.parametrize("selection,n_randoms", [(False, N_RANDOMS), (True, N_RANDOMS),
                                                  (True, N_SPARSE)])
def test_survey_scenario_recovers_the_holes(selection, n_randoms):
    rng = np.random.default_rng(20 + selection + (n_randoms == N_SPARSE))
    data = draw(N_DATA, selection, True, rng)
    randoms = draw(n_randoms, selection, False, rng)
    cat, fp, occ, alpha = cpu_survey(data, randoms, oracle)
    check_holes(cat, fp, f"selection={selection}, randoms={n_randoms}")


def test_survey_weights_scale_exactly():
    """Data weighing 2 and randoms 0.5 scale alpha by 4 and every compared sum
    by a power of two: the unweighted catalogue, footprint included, bit for
    bit (test_survey.c:343-414)."""
    rng = np.random.default_rng(30)
    data = draw(N_DATA // 2, True, True, rng)
    randoms = draw(N_RANDOMS // 2, True, False, rng)
    a, fa, _, _ = cpu_survey(data, randoms, oracle)
    wd = np.full(len(data[0]), 2.0, np.float32)
    wr = np.full(len(randoms[0]), 0.5, np.float32)
    b, fb, _, _ = cpu_survey(data, randoms, oracle, wd=wd, wr=wr)
    assert len(a) >= 3
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(fa.view(np.uint32), fb.view(np.uint32))
