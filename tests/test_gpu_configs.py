"""GPU parity at the workloads' own sizes.

- 128^3 Zel'dovich particles in a 250 Mpc/h box, 128^3 grid, 50^3 mesh
  (5 Mpc/h cells), threshold -0.8, the 48-rung geometric 5% ladder from two
  grid cells to 40 Mpc/h: the whole catalogue and its count column against
  the oracle's sequential commit on the same smoothed fields, witness on and
  off, the reference's quicksort tie order and this build's, under the default
  growth batch window and a small one; and the whole path from the same
  particles against the oracle's pipeline (float CIC in the reference's
  order, FP64 transform): the same catalogue bit for bit.
- 512^3 particles in a 1000 Mpc/h box, 512^3 grid, 100^3 mesh (10 Mpc/h):
  the chain mesh (offsets, order, positions) and the fixed-point CIC bit for
  bit against the oracle; the FP32 smoothed field of the top and bottom rungs
  against an FP64 transform, max and rms error stated; 200 sampled voids of
  the catalogue against a brute-force radius over every particle; and the
  catalogue bit-identical over two runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THRESH = -0.8


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def _ladder(box, n_cells, rmax):
    from sif_amd import zeldovich
    return zeldovich.ladder(2.0 * box / n_cells, rmax, 0.05)


def _grid(D, dx, dy, dz, n_part, n, box):
    g = D.cic(dx, dy, dz, n_part, n, box)
    D.density_contrast(g, n ** 3)
    return g


# --- 128^3, L = 250 -----------------------------------------------------------

C1 = dict(n_side=128, box=250.0, n=128, mesh=50, seed=1, rmax=40.0)


@pytest.fixture(scope="module")
def config1(native):
    import oracle
    from sif_amd import device as D
    from sif_amd import zeldovich
    c = C1
    _say("128^3 workload")
    x, y, z = zeldovich.generate(c["n_side"], c["box"], c["seed"])
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    radii = _ladder(c["box"], c["n"], c["rmax"])
    assert len(radii) == 48
    m = D.Mesh(dx, dy, dz, len(x), c["mesh"], c["box"])
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, c["mesh"], c["box"]), c["mesh"], c["box"])
    g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
    fields = D.exodus_fields(g, c["n"], c["box"], radii)
    return D, (x, y, z), (dx, dy, dz), m, om, radii, list(fields)


def test_config1_catalogue_equals_oracle(config1):
    import oracle
    D, (x, y, z), (dx, dy, dz), m, om, radii, fields = config1
    c = C1
    ref = None
    for witness in (True, False):
        for tie_rule in (1, 0):
            ocat, onin, orung, _ = oracle.exodus_run(fields, c["n"], c["box"], om, len(x), radii,
                                                     THRESH, tie_rule=tie_rule, witness=witness)
            if ref is None:
                ref = (ocat, onin, orung)
            # equal depths do occur at this size; both tie orders give the same
            # voids here (the GPU takes the (depth, flat) order)
            np.testing.assert_array_equal(ocat.view(np.uint32), ref[0].view(np.uint32))
            np.testing.assert_array_equal(onin, ref[1])
    ocat, onin, orung = ref
    _say(f"oracle: {len(ocat)} voids")
    for window in (None, (64, 512)):
        if window:
            D.set_batch_window(*window)
        try:
            g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
            cat, nin, stats = D.exodus(g, c["n"], c["box"], m, radii, THRESH)
        finally:
            D.set_batch_window(0, 0)
        _say(f"GPU (window {window}): {len(cat)} voids, "
             f"{sum(s['evaluated'] for s in stats)} growths")
        np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
        np.testing.assert_array_equal(nin, onin)
        # the rung of each void from the per-rung accept counts (the catalogue
        # is in commit order) is the oracle's
        acc = np.array([s["accepted"] for s in stats])
        np.testing.assert_array_equal(np.repeat(np.arange(len(radii)), acc), orung)
    assert len(ocat) > 1000
    # the smoothed fields written and scanned instead of the cell lists:
    # the same catalogue, and the grid restored to float precision
    D.set_cell_lists(False)
    try:
        g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
        delta0 = g.get()
        cat, nin, _ = D.exodus(g, c["n"], c["box"], m, radii, THRESH)
    finally:
        D.set_cell_lists(True)
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)
    assert np.abs(g.get() - delta0).max() < 1e-4 * max(1.0, np.abs(delta0).max())
    # with the lists the grid is never written
    g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
    delta0 = g.get()
    D.exodus(g, c["n"], c["box"], m, radii, THRESH)
    np.testing.assert_array_equal(g.get().view(np.uint32), delta0.view(np.uint32))


def test_config1_catalogue_independent_of_void_room(config1):
    """The device copy of the catalogue grown many times within a run (room
    for 16 voids at first, doubled as needed) gives the catalogue of a run
    whose room was allocated once, under the default batch window and a small
    one (more rung starts issued ahead of their rung's last commits). Each
    rung reserves its room before its overlap index is built, so the appends
    that follow never move the arrays the index points at (a launch with an
    index older than the arrays is refused)."""
    D, (x, y, z), (dx, dy, dz), m, om, radii, fields = config1
    c = C1
    for window in (None, (64, 512)):
        if window:
            D.set_batch_window(*window)
        try:
            D.set_void_room(0)
            g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
            ref, ref_nin, _ = D.exodus(g, c["n"], c["box"], m, radii, THRESH)
            D.set_void_room(16)
            g = _grid(D, dx, dy, dz, len(x), c["n"], c["box"])
            cat, nin, _ = D.exodus(g, c["n"], c["box"], m, radii, THRESH)
        finally:
            D.set_void_room(0)
            D.set_batch_window(0, 0)
        _say(f"void room 16 (window {window}): {len(cat)} voids against {len(ref)}")
        assert len(ref) > 1000
        np.testing.assert_array_equal(cat.view(np.uint32), ref.view(np.uint32))
        np.testing.assert_array_equal(nin, ref_nin)


def _match(cat_a, cat_b):
    """Voids of two catalogues on the same node: (indices into a, into b).
    Centres are grid nodes (node index times the cell length, one float
    product), so a void found at the same node has bit-identical centres."""
    key_a = {tuple(v.view(np.uint32)[:3]): i for i, v in enumerate(cat_a)}
    ia, ib = [], []
    for j, v in enumerate(cat_b):
        i = key_a.get(tuple(v.view(np.uint32)[:3]))
        if i is not None:
            ia.append(i)
            ib.append(j)
    return np.array(ia, np.int64), np.array(ib, np.int64)


def test_config1_end_to_end_against_oracle_pipeline(config1):
    """The whole path from the same particles on both sides: the GPU's
    fixed-point CIC, contrast, chain mesh, FP32 Stockham smoothing, scan,
    growth and commit, against the oracle's pipeline -- the reference's float
    CIC in its slab order, an FP64 transform, the same finder. The fields
    differ in their last bits (FP32 against FP64 smoothing, fixed-point
    against float deposit), which could reorder candidates of nearly equal
    depth and change a decision; on this workload none does. Reported: void
    counts, the fraction of voids found at the same node, the radius error
    of those, and the CIC's difference from the reference's deposit order.
    Bar: the two catalogues equal bit for bit, count column included, and
    the deposit within 1e-5 of the peak."""
    import oracle
    D, (x, y, z), (dx, dy, dz), m, om, radii, _ = config1
    c = C1
    n, box = c["n"], c["box"]
    # the deposits: fixed point (GPU) against the reference's float order
    g = D.cic(dx, dy, dz, len(x), n, box)
    got = g.get().astype(np.float64)
    ref = oracle.cic_float(x, y, z, n, box).astype(np.float64)
    cic_err = float(np.abs(got - ref).max() / ref.max())
    D.density_contrast(g, n ** 3)
    cat, nin, _ = D.exodus(g, n, box, m, radii, THRESH, restore=False)
    ocat, onin, stage = oracle.exodus_pipeline(x, y, z, n, box, c["mesh"], radii, THRESH)
    ia, ib = _match(cat, ocat)
    frac_gpu = len(ia) / len(cat)
    frac_orc = len(ib) / len(ocat)
    rel = np.abs(cat[ia, 3].astype(np.float64) - ocat[ib, 3]) / ocat[ib, 3]
    same_r = float(np.mean(cat[ia, 3] == ocat[ib, 3]))
    within = float(np.mean(rel <= 1e-5))
    same_n = float(np.mean(nin[ia] == onin[ib]))
    _say(f"end to end at 128^3: GPU {len(cat)} voids, oracle pipeline {len(ocat)} "
         f"(delta {len(cat) - len(ocat):+d}); matched {len(ia)} "
         f"({100 * frac_gpu:.2f}% of the GPU's, {100 * frac_orc:.2f}% of the oracle's); "
         f"radii bit-equal {100 * same_r:.2f}%, within 1e-5 {100 * within:.2f}%, "
         f"max rel {rel.max():.3e}; counts equal {100 * same_n:.2f}%; "
         f"CIC vs float order max {cic_err:.2e} of peak; oracle stages {np.round(stage, 2)}")
    assert cic_err <= 1e-5
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)


# --- 512^3, L = 1000 ----------------------------------------------------------

C2 = dict(n_side=512, box=1000.0, n=512, mesh=100, seed=2, rmax=40.0)


@pytest.fixture(scope="module")
def config2(native):
    from sif_amd import device as D
    from sif_amd import zeldovich
    c = C2
    _say("512^3 workload")
    x, y, z = zeldovich.generate(c["n_side"], c["box"], c["seed"])
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    return D, (x, y, z), (dx, dy, dz)


def test_config2_chain_mesh_bit_exact(config2):
    import oracle
    D, (x, y, z), (dx, dy, dz) = config2
    c = C2
    m = D.Mesh(dx, dy, dz, len(x), c["mesh"], c["box"])
    want = oracle.chain_mesh(x, y, z, c["mesh"], c["box"])
    np.testing.assert_array_equal(m.offsets.get(), want["offsets"])
    np.testing.assert_array_equal(m.orig.get(), want["orig"])
    for a in "xyz":
        np.testing.assert_array_equal(getattr(m, a).get(), want[a])
    _say("mesh bit-exact at 512^3 / 100^3")


def test_config2_cic_bit_exact(config2):
    import oracle
    D, (x, y, z), (dx, dy, dz) = config2
    c = C2
    got = D.cic(dx, dy, dz, len(x), c["n"], c["box"]).get()
    want, _, _ = oracle.cic_fixed(x, y, z, c["n"], c["box"])
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    cell = c["box"] / c["n"]
    assert abs(got.astype(np.float64).sum() * cell ** 3 - len(x)) <= 1e-4 * len(x)
    _say("CIC bit-exact at 512^3")


def test_config2_smoothed_field_against_fp64(config2):
    """The top and bottom rungs' FP32 smoothed fields against the same filter
    (the reference's float factors, fft.c:753-780) on an FP64 transform of the
    same contrast. Bar: 2e-5 of the field's range; the measured errors are
    printed."""
    import oracle
    import scipy.fft
    D, _, (dx, dy, dz) = config2
    c = C2
    n, box = c["n"], c["box"]
    radii = _ladder(box, n, c["rmax"])
    g = _grid(D, dx, dy, dz, dx.shape[0], n, box)
    delta = g.get().reshape(n, n, n).astype(np.float64)
    spec64 = scipy.fft.rfftn(delta, workers=-1)
    del delta
    spec = D.spectrum(g, n)
    axis = oracle.window_axis(n, 2, 0)
    for r in (float(radii[0]), float(radii[-1])):
        got = D.smooth(spec, n, box, 1, r, window_order=2).get()
        fac = oracle.mode_factors(oracle.filter_lut(1, r, 0.0, box, n), axis, n)
        want = scipy.fft.irfftn(spec64 * fac, s=(n, n, n), workers=-1).ravel()
        del fac
        err = np.abs(got.astype(np.float64) - want)
        span = want.max() - want.min()
        rms = float(np.sqrt(np.mean(err * err)))
        _say(f"R = {r:.3f}: max error {err.max():.3e}, rms {rms:.3e}, range {span:.4f} "
             f"(max/range {err.max() / span:.2e})")
        assert err.max() <= 2e-5 * span
        del got, want, err


def _brute_radius(px, py, pz, n_total, box, centre, rung, r_search, threshold):
    """The crossing of tests/test_finders.c:433-465 in FP64 over the given
    particles, at the mean density of n_total particles: the largest
    distance d in (rung, r_search] with (particles within d, the one at d
    excluded) <= (1 + t) n_total / L^3 (4/3) pi d^3; -1 if none."""
    L = float(box)
    d = []
    for p, cc in ((px, centre[0]), (py, centre[1]), (pz, centre[2])):
        a = np.abs(p.astype(np.float64) - float(cc))
        d.append(np.where(a > 0.5 * L, L - a, a))
    dist = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    core = int(np.count_nonzero(dist <= rung))
    shell = np.sort(dist[(dist > rung) & (dist <= r_search)])
    k = (1.0 + threshold) * n_total / L ** 3 * 4.0 / 3.0 * np.pi
    ok = (core + np.arange(len(shell))) <= k * shell ** 3
    idx = np.nonzero(ok)[0]
    return float(shell[idx[-1]]) if len(idx) else -1.0


def test_config2_catalogue_repeatable_and_radii_brute_force(config2):
    D, (x, y, z), (dx, dy, dz) = config2
    c = C2
    n, box = c["n"], c["box"]
    radii = _ladder(box, n, c["rmax"])
    m = D.Mesh(dx, dy, dz, len(x), c["mesh"], box, indices=False)
    runs = []
    for _ in range(2):
        g = _grid(D, dx, dy, dz, len(x), n, box)
        runs.append(D.exodus(g, n, box, m, radii, THRESH, restore=False))
    (cat, nin, stats), (cat2, nin2, _) = runs
    _say(f"512^3 catalogue: {len(cat)} voids")
    np.testing.assert_array_equal(cat.view(np.uint32), cat2.view(np.uint32))
    np.testing.assert_array_equal(nin, nin2)
    # the same catalogue from smoothed fields written and scanned
    D.set_cell_lists(False)
    try:
        g = _grid(D, dx, dy, dz, len(x), n, box)
        cat3, nin3, _ = D.exodus(g, n, box, m, radii, THRESH, restore=False)
    finally:
        D.set_cell_lists(True)
    np.testing.assert_array_equal(cat3.view(np.uint32), cat.view(np.uint32))
    np.testing.assert_array_equal(nin3, nin)
    del m
    # 200 voids against a brute force over every particle of the 3^3 coarse
    # cells around them (62.5 Mpc/h cells, wider than any search radius)
    rung_of = np.repeat(np.arange(len(radii)), [s["accepted"] for s in stats])
    assert len(rung_of) == len(cat)
    nc = 16
    cw = box / nc
    key = ((np.minimum((x / cw).astype(np.int64), nc - 1) * nc +
            np.minimum((y / cw).astype(np.int64), nc - 1)) * nc +
           np.minimum((z / cw).astype(np.int64), nc - 1))
    order = np.argsort(key, kind="stable")
    starts = np.searchsorted(key[order], np.arange(nc ** 3 + 1))
    cell = box / n
    picks = np.random.default_rng(2024).choice(len(cat), 200, replace=False)
    worst = 0.0
    for v in picks:
        cx, cy, cz, r = (float(a) for a in cat[v])
        k = int(rung_of[v])
        rung = float(radii[k])
        rs = rung + max(0.5 * rung, 3.0 * cell)
        if k > 0:
            rs = max(rs, float(radii[k - 1]))
        assert rs < cw
        ci = [int(a / cw) % nc for a in (cx, cy, cz)]
        sel = []
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for d in (-1, 0, 1):
                    q = (((ci[0] + a) % nc) * nc + (ci[1] + b) % nc) * nc + (ci[2] + d) % nc
                    sel.append(order[starts[q]:starts[q + 1]])
        sel = np.concatenate(sel)
        want = _brute_radius(x[sel], y[sel], z[sel], len(x), box, (cx, cy, cz), rung, rs,
                             THRESH)
        assert want > 0, (v, cat[v], rung, rs)
        rel = abs(r - want) / want
        worst = max(worst, rel)
        assert rel <= 1e-5, (v, cat[v], want)
    _say(f"200 voids: worst radius error vs brute force {worst:.2e}")
