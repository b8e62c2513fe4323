"""The survey finder on the GPU (survey.hip, grow_survey.h, the survey run of
exodus.hip, sif_finder_exodus_survey_raw) against the oracle
(oracle/oracle_survey.c), on the reference's own survey scenario
(tests/test_survey.c: a footprint ball with a masked stripe, carved holes,
a radial selection, dense and sparse randoms).

Bars: integer and byte work bit-exact (the observed nodes, their counts,
the footprint fractions); the survey radius and the catalogue bit-exact
against the oracle fed the GPU's own smoothed fields; the holes recovered as
the reference's test asks; weights of 2 and 0.5 reproducing the unweighted
catalogue bit for bit.
"""
import numpy as np
import pytest

import oracle
from survey_common import (BOX, CELL, N_DATA, N_GRID, N_RANDOMS, N_SPARSE, RADII, THRESHOLD,
                           check_holes, draw, suggest_mesh_cells)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[survey] {msg}", flush=True)


class Tracers:
    """Positions (host and device), their CIC density grid and mesh."""

    def __init__(self, D, pos, w=None, n=N_GRID, box=BOX):
        self.pos = pos
        self.w = w
        self.d = [D.to_device(a) for a in pos]
        self.dw = D.to_device(w) if w is not None else None
        self.n = len(pos[0])
        self.grid = D.cic(*self.d, self.n, n, box, dw=self.dw)
        self.grid0 = self.grid.get()  # a run restores its grid only up to rounding
        self.mesh_cells = suggest_mesh_cells(self.n, box, float(np.max(RADII)))
        self.mesh = D.Mesh(*self.d, self.n, self.mesh_cells, box, dw=self.dw, indices=False)

    def reset(self):
        self.grid.set(self.grid0)
        return self

    def oracle_mesh(self, box=BOX):
        x, y, z = self.pos
        m = oracle.chain_mesh(x, y, z, self.mesh_cells, box, w=self.w)
        return oracle.OracleMesh(m, self.mesh_cells, box)


CASES = {
    "flat-dense": (False, N_RANDOMS),
    "selection-dense": (True, N_RANDOMS),
    "selection-sparse": (True, N_SPARSE),
}


@pytest.fixture(scope="module")
def scenario(native):
    from sif_amd import device as D
    out = {}
    for name, (selection, n_r) in CASES.items():
        rng = np.random.default_rng(40 + len(out))
        data = draw(N_DATA, selection, True, rng)
        randoms = draw(n_r, selection, False, rng)
        out[name] = (Tracers(D, data), Tracers(D, randoms))
    return D, out


def _gpu_fields(D, data, randoms, occ_host, alpha):
    """The survey contrast of every rung from the GPU's own smoothing."""
    sd = D.exodus_fields(data.grid, N_GRID, BOX, RADII)
    sr = D.exodus_fields(randoms.grid, N_GRID, BOX, RADII)
    return [oracle.survey_contrast(a, b, occ_host, alpha) for a, b in zip(sd, sr)]


@pytest.mark.parametrize("case", list(CASES))
def test_occupancy_bit_exact(scenario, case):
    D, sc = scenario
    _, randoms = sc[case]
    m = randoms.mesh
    occ, k, lam, marked, observed = D.survey_occupancy(m.x, m.y, m.z, m.n, N_GRID, BOX)
    want, k2, lam2, marked2, observed2 = oracle.survey_occupancy(
        m.x.get(), m.y.get(), m.z.get(), N_GRID, BOX)
    _say(f"{case}: k {k}, lambda {lam:.3f}, {marked} marked, {observed} observed")
    assert (k, lam, marked, observed) == (k2, lam2, marked2, observed2)
    np.testing.assert_array_equal(occ.get(), want)
    assert k >= 1 and observed >= marked


def test_contrast_bit_exact(native):
    from sif_amd import device as D
    rng = np.random.default_rng(7)
    n = 32
    data = rng.uniform(0.0, 3.0, n ** 3).astype(np.float32)
    rand = rng.uniform(0.0, 2.0, n ** 3).astype(np.float32)
    rand[::17] = 0.0
    occ = (rng.random(n ** 3) < 0.8).astype(np.uint8)
    alpha = np.float32(0.137)
    dd, dr, do = D.to_device(data), D.to_device(rand), D.to_device(occ, np.uint8)
    D.survey_contrast(dd, dr, do, alpha)
    want = oracle.survey_contrast(data, rand, occ, alpha)
    np.testing.assert_array_equal(dd.get().view(np.uint32), want.view(np.uint32))


def test_footprint_bit_exact(scenario):
    D, sc = scenario
    _, randoms = sc["selection-sparse"]
    m = randoms.mesh
    occ, *_ = D.survey_occupancy(m.x, m.y, m.z, m.n, N_GRID, BOX)
    rng = np.random.default_rng(8)
    v = np.c_[rng.uniform(0.0, BOX, (300, 3)), rng.uniform(2.0, 45.0, 300)].astype(np.float32)
    fin, fsh = D.survey_footprint(occ, N_GRID, BOX, v)
    win, wsh = oracle.survey_footprint(occ.get(), N_GRID, BOX, v)
    np.testing.assert_array_equal(fin.view(np.uint32), win.view(np.uint32))
    np.testing.assert_array_equal(fsh.view(np.uint32), wsh.view(np.uint32))
    assert 0.0 < fin.mean() < 1.0


@pytest.mark.parametrize("weighted", [False, True])
def test_find_radius_survey_bit_exact(native, weighted):
    from sif_amd import device as D
    box = 200.0
    rng = np.random.default_rng(60 + weighted)
    dp = (rng.random((20000, 3)) * box).astype(np.float32)
    holes = rng.uniform(40.0, 160.0, (12, 3))
    for h in holes:
        d = np.sqrt(((dp - h) ** 2).sum(axis=1))
        dp = dp[(d > 18.0) | (rng.random(len(dp)) < 0.1)]
    rp = (rng.random((120000, 3)) * box).astype(np.float32)
    dw = rng.uniform(0.5, 2.0, len(dp)).astype(np.float32) if weighted else None
    rw = rng.uniform(0.2, 1.0, len(rp)).astype(np.float32) if weighted else None
    dev = [D.to_device(np.ascontiguousarray(dp[:, k])) for k in range(3)]
    rev = [D.to_device(np.ascontiguousarray(rp[:, k])) for k in range(3)]
    dmesh = D.Mesh(*dev, len(dp), 12, box, dw=D.to_device(dw) if weighted else None,
                   indices=False)
    rmesh = D.Mesh(*rev, len(rp), 20, box, dw=D.to_device(rw) if weighted else None,
                   indices=False)
    omd = oracle.OracleMesh(oracle.chain_mesh(dp[:, 0], dp[:, 1], dp[:, 2], 12, box, w=dw),
                            12, box)
    omr = oracle.OracleMesh(oracle.chain_mesh(rp[:, 0], rp[:, 1], rp[:, 2], 20, box, w=rw),
                            20, box)
    alpha = np.float32(omd.s.total_weight / omr.s.total_weight)
    K = float(np.float32(np.float32(1.0 - 0.7) * alpha))
    cen = np.vstack([holes + rng.normal(0.0, 3.0, holes.shape),
                     rng.uniform(0.0, box, (250, 3))]).astype(np.float32)
    for rmin, rs in ((4.0, 30.0), (9.0, 25.0)):
        got, reason, nin = D.find_radius_survey(dmesh, rmesh, cen, rmin, rs, K)
        want, wreason, wnin = oracle.find_radius_survey(omd, omr, cen, rmin, rs, K, window=8)
        _say(f"radius weighted={weighted} rmin {rmin}: {int((got > 0).sum())} found")
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(reason, wreason)
        np.testing.assert_array_equal(nin, wnin)
        assert (got > 0).sum() >= 5


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("window", [None, (16, 64)])
def test_survey_catalogue_matches_oracle(scenario, case, window):
    D, sc = scenario
    data, randoms = sc[case]
    data.reset()
    if window:
        D.set_batch_window(*window)
    try:
        cat, fp, n_in, stats = D.exodus_survey(data.grid, randoms.grid, N_GRID, BOX, data.mesh,
                                               randoms.mesh, RADII, THRESHOLD, restore=True)
    finally:
        D.set_batch_window(0, 0)
    m = randoms.mesh
    occ, *_ = D.survey_occupancy(m.x, m.y, m.z, m.n, N_GRID, BOX)
    occ_h = occ.get()
    alpha = np.float32(data.mesh.total_weight / randoms.mesh.total_weight)
    fields = _gpu_fields(D, data.reset(), randoms, occ_h, alpha)
    K = float(np.float32(np.float32(1.0 + np.float32(THRESHOLD)) * alpha))
    desc = np.sort(RADII)[::-1].copy()
    want, wnin, _, _ = oracle.exodus_run_survey(fields, N_GRID, BOX, data.oracle_mesh(),
                                                randoms.oracle_mesh(), K, desc, THRESHOLD)
    wfin, wfsh = oracle.survey_footprint(occ_h, N_GRID, BOX, want)
    _say(f"{case} window={window}: {len(cat)} voids")
    np.testing.assert_array_equal(cat.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(fp[:, 0].view(np.uint32), wfin.view(np.uint32))
    np.testing.assert_array_equal(fp[:, 1].view(np.uint32), wfsh.view(np.uint32))
    np.testing.assert_array_equal(n_in, wnin)
    check_holes(cat, fp, case)


def test_survey_weights_scale_exactly(scenario):
    from sif_amd import device as D
    _, sc = scenario
    data, randoms = sc["selection-dense"]
    data.reset()
    a, fa, _, _ = D.exodus_survey(data.grid, randoms.grid, N_GRID, BOX, data.mesh, randoms.mesh,
                                  RADII, THRESHOLD)
    wd = Tracers(D, data.pos, w=np.full(data.n, 2.0, np.float32))
    wr = Tracers(D, randoms.pos, w=np.full(randoms.n, 0.5, np.float32))
    b, fb, _, _ = D.exodus_survey(wd.grid, wr.grid, N_GRID, BOX, wd.mesh, wr.mesh, RADII,
                                  THRESHOLD)
    assert len(a) >= 3
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(fa.view(np.uint32), fb.view(np.uint32))


def test_survey_restores_the_data_grid(scenario):
    from sif_amd import device as D
    _, sc = scenario
    data, randoms = sc["flat-dense"]
    data.reset()
    before_d, before_r = data.grid.get(), randoms.grid.get()
    D.exodus_survey(data.grid, randoms.grid, N_GRID, BOX, data.mesh, randoms.mesh, RADII,
                    THRESHOLD, restore=True)
    after_d, after_r = data.grid.get(), randoms.grid.get()
    np.testing.assert_array_equal(after_r, before_r)  # only ever read
    scale = np.abs(before_d).max()
    assert np.abs(after_d - before_d).max() <= 1e-3 * scale  # test_survey.c:315-325


def _largest_prime(n):
    p, best = 2, 1
    while n > 1:
        while n % p == 0:
            n //= p
            best = p
        p += 1
    return best


def _pysif_tracers(pysif, pos, box, n=N_GRID):
    f = pysif.field_from_numpy(*pos)
    g = pysif.Grid(n, box)
    g.assign_cic(f)
    m = pysif.ChainMesh(suggest_mesh_cells(len(pos[0]), box, float(np.max(RADII))), box, f,
                        drop_indices=True)
    return f, g, m


def test_survey_dropin_in_the_callers_frame(native):
    """survey_box, translate, exodus_survey through the host C API: the voids
    come back in the caller's frame, as the device run on the moved tracers
    gives them moved back (test_survey.c:442-560)."""
    from sif_amd import device as D
    from sif_amd import pysif
    from sif_amd.pysif import finders
    pysif.system_init()
    rng = np.random.default_rng(70)
    shift = np.float32(-500.0)  # the caller's frame: the survey around (-300, -300, -300)
    data = draw(N_DATA, True, True, rng, offset=shift)
    randoms = draw(N_RANDOMS, True, False, rng, offset=shift)
    fd = pysif.field_from_numpy(*data)
    fr = pysif.field_from_numpy(*randoms)
    offset, box = finders.survey_box(fd, fr, RADII)
    _say(f"survey box {box:.1f}, offset {offset}")
    fd.translate(offset)
    fr.translate(offset)
    moved_d = [np.asarray(a, np.float32) + np.float32(o) for a, o in zip(data, offset)]
    moved_r = [np.asarray(a, np.float32) + np.float32(o) for a, o in zip(randoms, offset)]
    n = int(np.ceil(box / CELL))
    while _largest_prime(n) > 7:
        n += 1  # a grid size the FFT factors into small radices
    gd, gr = pysif.Grid(n, box), pysif.Grid(n, box)
    gd.assign_cic(fd)
    gr.assign_cic(fr)
    md = pysif.ChainMesh(suggest_mesh_cells(len(data[0]), box, float(np.max(RADII))), box, fd,
                         drop_indices=True)
    mr = pysif.ChainMesh(suggest_mesh_cells(len(randoms[0]), box, float(np.max(RADII))), box,
                         fr, drop_indices=True)
    cat = finders.exodus_survey_raw(gd, gr, md, mr, offset, RADII, THRESHOLD, checkpoint=False)
    got = np.c_[cat.centres, cat.radii]
    fp = np.c_[cat.footprint, cat.footprint_shell]
    td, tr = Tracers(D, moved_d, n=n, box=box), Tracers(D, moved_r, n=n, box=box)
    want, wfp, _, _ = D.exodus_survey(td.grid, tr.grid, n, box, td.mesh, tr.mesh, RADII,
                                      THRESHOLD)
    want[:, :3] -= np.asarray(offset, np.float32)
    np.testing.assert_array_equal(got.astype(np.float32), want)
    np.testing.assert_array_equal(fp.astype(np.float32), wfp)
    inner = [h + np.r_[shift, shift, shift, 0.0] for h in
             np.array([[160.0, 160.0, 160.0, 26.0], [240.0, 160.0, 240.0, 22.0],
                       [150.0, 240.0, 220.0, 20.0]])]
    for h in inner:
        d = np.sqrt(((got[:, :3] - h[:3]) ** 2).sum(axis=1))
        assert d.min() <= 2.0 * CELL


def test_survey_refusals(native):
    """test_survey.c:762-828: box-only settings, a contrast grid, a missing
    mesh, too little padding -- each refused before anything runs."""
    from sif_amd import pysif
    from sif_amd.pysif import finders
    pysif.system_init()
    rng = np.random.default_rng(80)
    data = draw(20000, False, True, rng)
    randoms = draw(100000, False, False, rng)
    fd, gd, md = _pysif_tracers(pysif, data, BOX)
    fr, gr, mr = _pysif_tracers(pysif, randoms, BOX)
    with pytest.raises(ValueError):
        finders.exodus_survey_raw(gd, gr, md, mr, (0, 0, 0), RADII, THRESHOLD, seeds="minima",
                                  checkpoint=False)
    # the footprint ball reaches 60 from the faces; a ladder to 60 needs more
    with pytest.raises(ValueError, match="of padding on every side"):
        finders.exodus_survey_raw(gd, gr, md, mr, (0, 0, 0), [60.0, 40.0], THRESHOLD,
                                  checkpoint=False)
    gd.to_density_contrast()
    with pytest.raises(ValueError, match="holds a density contrast"):
        finders.exodus_survey_raw(gd, gr, md, mr, (0, 0, 0), RADII, THRESHOLD,
                                  checkpoint=False)


def test_survey_checkpoint_kept_resumed_and_refused(native, tmp_path):
    """A survey's checkpoint (exodus_finder.c:1435-1575, exodus_checkpoint.c:
    417-430): kept on request, read back in the caller's frame, resumed to the
    same catalogue, refused for another offset."""
    from sif_amd import pysif
    from sif_amd.pysif import finders
    pysif.system_init()
    rng = np.random.default_rng(90)
    data = draw(40000, True, True, rng)
    randoms = draw(400000, True, False, rng)
    fd, gd, md = _pysif_tracers(pysif, data, BOX)
    fr, gr, mr = _pysif_tracers(pysif, randoms, BOX)
    off = (0.0, 0.0, 0.0)
    path = tmp_path / "survey.ckpt"
    cat = finders.exodus_survey_raw(gd, gr, md, mr, off, RADII, THRESHOLD, checkpoint=str(path),
                                    keep_checkpoint=True)
    full = np.c_[cat.centres, cat.radii]
    fp = np.c_[cat.footprint, cat.footprint_shell]
    assert path.exists() and len(full) >= 3
    voids, info = finders.read_checkpoint(path)
    assert info["mode"] == "survey" and info["rungs_done"] == len(RADII)
    np.testing.assert_array_equal(np.c_[voids.centres, voids.radii], full)
    again = finders.exodus_survey_raw(gd, gr, md, mr, off, RADII, THRESHOLD,
                                      checkpoint=str(path), keep_checkpoint=True)
    np.testing.assert_array_equal(np.c_[again.centres, again.radii], full)
    np.testing.assert_array_equal(np.c_[again.footprint, again.footprint_shell], fp)
    with pytest.raises(ValueError, match="offset"):
        finders.exodus_survey_raw(gd, gr, md, mr, (1.0, 0.0, 0.0), RADII, THRESHOLD,
                                  checkpoint=str(path))
