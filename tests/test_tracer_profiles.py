"""Stacked radial profiles over the tracers, sif_profiles
(src/measure/profiles.c:354-673, include/sif/measure/profiles.h).

CPU: the oracle's restatement against a brute force over every tracer in
FP64 (minimum-image or open distances; counts per bin equal but for tracers
within float rounding of a bin edge, shell-mean radial velocities to 1e-4 of
the largest speed), against the identity between its cumulative and
differential profiles, on a periodic and an open box, on voids large enough
that the separation is wrapped per tracer (profiles.c:521-530); the C API's
refusals (profiles.c:97-142), which come before any device work.

GPU: the device path against the oracle on the same mesh -- unweighted
densities bit for bit (counts are integers on both sides), weighted
densities and velocities to the rounding of the reference's float running
sums (1e-5 of the enclosed mass; 1e-4 of the largest speed) -- periodic and
open, cumulative and differential, more bins than one launch accumulates,
a void of no radius keeping its row; and pysif.measure.profiles over a field
with velocities and weights.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

N_SIDE, BOX, SEED, MESH = 32, 100.0, 11, 10
PBC_OPEN, DIFFERENTIAL, VEL_WEIGHTED = 1 << 0, 1 << 8, 1 << 9


@pytest.fixture(scope="module")
def tracers():
    from sif_amd import zeldovich
    x, y, z = zeldovich.generate(N_SIDE, BOX, SEED)
    rng = np.random.default_rng(7)
    v = rng.normal(0.0, 300.0, (3, len(x))).astype(np.float32)
    w = rng.uniform(0.5, 2.0, len(x)).astype(np.float32)
    return x, y, z, v, w


def oracle_mesh(tracers, weighted=False):
    x, y, z, v, w = tracers
    m = oracle.chain_mesh(x, y, z, MESH, BOX, w=w if weighted else None)
    o = m["orig"].astype(np.int64)
    m["vx"], m["vy"], m["vz"] = (np.ascontiguousarray(v[k][o]) for k in range(3))
    return m, oracle.OracleMesh(m, MESH, BOX)


def some_voids(seed=3, n=40):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, BOX, (n, 3))
    r = rng.uniform(3.0, 12.0, n)
    voids = np.c_[c, r].astype(np.float32)
    voids[5, 3] = 9.5    # r_max 47.5: with a 10-unit cell, wrapped per tracer
    voids[7, 3] = 0.0    # no radius: the row is not written
    voids[8, :3] = (0.2, 99.9, 50.0)  # at the faces
    return voids


def brute(tracers, voids, ext, n_bins, periodic=True, weights=None):
    """FP64 over every tracer: counts (and weight) per bin and the shell mean
    radial velocity, for each void."""
    x, y, z, v, w = tracers
    p = np.c_[x, y, z].astype(np.float64)
    vv = v.T.astype(np.float64)
    counts, means, mass = [], [], []
    for cx, cy, cz, rv in voids.astype(np.float64):
        d = p - (cx, cy, cz)
        if periodic:
            d = d - BOX * np.round(d / BOX)
        r = np.sqrt((d * d).sum(1))
        rmax = rv * ext
        m = (r <= rmax) & (rv > 0)
        b = np.floor(r[m] / (rmax / n_bins)).astype(np.int64)
        keep = b < n_bins
        idx = np.nonzero(m)[0][keep]
        b = b[keep]
        counts.append(np.bincount(b, minlength=n_bins))
        mass.append(np.bincount(b, weights=(weights[idx] if weights is not None else None),
                                minlength=n_bins))
        rr = np.maximum(r[idx], 1e-300)
        vr = (vv[idx] * d[idx]).sum(1) / rr
        s = np.bincount(b, weights=vr, minlength=n_bins)
        means.append(np.where(counts[-1] > 0, s / np.maximum(counts[-1], 1), 0.0))
    return np.array(counts), np.array(means), np.array(mass)


@pytest.mark.parametrize("periodic", [True, False])
def test_oracle_against_brute_force(tracers, periodic):
    m, om = oracle_mesh(tracers)
    voids = some_voids()
    ext, n_bins = 3.0, 12
    opt = 0 if periodic else PBC_OPEN
    d, u, cnt = oracle.profiles(om, voids, ext, n_bins, opt, dens=True, vel=True)
    bc, bmean, _ = brute(tracers, voids, ext, n_bins, periodic)
    diff = np.abs(cnt.astype(np.int64) - bc)
    assert diff.sum() <= 4 and cnt.sum() > 10000  # tracers on a bin edge in float
    assert np.all(cnt[7] == 0) and np.all(d[7] == 0.0)
    vmax = np.abs(tracers[3]).max()
    same = (diff == 0) & (bc > 0)
    assert np.abs(u[same] - bmean[same]).max() <= 1e-4 * vmax
    # the contrast from the counts: cumulative weight over the sphere's volume
    x = tracers[0]
    mean = len(x) / BOX ** 3
    ok = voids[:, 3] > 0
    edges = (np.arange(1, n_bins + 1) * (ext / n_bins))[None, :] * voids[ok, 3:4]
    want = np.cumsum(cnt[ok], 1) / (4.0 / 3.0 * np.pi * edges ** 3) / mean - 1.0
    np.testing.assert_allclose(d[ok], want, rtol=2e-6, atol=2e-6)


def test_oracle_cumulative_and_differential_agree(tracers):
    m, om = oracle_mesh(tracers)
    voids = some_voids(4)
    ext, n_bins = 2.5, 10
    dc, _, _ = oracle.profiles(om, voids, ext, n_bins, 0)
    dd, _, _ = oracle.profiles(om, voids, ext, n_bins, DIFFERENTIAL)
    k = np.arange(n_bins + 1) * (ext / n_bins)
    shell = k[1:] ** 3 - k[:-1] ** 3
    ok = voids[:, 3] > 0
    lhs = (dc[ok].astype(np.float64) + 1.0) * k[1:] ** 3
    rhs = np.cumsum((dd[ok].astype(np.float64) + 1.0) * shell, axis=1)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-5, atol=1e-5)


def _structs(tracers, velocities=True, n_voids=3):
    import sif_amd.pysif as pysif
    m, _ = oracle_mesh(tracers)
    fp = C.POINTER(C.c_float)
    s = pysif._Mesh()
    s.n_cells, s.total_cells = MESH, MESH ** 3
    s.box_length, s.cell_length = BOX, BOX / MESH
    s.x, s.y, s.z = (m[a].ctypes.data_as(fp) for a in "xyz")
    if velocities:
        s.vx, s.vy, s.vz = (m[a].ctypes.data_as(fp) for a in ("vx", "vy", "vz"))
    s.total_weight = float(len(m["x"]))
    s.n_particles = len(m["x"])
    s.cell_offsets = m["offsets"].ctypes.data_as(C.POINTER(C.c_uint64))
    v = some_voids()[:n_voids]
    cat = pysif.catalogue_from_numpy(v[:, 0], v[:, 1], v[:, 2], v[:, 3])
    return s, cat, m


def test_api_refusals(tracers):
    """profiles.c:97-142 and 366-373, before any device work."""
    from sif_amd import _native as N
    from sif_amd.pysif import measure
    s, cat, keep = _structs(tracers, velocities=False)
    d = C.POINTER(measure._Density)()
    v = C.POINTER(measure._Velocity)()
    L = measure.L

    def run(mesh, catalogue, n_bins=8, dens=True, vel=False):
        return L.sif_profiles(catalogue, mesh, 3.0, n_bins, 0,
                              C.byref(d) if dens else None, C.byref(v) if vel else None), \
            (N.lib.sif_error_message() or b"").decode()

    st, msg = run(C.byref(s), cat._p, n_bins=0)
    assert st == N.SIF_ERR_INVALID and "n_bins" in msg
    st, msg = run(C.byref(s), cat._p, dens=False)
    assert st == N.SIF_ERR_INVALID and "nothing to compute" in msg
    st, msg = run(C.byref(s), cat._p, vel=True)
    assert st == N.SIF_ERR_INVALID and "velocities" in msg
    st, msg = run(None, cat._p)
    assert st == N.SIF_ERR_INVALID and "mesh" in msg
    cat._p.contents.units = 1  # sky coordinates
    st, msg = run(C.byref(s), cat._p)
    assert st == N.SIF_ERR_INVALID and "sky" in msg
    assert not d and not v


def test_suggest_mesh_cells():
    from sif_amd.pysif import measure
    assert measure.profiles_suggest_mesh_cells(0) == 1
    assert measure.profiles_suggest_mesh_cells(36 * 50 ** 3) == 50
    assert measure.profiles_suggest_mesh_cells(10 ** 12) == 256
    assert measure.profiles_suggest_mesh_cells(100) == 1


# ---- GPU --------------------------------------------------------------------

@pytest.fixture(scope="module")
def dev_mesh(native, tracers):
    from sif_amd import device as D
    x, y, z, v, w = tracers
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    m = D.Mesh(dx, dy, dz, len(x), MESH, BOX)
    mw = D.Mesh(dx, dy, dz, len(x), MESH, BOX, dw=D.to_device(w))
    o = m.orig.get().astype(np.int64)
    vel = tuple(D.to_device(np.ascontiguousarray(v[k][o])) for k in range(3))
    return D, m, mw, vel


@pytest.mark.gpu
@pytest.mark.parametrize("opt,ext,n_bins", [(0, 3.0, 12), (DIFFERENTIAL, 2.0, 16),
                                             (PBC_OPEN, 3.0, 12), (0, 5.0, 40)])
def test_device_density_bit_exact(dev_mesh, tracers, opt, ext, n_bins):
    D, m, _, vel = dev_mesh
    _, om = oracle_mesh(tracers)
    voids = some_voids(5)
    print(f"[tracer profiles] opt {opt} ext {ext} bins {n_bins}", flush=True)
    got, gv = D.tracer_profiles(m, voids, ext, n_bins, opt, velocities=vel)
    want, wv, cnt = oracle.profiles(om, voids, ext, n_bins, opt, dens=True, vel=True)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.all(got[7] == 0.0) and cnt.sum() > 10000
    vmax = np.abs(tracers[3]).max()
    assert np.abs(gv - wv).max() <= 1e-4 * vmax


@pytest.mark.gpu
def test_device_weighted(dev_mesh, tracers):
    D, _, mw, vel = dev_mesh
    _, om = oracle_mesh(tracers, weighted=True)
    voids = some_voids(6)
    ext, n_bins = 3.0, 12
    for opt in (0, VEL_WEIGHTED):
        got, gv = D.tracer_profiles(mw, voids, ext, n_bins, opt, velocities=vel)
        want, wv, _ = oracle.profiles(om, voids, ext, n_bins, opt, dens=True, vel=True)
        # 1 + contrast is the enclosed weight over its expectation
        np.testing.assert_allclose(got + 1.0, want + 1.0, rtol=1e-5, atol=1e-6)
        vmax = np.abs(tracers[3]).max()
        assert np.abs(gv - wv).max() <= 1e-4 * vmax


@pytest.mark.gpu
def test_pysif_profiles(native, tracers):
    import sif_amd.pysif as pysif
    from sif_amd.pysif import measure
    x, y, z, v, w = tracers
    pysif.system_init()
    field = pysif.field_from_numpy(x, y, z, vx=v[0], vy=v[1], vz=v[2])
    mesh = pysif.ChainMesh(MESH, BOX, field)
    voids = some_voids(8)
    cat = pysif.catalogue_from_numpy(voids[:, 0], voids[:, 1], voids[:, 2], voids[:, 3])
    print("[tracer profiles] pysif", flush=True)
    p = measure.profiles(cat, mesh, 12, ext=3.0, compute_velocity=True)
    _, om = oracle_mesh(tracers)
    want, wv, _ = oracle.profiles(om, voids, 3.0, 12, 0, dens=True, vel=True)
    np.testing.assert_array_equal(p.density.view(np.uint32), want.view(np.uint32))
    assert np.abs(p.velocity - wv).max() <= 1e-4 * np.abs(v).max()
    assert p.has_density and p.has_velocity and p.n_voids == len(voids) and p.n_bins == 12
    np.testing.assert_allclose(p.r_edges, np.arange(13) * np.float32(0.25), rtol=1e-6)
    np.testing.assert_array_equal(p.r_bins, p.r_edges[1:])
    assert p.differential is False
    # a differential set, its bins at their middles
    q = measure.profiles(cat, mesh, 12, ext=3.0, differential=True)
    assert q.differential and not q.has_velocity
    np.testing.assert_allclose(q.r_bins, (q.r_edges[1:] + q.r_edges[:-1]) / 2)
    # weights carried from the field
    fw = pysif.field_from_numpy(x, y, z, weights=w)
    qw = measure.profiles(cat, pysif.ChainMesh(MESH, BOX, fw), 12, ext=3.0)
    _, omw = oracle_mesh(tracers, weighted=True)
    ww, _, _ = oracle.profiles(omw, voids, 3.0, 12, 0)
    np.testing.assert_allclose(qw.density + 1.0, ww + 1.0, rtol=1e-5, atol=1e-6)
