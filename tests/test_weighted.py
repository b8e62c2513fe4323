"""Weighted tracers on the exodus path: sif__exodus_find_radius_weighted
(exodus_rescale.c:1029-1037, find_exact_radius_impl with `weighted`, and
resolve_bin's pair walk, :391-507), the mean weight density of the density
tests (exodus_finder.c:1634-1642) and the weight checks
(exodus_finder.c:1039-1106).

Parity bar. The reference sums weights in double, in an order of its own
(tracers in template-cell order; windows of bins whose width adapts from one
candidate to the next on each thread). Where every partial sum is exact in
double the order cannot matter, and the device result is the oracle's bit
for bit: floats in [0.5, 2) over fewer than 2^28 tracers a sphere are such
weights (every partial sum a multiple of 2^-24 below 2^29), and so are
weights some of which are zero. Weights spread over many octaves round in
the last bit of a double, in the reference too, in an order its thread
schedule decides; there the device and the oracle are compared as the
catalogue they give on one seed.

CPU: the oracle against an FP64 brute force over every tracer; its
independence of the resolve window; all-ones weights reduce it to the
unweighted restatement; the C API's refusals of unusable weights.
GPU: the kernel's radii, reasons and counts bit for bit against the oracle;
the catalogue of a weighted run against the oracle's on the same smoothed
fields; all-ones weights give the unweighted catalogue; the pysif pipeline
on a weighted field; repeatability.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

N_SIDE, BOX, SEED = 32, 100.0, 11
MESH_CELLS = 10
THRESH = np.float32(-0.8)
PI_F = np.float32(3.14159265358979323846)


def vol_factor(n):
    return (np.float32(4) / np.float32(3)) * PI_F * (np.float32(n) / (np.float32(BOX) ** 3))


def mass_factor(total_weight):
    """exodus_finder.c:1639-1642, rounded as the reference rounds it."""
    return (np.float32(4) / np.float32(3)) * PI_F * (np.float32(total_weight) /
                                                     (np.float32(BOX) ** 3))


@pytest.fixture(scope="module")
def particles():
    from sif_amd import zeldovich
    return zeldovich.generate(N_SIDE, BOX, SEED)


def exact_weights(n, seed=5):
    """Floats in [0.5, 2), a fifth of them zero: double sums exact."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    w[rng.random(n) < 0.2] = 0.0
    return w


def spread_weights(n, seed=6):
    """Log-normal over about twelve octaves: double sums round."""
    return np.exp(np.random.default_rng(seed).normal(0.0, 1.5, n)).astype(np.float32)


def brute_force_weighted(x, y, z, w, centre, rung, rs, t, mass):
    """FP64 over every tracer: the outermost tracer distance d in (rung, rs]
    where the weight strictly inside d, over mass * d^3, is at most 1 + t."""
    p = np.c_[x, y, z].astype(np.float64)
    d = np.abs(p - np.asarray(centre, np.float64))
    d = np.minimum(d, BOX - d)
    d = np.sqrt((d * d).sum(1))
    order = np.argsort(d, kind="stable")
    d, ww = d[order], w[order].astype(np.float64)
    inside = np.concatenate([[0.0], np.cumsum(ww)[:-1]])
    ok = (d > rung) & (d <= rs) & (inside <= (1.0 + t) * float(mass) * d ** 3)
    return d[np.nonzero(ok)[0][-1]] if ok.any() else -1.0


def search_radius(rung, factor):
    return rung + max((factor - 1.0) * rung, 3.0 * BOX / N_SIDE)


# ---- CPU --------------------------------------------------------------------

def test_oracle_weighted_against_brute_force(particles):
    x, y, z = particles
    w = exact_weights(len(x))
    m = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX, w=w)
    om = oracle.OracleMesh(m, MESH_CELLS, BOX)
    mass = mass_factor(m["total_weight"])
    rung, rs = 6.0, search_radius(6.0, 1.5)
    centres = np.random.default_rng(2).uniform(0, BOX, (120, 3)).astype(np.float32)
    r, why, _ = oracle.find_radius(om, centres, rung, rs, THRESH, vol_factor(len(x)),
                                   mass_factor=mass)
    found = np.nonzero(r > 0)[0]
    assert len(found) >= 10
    for i in found:
        want = brute_force_weighted(x, y, z, w, centres[i], rung, rs, float(THRESH), mass)
        assert want > 0 and abs(r[i] - want) <= 1e-5 * want
    # a refusal below the rung means no tracer in the annulus passes
    for i in np.nonzero(why == 3)[0][:10]:
        assert brute_force_weighted(x, y, z, w, centres[i], rung, rs, float(THRESH),
                                    mass) < 0


def test_oracle_weighted_window_invariance(particles):
    """exodus_rescale.c:384-390: the radius does not depend on how many bins
    one resolution covers; weighted as unweighted."""
    x, y, z = particles
    m = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX, w=exact_weights(len(x)))
    om = oracle.OracleMesh(m, MESH_CELLS, BOX)
    mass = mass_factor(m["total_weight"])
    centres = np.random.default_rng(9).uniform(0, BOX, (80, 3)).astype(np.float32)
    base = oracle.find_radius(om, centres, 6.0, 9.375, THRESH, vol_factor(len(x)), window=1,
                              mass_factor=mass)[0]
    assert (base > 0).sum() >= 5
    for win in (2, 8, 64):
        got = oracle.find_radius(om, centres, 6.0, 9.375, THRESH, vol_factor(len(x)),
                                 window=win, mass_factor=mass)[0]
        np.testing.assert_array_equal(got, base)


def test_oracle_unit_weights_are_the_unweighted_run(particles):
    """exodus_finder.c:1634-1638: all-ones weights reproduce the unweighted
    rescaling bit for bit."""
    x, y, z = particles
    m0 = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX)
    m1 = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX, w=np.ones(len(x), np.float32))
    assert m1["total_weight"] == len(x) and mass_factor(len(x)) == vol_factor(len(x))
    centres = np.random.default_rng(4).uniform(0, BOX, (120, 3)).astype(np.float32)
    for rung, factor in ((6.0, 1.5), (9.0, 2.0)):
        rs = search_radius(rung, factor)
        a = oracle.find_radius(oracle.OracleMesh(m0, MESH_CELLS, BOX), centres, rung, rs,
                               THRESH, vol_factor(len(x)))
        b = oracle.find_radius(oracle.OracleMesh(m1, MESH_CELLS, BOX), centres, rung, rs,
                               THRESH, vol_factor(len(x)), mass_factor=vol_factor(len(x)))
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)


def _hand_built(x, y, z, w, cell_table=True):
    """A weighted sif_chain_mesh_t and a small sif_grid_t, filled from the
    oracle's mesh in host memory: what the C API checks before any device
    work."""
    import sif_amd.pysif as pysif
    m = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX, w=w)
    fp = C.POINTER(C.c_float)
    s = pysif._Mesh()
    s.n_cells, s.total_cells = MESH_CELLS, MESH_CELLS ** 3
    s.box_length, s.cell_length = BOX, BOX / MESH_CELLS
    s.x, s.y, s.z = (m[a].ctypes.data_as(fp) for a in "xyz")
    s.weights = m["w"].ctypes.data_as(fp)
    s.total_weight = m["total_weight"]
    s.cell_weights = m["cell_weights"].ctypes.data_as(fp) if cell_table else None
    s.n_particles = len(x)
    s.cell_offsets = m["offsets"].ctypes.data_as(C.POINTER(C.c_uint64))
    values = np.zeros(8 ** 3, np.float32)
    g = pysif._Grid()
    g.values = values.ctypes.data_as(fp)
    g.n_cells, g.total_cells, g.box_length, g.cell_length = 8, 8 ** 3, BOX, BOX / 8
    return s, g, (m, values)


def test_api_refuses_unusable_weights(particles):
    """exodus_finder.c:1039-1106, before any device work: negative or
    non-finite weights (counted), no cell table, a zero total. -0 is a zero,
    and passes (the run then stops at an option this build does not take)."""
    from sif_amd import _native as N
    from sif_amd.pysif import finders
    x, y, z = (a[:4000] for a in particles)
    radii = np.array([10.0, 8.0], np.float32)
    rp = radii.ctypes.data_as(C.POINTER(C.c_float))

    def run(w, cell_table=True, opt=finders.NO_CHECKPOINT):
        s, g, keep = _hand_built(x, y, z, w, cell_table)
        ptr = finders.L.sif_finder_exodus_raw(C.byref(g), C.byref(s), rp, 2, -0.8, 0.0, opt,
                                              None)
        assert not ptr
        return N.lib.sif_error_status(), (N.lib.sif_error_message() or b"").decode()

    base = np.ones(len(x), np.float32)
    w = base.copy()
    w[[3, 7]] = -1.0
    w[11] = np.inf
    w[12] = np.nan
    st, msg = run(w)
    assert st == N.SIF_ERR_INVALID and "2 negative, 2 not finite" in msg
    st, msg = run(base, cell_table=False)
    assert st == N.SIF_ERR_INVALID and "per-cell weight table" in msg
    st, msg = run(np.zeros(len(x), np.float32))
    assert st == N.SIF_ERR_INVALID and "sum to zero" in msg
    w = base.copy()
    w[5] = -0.0
    st, msg = run(w, opt=finders.NO_CHECKPOINT | finders.REFINE["gaussian"])
    assert st == N.SIF_ERR_UNSUPPORTED and "REFINE_GAUSSIAN" in msg


# ---- GPU --------------------------------------------------------------------

@pytest.fixture(scope="module")
def dev(native, particles):
    from sif_amd import device as D
    x, y, z = particles
    return D, D.to_device(x), D.to_device(y), D.to_device(z)


def _weighted_pair(dev, particles, w):
    D, dx, dy, dz = dev
    x, y, z = particles
    m = D.Mesh(dx, dy, dz, len(x), MESH_CELLS, BOX, dw=D.to_device(w))
    om_arrays = oracle.chain_mesh(x, y, z, MESH_CELLS, BOX, w=w)
    assert m.total_weight == om_arrays["total_weight"]
    return m, oracle.OracleMesh(om_arrays, MESH_CELLS, BOX)


@pytest.mark.gpu
@pytest.mark.parametrize("rung,factor,t", [(6.0, 1.5, -0.8), (9.0, 2.0, -0.7),
                                            (4.0, 1.25, -0.4)])
def test_weighted_kernel_bit_exact(dev, particles, rung, factor, t):
    D = dev[0]
    x, y, z = particles
    w = exact_weights(len(x))
    m, om = _weighted_pair(dev, particles, w)
    mass = mass_factor(m.total_weight)
    rs = search_radius(rung, factor)
    centres = np.random.default_rng(int(rung * 10)).uniform(0, BOX, (400, 3)).astype(np.float32)
    print(f"[weighted kernel] rung {rung} search {rs} t {t}", flush=True)
    gr, gre, gn = D.find_radius(m, centres, rung, rs, np.float32(t), vol_factor(len(x)),
                                mass_factor=mass)
    orr, ore, on = oracle.find_radius(om, centres, rung, rs, np.float32(t),
                                      vol_factor(len(x)), mass_factor=mass)
    np.testing.assert_array_equal(gre, ore)
    np.testing.assert_array_equal(gr.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(gn[gr > 0], on[orr > 0])
    found = np.nonzero(gr > 0)[0]
    assert len(found) >= 10
    for i in found[:25]:
        want = brute_force_weighted(x, y, z, w, centres[i], rung, rs, t, mass)
        assert want > 0 and abs(gr[i] - want) <= 1e-5 * want


def _fields_and_grid(dev, particles, w, n=32):
    D, dx, dy, dz = dev
    x, y, z = particles
    g = D.cic(dx, dy, dz, len(x), n, BOX, dw=D.to_device(w) if w is not None else None)
    D.density_contrast(g, n ** 3)
    return g


RADII = np.array([14.0, 11.0, 9.0, 7.5, 6.25], np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["exact", "spread"])
def test_weighted_catalogue_matches_oracle(dev, particles, kind):
    D = dev[0]
    x, y, z = particles
    w = exact_weights(len(x)) if kind == "exact" else spread_weights(len(x))
    n = 32
    g = _fields_and_grid(dev, particles, w, n)
    fields = D.exodus_fields(g, n, BOX, RADII, 0)
    m, om = _weighted_pair(dev, particles, w)
    print(f"[weighted exodus] {kind} weights", flush=True)
    cat, nin, stats = D.exodus(g, n, BOX, m, RADII, THRESH)
    ocat, onin, orung, _ = oracle.exodus_run(list(fields), n, BOX, om, len(x), RADII, THRESH,
                                             tie_rule=1, witness=False)
    assert len(cat) > 5
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)
    assert sum(s["accepted"] for s in stats) == len(cat)
    again, _, _ = D.exodus(g, n, BOX, m, RADII, THRESH)
    np.testing.assert_array_equal(again.view(np.uint32), cat.view(np.uint32))
    if kind == "exact":
        mass = mass_factor(m.total_weight)
        for (cx, cy, cz, r), k in zip(cat[:15], orung[:15]):
            rung = float(RADII[k])
            rs = search_radius(rung, 1.5)
            if k > 0:
                rs = max(rs, float(RADII[k - 1]))
            want = brute_force_weighted(x, y, z, w, (cx, cy, cz), rung, rs, float(THRESH), mass)
            assert abs(r - want) <= 1e-5 * want


@pytest.mark.gpu
def test_unit_weights_give_the_unweighted_catalogue(dev, particles):
    D, dx, dy, dz = dev
    x, y, z = particles
    n = 32
    g = _fields_and_grid(dev, particles, None, n)
    m0 = D.Mesh(dx, dy, dz, len(x), MESH_CELLS, BOX, indices=False)
    m1 = D.Mesh(dx, dy, dz, len(x), MESH_CELLS, BOX, dw=D.to_device(np.ones(len(x), np.float32)),
                indices=False)
    print("[weighted exodus] unit weights", flush=True)
    a, na, _ = D.exodus(g, n, BOX, m0, RADII, THRESH)
    b, nb, _ = D.exodus(g, n, BOX, m1, RADII, THRESH)
    assert len(a) > 5
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(na, nb)


@pytest.mark.gpu
def test_pysif_weighted_pipeline(dev, particles):
    """Weights carried from the field through the CIC grid and the mesh
    (field.h, grid.h, chain_mesh.h) into the finder."""
    import sif_amd.pysif as pysif
    from sif_amd.pysif import finders
    D = dev[0]
    x, y, z = particles
    w = exact_weights(len(x))
    pysif.system_init()
    field = pysif.field_from_numpy(x, y, z, weights=w)
    grid = pysif.Grid(32, BOX)
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(MESH_CELLS, BOX, field, drop_indices=True)
    print("[weighted exodus] pysif", flush=True)
    got = finders.exodus_raw(grid, mesh, RADII, THRESH, checkpoint=False)
    got = np.c_[got.centres, got.radii]
    g = _fields_and_grid(dev, particles, w, 32)
    m, _ = _weighted_pair(dev, particles, w)
    want, _, _ = D.exodus(g, 32, BOX, m, RADII, THRESH, restore=False)
    assert len(want) > 5
    np.testing.assert_array_equal(got, want)
