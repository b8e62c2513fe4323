"""GPU parity of every stage of the exodus path against the oracle, through
the C ABI, on seeded Zel'dovich particle sets small enough for the oracle.

Bars: integer and index work bit-exact (mesh, candidate order, counts);
the fixed-point CIC bit-exact against its integer model and within 1e-5 of
the peak against the textbook double deposit and the reference float order
(tests/test_field_grid.c:812); the FP32 smoothed field against an FP64
transform within the tolerance stated in the test; radii bit-exact against
the oracle's restatement on the same centres, and within 1e-5 relative of
the brute-force radius of tests/test_finders.c:433-465.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SIDE, BOX, SEED = 32, 100.0, 11
MESH_CELLS = 10
THRESH = -0.8


@pytest.fixture(scope="module")
def particles():
    from sif_amd import zeldovich
    return zeldovich.generate(N_SIDE, BOX, SEED)


@pytest.fixture(scope="module")
def dev(native, particles):
    from sif_amd import device as D
    x, y, z = particles
    return D, D.to_device(x), D.to_device(y), D.to_device(z)


def test_check_in_box(dev, particles):
    D, dx, dy, dz = dev
    import oracle
    x, y, z = particles
    assert D.check_in_box(dx, dy, dz, len(x), BOX)[:2] == oracle.check_in_box(x, y, z, BOX)[:2]
    bad = x.copy()
    bad[7] = BOX * 1.01
    bad[9] = -1.0
    bad[11] = np.float32(BOX)  # on the face: accepted within tolerance
    got = D.check_in_box(D.to_device(bad), dy, dz, len(x), BOX)
    assert got == (2, 1, 7)
    assert got == oracle.check_in_box(bad, y, z, BOX)


@pytest.mark.parametrize("n_cells", [16, 32, 48])
def test_cic_matches_fixed_point_model_bit_exact(dev, particles, n_cells):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    got = D.cic(dx, dy, dz, len(x), n_cells, BOX).get()
    want, _, _ = oracle.cic_fixed(x, y, z, n_cells, BOX)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_cic_against_textbook_and_reference_order(dev, particles):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    n_cells = 32
    got = D.cic(dx, dy, dz, len(x), n_cells, BOX).get().astype(np.float64)
    brute = oracle.cic_brute(x, y, z, n_cells, BOX)
    ref = oracle.cic_float(x, y, z, n_cells, BOX).astype(np.float64)
    peak = brute.max()
    assert np.abs(got - brute).max() <= 1e-5 * peak
    assert np.abs(got - ref).max() <= 1e-5 * peak
    # mass: every particle deposits 1 (test_field_grid.c:260-309)
    cell = BOX / n_cells
    assert abs(got.sum() * cell ** 3 - len(x)) <= 1e-4 * len(x)


def test_cic_weighted_and_repeatable(dev, particles):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    w = np.random.default_rng(3).uniform(0.25, 4.0, len(x)).astype(np.float32)
    dw = D.to_device(w)
    a = D.cic(dx, dy, dz, len(x), 32, BOX, dw=dw).get()
    b = D.cic(dx, dy, dz, len(x), 32, BOX, dw=dw).get()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    want, _, _ = oracle.cic_fixed(x, y, z, 32, BOX, w=w)
    np.testing.assert_array_equal(a.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("scale", [1e-6, 1e9, 1e10])
def test_cic_weights_far_from_one(dev, particles, scale):
    """Weights of 1e-6 and of 1e9 to 1e10 (halo masses): the unit of the fixed
    point follows them, so nothing overflows and small weights keep their
    digits; bit-exact against the oracle's model, within 1e-5 of the peak
    against the textbook double deposit."""
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    w = (np.random.default_rng(8).uniform(0.1, 10.0, len(x)) * scale).astype(np.float32)
    got = D.cic(dx, dy, dz, len(x), 32, BOX, dw=D.to_device(w)).get()
    want, _, _ = oracle.cic_fixed(x, y, z, 32, BOX, w=w)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    brute = oracle.cic_brute(x, y, z, 32, BOX, w=w)
    assert np.abs(got - brute).max() <= 1e-5 * np.abs(brute).max()


def test_cic_refuses_a_weight_that_is_not_finite(dev, particles):
    D, dx, dy, dz = dev
    x = particles[0]
    w = np.ones(len(x), np.float32)
    w[17] = np.inf
    with pytest.raises(ValueError, match="not finite"):
        D.cic(dx, dy, dz, len(x), 32, BOX, dw=D.to_device(w))


def test_density_contrast(dev, particles):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    g = D.cic(dx, dy, dz, len(x), 32, BOX)
    host = g.get()
    mean = D.density_contrast(g, 32 ** 3)
    assert abs(mean - host.astype(np.float64).mean()) <= 1e-12 * mean
    want = oracle.density_contrast_with_mean(host, mean)
    np.testing.assert_array_equal(g.get().view(np.uint32), want.view(np.uint32))
    ref, ref_mean = oracle.density_contrast(host)
    assert abs(ref_mean - mean) <= 1e-12 * mean


@pytest.mark.parametrize("mesh_cells", [1, 2, 3, 7, 10, 16])
def test_chain_mesh_bit_exact(dev, particles, mesh_cells):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    m = D.Mesh(dx, dy, dz, len(x), mesh_cells, BOX)
    want = oracle.chain_mesh(x, y, z, mesh_cells, BOX)
    np.testing.assert_array_equal(m.offsets.get(), want["offsets"])
    np.testing.assert_array_equal(m.orig.get(), want["orig"])
    for a in "xyz":
        np.testing.assert_array_equal(getattr(m, a).get(), want[a])
    assert m.total_weight == len(x)


@pytest.mark.parametrize("mesh_cells", [3, 10])
def test_chain_mesh_weighted(dev, particles, mesh_cells):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    w = (np.random.default_rng(5).integers(1, 8, len(x)) * 0.25).astype(np.float32)
    m = D.Mesh(dx, dy, dz, len(x), mesh_cells, BOX, dw=D.to_device(w), indices=True)
    want = oracle.chain_mesh(x, y, z, mesh_cells, BOX, w=w)
    np.testing.assert_array_equal(m.orig.get(), want["orig"])
    np.testing.assert_array_equal(m.x.get(), want["x"])
    np.testing.assert_array_equal(m.w.get(), want["w"])
    np.testing.assert_array_equal(m.cell_weights.get(), want["cell_weights"])
    assert m.total_weight == want["total_weight"]


def test_chain_mesh_edge_particles(native):
    """Coordinates within rounding of the faces are folded (field_internal.h)."""
    import oracle
    from sif_amd import device as D
    L = np.float32(BOX)
    x = np.array([0.0, L, np.nextafter(L, 0), 50.0, 99.99, -1e-6], np.float32)
    y = np.array([L, 0.0, 3.0, 50.0, 0.5, 2.0], np.float32)
    z = np.array([1.0, 2.0, L, 50.0, 0.0, 0.0], np.float32)
    m = D.Mesh(D.to_device(x), D.to_device(y), D.to_device(z), len(x), 4, BOX)
    want = oracle.chain_mesh(x, y, z, 4, BOX)
    np.testing.assert_array_equal(m.offsets.get(), want["offsets"])
    np.testing.assert_array_equal(m.orig.get(), want["orig"])
    np.testing.assert_array_equal(m.x.get(), want["x"])


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
def test_fft_round_trip(native, n):
    """test_fft.c:49-125: r2c then c2r returns the field to 1e-4."""
    from sif_amd import device as D
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n ** 3).astype(np.float32)
    dv = D.to_device(v)
    spec = D.spectrum(dv, n)
    p = D.spectrum_pitch(n)
    assert p % 16 == 0 and p >= n // 2 + 1  # rows padded to lines
    s = D.spectrum_modes(spec, n)
    import scipy.fft
    want = scipy.fft.rfftn(v.reshape(n, n, n).astype(np.float64), workers=-1)
    got = s[..., 0] + 1j * s[..., 1]
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    del want, got, s
    back = D.smooth(spec, n, BOX, 0, 0.0, window_order=0).get()
    assert np.abs(back - v).max() < 1e-4
    # a single mode: the spectrum has one nonzero entry at the right index,
    # of magnitude n^3 / 2 (pins the twiddle sign and the axis order)
    k = (1, 2, 3)
    i = np.arange(n, dtype=np.float64)
    phase = (k[0] * i)[:, None, None] + (k[1] * i)[None, :, None] + (k[2] * i)[None, None, :]
    wave = np.cos(2 * np.pi * phase / n).astype(np.float32).ravel()
    del phase
    s1 = D.spectrum_modes(D.spectrum(D.to_device(wave), n), n)
    mag = np.hypot(s1[..., 0], s1[..., 1])
    assert abs(mag[k] - n ** 3 / 2) <= 1e-4 * n ** 3
    mag[k] = 0.0
    assert mag.max() <= 1e-4 * n ** 3


@pytest.mark.parametrize("n", [15, 24, 40, 72, 100, 136, 200, 248, 384])
def test_fft_any_length(native, n):
    """Lengths that are not powers of two (the mixed-radix kernels: radices 8,
    4, 2 and odd primes up to 31; n odd through whole rows): against scipy's
    FP64 transform, the round trip of test_fft.c:49-125 (which round-trips
    n = 15) and a single mode in one bin."""
    import scipy.fft
    from sif_amd import device as D
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n ** 3).astype(np.float32)
    spec = D.spectrum(D.to_device(v), n)
    s = D.spectrum_modes(spec, n)
    want = scipy.fft.rfftn(v.reshape(n, n, n).astype(np.float64), workers=-1)
    got = s[..., 0] + 1j * s[..., 1]
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    back = D.smooth(spec, n, BOX, 0, 0.0, window_order=0).get()
    assert np.abs(back - v).max() < 1e-4
    k = (1, 2, 3)
    i = np.arange(n, dtype=np.float64)
    phase = (k[0] * i)[:, None, None] + (k[1] * i)[None, :, None] + (k[2] * i)[None, None, :]
    wave = np.cos(2 * np.pi * phase / n).astype(np.float32).ravel()
    s1 = D.spectrum_modes(D.spectrum(D.to_device(wave), n), n)
    mag = np.hypot(s1[..., 0], s1[..., 1])
    assert abs(mag[k] - n ** 3 / 2) <= 1e-4 * n ** 3
    mag[k] = 0.0
    assert mag.max() <= 1e-4 * n ** 3


@pytest.mark.parametrize("n,radius", [(72, 6.0), (100, 9.0), (136, 12.5)])
def test_smoothing_any_length_against_fp64(native, n, radius):
    """The filtered c2r (top hat, CIC window divided out) of a field on a
    grid that is not a power of two, against the same factors on an FP64
    transform; 1e-5 of the field's range."""
    import oracle
    from sif_amd import device as D
    v = np.random.default_rng(n + 1).standard_normal(n ** 3).astype(np.float32)
    spec = D.spectrum(D.to_device(v), n)
    got = D.smooth(spec, n, BOX, 1, radius, window_order=2).get()
    fac = oracle.mode_factors(oracle.filter_lut(1, radius, 0.0, BOX, n),
                              oracle.window_axis(n, 2, 0), n).astype(np.float64)
    want = np.fft.irfftn(np.fft.rfftn(v.reshape(n, n, n).astype(np.float64)) * fac,
                         s=(n, n, n), axes=(0, 1, 2)).reshape(-1)
    assert np.abs(got - want).max() <= 1e-5 * (want.max() - want.min())


def test_exodus_on_a_grid_that_is_not_a_power_of_two(native):
    """40^3 Zel'dovich particles on a 40^3 grid (the grid matching the
    lattice) in a 62.5 Mpc/h box: the catalogue equals the oracle's on the
    same smoothed fields."""
    import oracle
    from sif_amd import device as D
    from sif_amd import zeldovich
    n, box, mc = 40, 62.5, 6
    x, y, z = zeldovich.generate(40, box, 31)
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    g = D.cic(dx, dy, dz, len(x), n, box)
    D.density_contrast(g, n ** 3)
    radii = zeldovich.ladder(2.0 * box / n, 10.0, 0.05)
    fields = D.exodus_fields(g, n, box, radii)
    m = D.Mesh(dx, dy, dz, len(x), mc, box)
    cat, nin, _ = D.exodus(g, n, box, m, radii, THRESH)
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mc, box), mc, box)
    ocat, onin, _, _ = oracle.exodus_run(list(fields), n, box, om, len(x), radii, THRESH,
                                        tie_rule=1)
    assert len(cat) > 10
    np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
    np.testing.assert_array_equal(nin, onin)


def test_fft_1024_properties(native):
    """n = 1024, where the column passes hold 73.8 KB of LDS a block: the
    round trip to 1e-4 (test_fft.c:49-125), Parseval's identity over the
    half spectrum, and a single mode in exactly one bin of magnitude n^3 / 2.
    (The transform's values against scipy are pinned up to n = 512 above.)"""
    from sif_amd import device as D
    n = 1024
    print("[fft] 1024^3 round trip", flush=True)
    v = np.random.default_rng(n).standard_normal(n ** 3, dtype=np.float32)
    dv = D.to_device(v)
    spec = D.spectrum(dv, n)
    s = D.spectrum_modes(spec, n)
    # Parseval: sum |v|^2 = (1/n^3) sum over the full spectrum |F|^2; the half
    # spectrum counts every kz but 0 and n/2 twice
    power = 0.0
    for ix in range(0, n, 64):
        blk = s[ix:ix + 64].astype(np.float64)
        m2 = blk[..., 0] ** 2 + blk[..., 1] ** 2
        power += 2.0 * m2.sum() - m2[..., 0].sum() - m2[..., n // 2].sum()
    energy = float(np.dot(v.astype(np.float64), v.astype(np.float64)))
    assert abs(power / n ** 3 - energy) <= 1e-5 * energy
    del s
    back = D.smooth(spec, n, BOX, 0, 0.0, window_order=0).get()
    assert np.abs(back - v).max() < 1e-4
    del back, dv, spec
    k = (1, 2, 3)
    i = np.arange(n, dtype=np.int32)
    phase = (k[0] * i[:, None, None] + k[1] * i[None, :, None] + k[2] * i[None, None, :]) % n
    table = np.cos(2 * np.pi * np.arange(n) / n).astype(np.float32)
    wave = table[phase].ravel()
    del phase
    s1 = D.spectrum_modes(D.spectrum(D.to_device(wave), n), n)
    del wave
    mag = np.hypot(s1[..., 0], s1[..., 1])
    del s1
    assert abs(mag[k] - n ** 3 / 2) <= 1e-4 * n ** 3
    mag[k] = 0.0
    assert mag.max() <= 1e-4 * n ** 3


def test_top_hat_preserves_constant(native):
    """test_fft.c:129-173."""
    from sif_amd import device as D
    n = 32
    v = np.full(n ** 3, 3.25, np.float32)
    spec = D.spectrum(D.to_device(v), n)
    out = D.smooth(spec, n, BOX, 1, 7.0).get()
    assert np.abs(out - 3.25).max() < 1e-4 * 3.25


@pytest.mark.parametrize("radius,window", [(6.0, 2), (12.5, 2), (6.0, 0)])
def test_smoothing_against_fp64(dev, particles, radius, window):
    """The FP32 smoothed field against the same filter applied to an FP64
    transform of the same contrast; the bar (1e-5 of the field's range) is the
    FP32 transform's rounding at these sizes."""
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    n = 32
    g = D.cic(dx, dy, dz, len(x), n, BOX)
    D.density_contrast(g, n ** 3)
    delta = g.get()
    spec = D.spectrum(g, n)
    got = D.smooth(spec, n, BOX, 1, radius, window_order=window).get()
    lut = oracle.filter_lut(1, radius, 0.0, BOX, n)
    axis = oracle.window_axis(n, 2, 0) if window else None
    fac = oracle.mode_factors(lut, axis, n).astype(np.float64)
    want = np.fft.irfftn(np.fft.rfftn(delta.reshape(n, n, n).astype(np.float64)) * fac,
                         s=(n, n, n)).reshape(-1)
    span = want.max() - want.min()
    err = np.abs(got - want)
    assert err.max() <= 1e-5 * span, (err.max(), span)


def test_scan_candidates_match_oracle(dev, particles):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    n = 32
    g = D.cic(dx, dy, dz, len(x), n, BOX)
    D.density_contrast(g, n ** 3)
    spec = D.spectrum(g, n)
    f = D.smooth(spec, n, BOX, 1, 8.0)
    field = f.get()
    mask = np.zeros((n ** 3 + 63) // 64, np.uint64)
    mask[3] = 0xFFFF0000FFFF0000
    dmask = D.to_device(mask, np.uint64)
    for t in (-0.2, -0.5):
        gd, gf = D.scan_candidates(f, n ** 3, t, dmask=dmask)
        od, of = oracle.scan_candidates(field, t, mask=mask, tie_rule=1)
        np.testing.assert_array_equal(gf, of.astype(np.uint32))
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
        # no exact ties on this field: the reference quicksort gives the same
        # order
        rd, rf = oracle.scan_candidates(field, t, mask=mask, tie_rule=0)
        if len(np.unique(od)) == len(od):
            np.testing.assert_array_equal(rf, of)


def _mesh_pair(dev, particles, mesh_cells):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    m = D.Mesh(dx, dy, dz, len(x), mesh_cells, BOX)
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mesh_cells, BOX), mesh_cells, BOX)
    return m, om


@pytest.mark.parametrize("rung,factor", [(6.0, 1.5), (9.0, 2.0), (4.0, 1.25)])
def test_find_radius_bit_exact_and_brute_force(dev, particles, rung, factor):
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    m, om = _mesh_pair(dev, particles, MESH_CELLS)
    rng = np.random.default_rng(int(rung * 10))
    centres = rng.uniform(0, BOX, (300, 3)).astype(np.float32)
    cell = BOX / 32
    r_search = rung + max((factor - 1.0) * rung, 3.0 * cell)
    vol = (np.float32(4.0) / np.float32(3.0)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) * np.float32(BOX) * np.float32(BOX)))
    t = np.float32(THRESH)
    gr, gre, gn = D.find_radius(m, centres, rung, r_search, t, vol)
    orr, ore, on = oracle.find_radius(om, centres, rung, r_search, t, vol)
    np.testing.assert_array_equal(gre, ore)
    np.testing.assert_array_equal(gr.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(gn[gr > 0], on[orr > 0])
    found = np.nonzero(gr > 0)[0]
    assert len(found) > 0
    for i in found[:40]:
        want = oracle.brute_force_radius(x, y, z, BOX, centres[i], rung, r_search, THRESH)
        assert want > 0 and abs(gr[i] - want) <= 1e-5 * want


def test_find_radius_split_first_pass(dev, particles):
    """A launch of fewer candidates than the chip holds waves shares each
    candidate's first pass among several waves; a large launch does not.
    The same centres through both give the same bits, and the oracle's."""
    import oracle
    D = dev[0]
    x, y, z = particles
    m, om = _mesh_pair(dev, particles, MESH_CELLS)
    centres = np.random.default_rng(23).uniform(0, BOX, (2400, 3)).astype(np.float32)
    vol = (np.float32(4.0) / np.float32(3.0)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    for rung, rs in ((6.0, 9.375), (9.0, 18.0)):
        whole = D.find_radius(m, centres, rung, rs, THRESH, vol)        # one wave each
        split = D.find_radius(m, centres[:150], rung, rs, THRESH, vol)  # 16 waves each
        for a, b in zip(whole, split):
            np.testing.assert_array_equal(a[:150], b)
        orr, ore, _ = oracle.find_radius(om, centres[:150], rung, rs, THRESH, vol)
        np.testing.assert_array_equal(split[0].view(np.uint32), orr.view(np.uint32))
        np.testing.assert_array_equal(split[1], ore)


def test_find_radius_window_invariance(dev, particles):
    """The oracle's radius does not depend on the resolve window (the
    reference's claim, exodus_rescale.c:384-390), so the GPU's choice of
    window is free."""
    import oracle
    _, om = _mesh_pair(dev, particles, MESH_CELLS)
    centres = np.random.default_rng(9).uniform(0, BOX, (100, 3)).astype(np.float32)
    vol = (np.float32(4.0) / np.float32(3.0)) * np.float32(3.14159265358979323846) * \
        (np.float32(N_SIDE ** 3) / (np.float32(BOX) ** 3))
    base = oracle.find_radius(om, centres, 6.0, 9.375, THRESH, vol, window=1)[0]
    for w in (2, 8, 64):
        np.testing.assert_array_equal(
            oracle.find_radius(om, centres, 6.0, 9.375, THRESH, vol, window=w)[0], base)


def _exodus_setup(dev, particles, n=32):
    D, dx, dy, dz = dev
    x, y, z = particles
    g = D.cic(dx, dy, dz, len(x), n, BOX)
    D.density_contrast(g, n ** 3)
    return g


# growth batch windows (floor, ceiling) small enough that a 32^3 rung spans
# many batches: the pipelined slots, the pair effects between a batch's
# candidates, the re-tests at batch assembly and the rung-start index are
# all crossed
SMALL_WINDOWS = [(16, 16), (64, 256), (256, 1024)]


@pytest.mark.parametrize("overlap,opts", [(0.0, 0), (0.0, 1 << 18), (0.5, 0), (0.0, 1 << 9),
                                          (0.5, 1 << 9), (1.0, (1 << 9) | (1 << 21))])
def test_exodus_catalogue_matches_oracle_on_same_fields(dev, particles, overlap, opts):
    """Cells, Lagrangian overlap, a partial overlap, the minima seeds
    (SIF_FINDER_SEEDS_MINIMA, bit 9) and the extended reach
    (SIF_FINDER_EXTEND_REACH, bit 21); with the default batch window (one
    batch a rung at this size) and with small ones (many)."""
    import oracle
    D, dx, dy, dz = dev
    x, y, z = particles
    n = 32
    radii = np.array([14.0, 11.0, 9.0, 7.5, 6.25], np.float32)
    g = _exodus_setup(dev, particles, n)
    delta0 = g.get()
    fields = D.exodus_fields(g, n, BOX, radii, opts)
    m, om = _mesh_pair(dev, particles, MESH_CELLS)
    cat, nin, stats = D.exodus(g, n, BOX, m, radii, THRESH, overlap=overlap, options=opts)
    # the grid comes back restored to float precision
    assert np.abs(g.get() - delta0).max() < 1e-4 * max(1.0, np.abs(delta0).max())
    lag = bool(opts & (1 << 18))
    for witness in (True, False):
        ocat, onin, orung, ostats = oracle.exodus_run(
            list(fields), n, BOX, om, len(x), radii, THRESH, overlap=overlap,
            lagrangian=lag, search_opt=opts & ((1 << 9) | (1 << 21)), tie_rule=1,
            witness=witness)
        np.testing.assert_array_equal(cat.view(np.uint32), ocat.view(np.uint32))
        np.testing.assert_array_equal(nin, onin)
    assert len(cat) > 0
    assert sum(s["accepted"] for s in stats) == len(cat)
    try:
        for floor, ceiling in SMALL_WINDOWS:
            D.set_batch_window(floor, ceiling)
            g2 = _exodus_setup(dev, particles, n)  # the same bits as g before the run
            cat2, nin2, stats2 = D.exodus(g2, n, BOX, m, radii, THRESH, overlap=overlap,
                                          options=opts)
            np.testing.assert_array_equal(cat2.view(np.uint32), ocat.view(np.uint32))
            np.testing.assert_array_equal(nin2, onin)
            # the reason a candidate is refused for can depend on the batching
            # (as in the reference, whose phase 1 tests against the catalogue
            # frozen at the batch's start); what is refused cannot
            refused = ("masked", "proxy", "overlap", "rescale_fail")
            for a, b in zip(stats, stats2):
                assert a["candidates"] == b["candidates"] and a["accepted"] == b["accepted"]
                assert sum(a[k] for k in refused) == sum(b[k] for k in refused), (floor, a, b)
            if ceiling == 16:  # a batch holds at most 16: such a rung spans several
                assert max(s["evaluated"] for s in stats2) > ceiling
    finally:
        D.set_batch_window(0, 0)
    # radii against brute force at the rung they came from
    for (cx, cy, cz, r), k in zip(cat[:20], orung[:20]):
        rung = float(radii[k])
        rs = rung + max(0.5 * rung, 3.0 * BOX / n)
        if k > 0:
            rs = max(rs, float(radii[k - 1]))
        if opts & (1 << 21):  # the extended reach: up to the top rung's
            rs = max(rs, float(radii[0]) + max(0.5 * float(radii[0]), 3.0 * BOX / n))
        want = oracle.brute_force_radius(x, y, z, BOX, (cx, cy, cz), rung, rs, THRESH)
        assert abs(r - want) <= 1e-5 * want


@pytest.mark.parametrize("n,box,nodes", [(64, 100.0, True), (100, 333.3, False)])
def test_device_painting_matches_reference_bits(native, n, box, nodes):
    """sif__mark_sphere (utils.c:965-1033) on the device, bit for bit
    against the oracle's node-by-node restatement."""
    import ctypes as C
    import oracle
    from sif_amd import device as D
    from sif_amd._native import DeviceArray
    rng = np.random.default_rng(n)
    l = np.float32(box) / np.float32(n)
    k = 200
    cen = (rng.integers(0, n, (k, 3)).astype(np.float32) * l if nodes
           else rng.uniform(0, box, (k, 3)).astype(np.float32))
    r = np.concatenate([rng.uniform(0.3, 3, k // 2) * l,
                        rng.uniform(3, 20, k - k // 2) * l]).astype(np.float32)
    scale = np.float32(np.cbrt(np.float32(0.2)))
    words = (n ** 3 + 63) // 64
    dm = DeviceArray((words,), np.uint64)
    native.check(native.lib.sif_amd_memset(dm.ptr, 0, words * 8))
    dx, dy, dz, dr = (D.to_device(np.ascontiguousarray(a)) for a in
                      (cen[:, 0], cen[:, 1], cen[:, 2], r))
    native.check(native.lib.sif_amd_mark_spheres(dm.ptr, n, box, dx.ptr, dy.ptr, dz.ptr, dr.ptr,
                                                 k, scale))
    want = np.zeros(words, np.uint64)
    for (cx, cy, cz), rr in zip(cen, r):
        oracle.lib.orc_mark_sphere(C.c_void_p(want.ctypes.data), C.c_float(cx), C.c_float(cy),
                                   C.c_float(cz), C.c_float(np.float32(scale * rr)),
                                   C.c_uint32(n), C.c_float(l))
    np.testing.assert_array_equal(dm.get(), want)
