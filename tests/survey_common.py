"""The survey scenario of the reference's tests/test_survey.c (footprint,
holes, selection) and the CPU survey pipeline of the oracle, shared by the
oracle's own tests and the GPU parity tests.

The scenario: a 400 Mpc/h box, a 64^3 grid, a footprint ball of radius 140
at the centre less a masked stripe along z, three voids carved well inside
it and one cut by its edge; data keep 3% of what falls in a hole. Positions
are drawn with numpy's generator (the reference's own LCG stream is not
reproduced: its checks are on the geometry, not on particular tracers).
"""
import numpy as np

BOX = 400.0
N_GRID = 64
CELL = BOX / N_GRID
FOOT = np.array([200.0, 200.0, 200.0])
FOOT_R = 140.0
N_DATA = 100_000
N_RANDOMS = 10 * N_DATA
N_SPARSE = 3 * N_DATA
RADII = np.array([30.0, 24.0, 19.0, 15.0], np.float32)
THRESHOLD = -0.7
INNER = np.array([[160.0, 160.0, 160.0, 26.0], [240.0, 160.0, 240.0, 22.0],
                  [150.0, 240.0, 220.0, 20.0]])
EDGE = np.array([200.0, 200.0, 325.0, 25.0])


def in_footprint(p):
    d = p - FOOT
    inside = (d * d).sum(axis=1) <= FOOT_R * FOOT_R
    stripe = (np.abs(p[:, 0] - 200.0) < 12.0) & (np.abs(p[:, 1] - 260.0) < 12.0)
    return inside & ~stripe


def kept_fraction(z, selection):
    """1 at the bottom of the footprint, 0.4 at the top (test_survey.c:109-114)."""
    if not selection:
        return np.ones_like(z)
    return 1.0 - 0.6 * (z - (FOOT[2] - FOOT_R)) / (2.0 * FOOT_R)


def draw(n, selection, carve, rng, offset=0.0):
    """n points over the footprint under the selection; with `carve`, all but
    3% of what falls in a hole dropped (test_survey.c:119-147)."""
    out = []
    got = 0
    holes = np.vstack([INNER, EDGE[None, :]])
    while got < n:
        m = max(2 * (n - got), 4096)
        p = FOOT + (2.0 * rng.random((m, 3)) - 1.0) * FOOT_R
        keep = in_footprint(p)
        keep &= rng.random(m) <= kept_fraction(p[:, 2], selection)
        if carve:
            d2 = ((p[:, None, :] - holes[None, :, :3]) ** 2).sum(axis=2)
            carved = (d2 < holes[None, :, 3] ** 2).any(axis=1)
            keep &= ~carved | (rng.random(m) <= 0.03)
        p = p[keep][: n - got]
        out.append(p)
        got += len(p)
    p = np.vstack(out) + offset
    return [np.ascontiguousarray(p[:, k], dtype=np.float32) for k in range(3)]


def suggest_mesh_cells(n_particles, box, max_radius):
    """sif_finder_suggest_mesh_cells (exodus_finder.h:654-685)."""
    cells = (n_particles / 30.0) ** (1.0 / 3.0)
    least = box / (box - 2.0 * max_radius) + 1.0
    cells = max(cells, least, 8.0)
    return int(min(cells, 1024.0))


def smoothed_densities(rho, n, box, radii, oracle):
    """The oracle's smoothing of a CIC density grid: r2c, top-hat times the
    CIC window (1/W), c2r in double, one float field per rung."""
    spec = oracle.fft_r2c(np.asarray(rho, np.float64), n)
    axis = oracle.window_axis(n, 2, 0)
    out = []
    for r in radii:
        lut = oracle.filter_lut(1, float(r), 0.0, box, n)
        fac = oracle.mode_factors(lut, axis, n)
        out.append((oracle.fft_c2r(spec * fac, n) / n ** 3).astype(np.float32))
    return out


def cpu_survey(data, randoms, oracle, n=N_GRID, box=BOX, radii=RADII, threshold=THRESHOLD,
               wd=None, wr=None, overlap=0.0):
    """The whole survey run on the CPU oracle: (catalogue (k, 4), footprint
    (k, 2), occupancy, alpha)."""
    x, y, z = data
    rx, ry, rz = randoms
    rho_d = oracle.cic_float(x, y, z, n, box, w=wd)
    rho_r = oracle.cic_float(rx, ry, rz, n, box, w=wr)
    desc = np.sort(np.asarray(radii, np.float32))[::-1].copy()
    md = suggest_mesh_cells(len(x), box, float(desc[0]))
    mr = suggest_mesh_cells(len(rx), box, float(desc[0]))
    dm = oracle.chain_mesh(x, y, z, md, box, w=wd)
    rm = oracle.chain_mesh(rx, ry, rz, mr, box, w=wr)
    om_d = oracle.OracleMesh(dm, md, box)
    om_r = oracle.OracleMesh(rm, mr, box)
    occ, k, lam, marked, observed = oracle.survey_occupancy(rm["x"], rm["y"], rm["z"], n, box)
    alpha = np.float32(dm["total_weight"] / rm["total_weight"])
    sd = smoothed_densities(rho_d, n, box, desc, oracle)
    sr = smoothed_densities(rho_r, n, box, desc, oracle)
    fields = [oracle.survey_contrast(a, b, occ, alpha) for a, b in zip(sd, sr)]
    K = np.float32(np.float32(1.0 + np.float32(threshold)) * alpha)
    cat, n_in, rung, stats = oracle.exodus_run_survey(fields, n, box, om_d, om_r, float(K), desc,
                                                      threshold, overlap=overlap)
    fin, fsh = oracle.survey_footprint(occ, n, box, cat)
    return cat, np.stack([fin, fsh], axis=1), occ, float(alpha)


def nearest(cat, hole):
    if len(cat) == 0:
        return -1, np.inf
    d = np.sqrt(((cat[:, :3].astype(np.float64) - hole[:3]) ** 2).sum(axis=1))
    i = int(np.argmin(d))
    return i, float(d[i])


def check_holes(cat, fp, label):
    """test_survey.c:254-312: the three inner voids and perhaps the edge one,
    nothing else; the inner ones whole in the footprint, the edge one partly."""
    assert len(cat) in (3, 4), f"{label}: {len(cat)} voids, expected the 3 inner ones and " \
                               f"perhaps the edge one"
    for h in INNER:
        i, d = nearest(cat, h)
        assert d <= 2.0 * CELL, f"{label}: hole at {h[:3]} not recovered ({d:.1f} away)"
        r = float(cat[i, 3])
        assert 0.9 * h[3] <= r <= 1.5 * h[3], f"{label}: hole r {r:.2f} vs {h[3]}"
        assert fp[i, 0] == 1.0 and fp[i, 1] > 0.99, f"{label}: inner footprint {fp[i]}"
    assert ((fp >= 0.0) & (fp <= 1.0)).all()
    e, d = nearest(cat, EDGE)
    if d <= 2.0 * CELL:
        assert 0.3 < fp[e, 0] < 0.99, f"{label}: edge void footprint {fp[e]}"
        assert fp[e, 1] < fp[e, 0]
    else:
        assert len(cat) == 3, f"{label}: an extra void that is not the edge one"
