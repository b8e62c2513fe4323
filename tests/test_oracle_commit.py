"""The oracle's commit path pinned by brute force, on the CPU.

The reference's exodus commit (src/finder/exodus_finder.c:2174-2329) refuses
a candidate whose node is painted (sif__mark_sphere, src/finder/utils.c:
965-1033), whose axis poles are painted (sif__overlap_quick, utils.c:
819-848), or which collides with an accepted void (sif__overlap_exact,
utils.c:850-948) at the rung radius or at its grown radius. The painting
and the poles are shortcuts of the collision test: a painted node lies within
s r_j of void j, a painted pole within s R (1 - f) - l of the node, and both
imply d < s (r + r_j - f min(r, r_j)). So the catalogue is that of a plain
greedy pass: candidates in depth order, each accepted when it collides with
no earlier void at its grown radius (or at the rung's). That pass is
restated here with no cell list, no mask and no batching -- a KD-tree over
the voids (scipy's, periodic) finds the pairs and the collision is the
float test itself -- and must reproduce orc_exodus_run's catalogue exactly.

Also here: the painting against a node-by-node FP64 minimum-image brute
force, the exact overlap test and the pole probe on constructed pairs at the
collision distance, and the tie order of equal depths.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle

THRESH = -0.8


@pytest.fixture(scope="module")
def config1():
    """128^3 Zel'dovich particles (seed 1) in a 250 Mpc/h box, the oracle's
    float pipeline (float CIC, FP64 transform, the reference's float filter
    factors) on a 128^3 grid, a 50^3 mesh, the 48-rung 5% ladder from two
    grid cells to 40 Mpc/h."""
    from sif_amd import zeldovich
    n, box, mc = 128, 250.0, 50
    x, y, z = zeldovich.generate(128, box, 1)
    radii = zeldovich.ladder(2.0 * box / n, 40.0, 0.05)
    om = oracle.OracleMesh(oracle.chain_mesh(x, y, z, mc, box), mc, box)
    g = oracle.cic_float(x, y, z, n, box)
    d, _ = oracle.density_contrast(g)
    spec = np.fft.rfftn(d.reshape(n, n, n).astype(np.float64))
    axis = oracle.window_axis(n, 2, 0)
    fields = []
    for r in radii:
        fac = oracle.mode_factors(oracle.filter_lut(1, r, 0.0, box, n), axis, n)
        fields.append(np.fft.irfftn(spec * fac, s=(n, n, n), axes=(0, 1, 2))
                      .astype(np.float32).ravel())
    return dict(x=x, y=y, z=z, n=n, box=box, om=om, radii=radii, fields=fields)


def _collides(x, y, z, r, vx, vy, vz, vr, f, s, box, pair_min=np.minimum):
    """Any collision of the sphere (x, y, z, r) with the voids (vx, vy, vz, vr):
    minimum-image d^2 < (s (r + r_j - f min(r, r_j)))^2 in float32, the
    operations of the reference's test."""
    if len(vx) == 0:
        return False
    f32 = np.float32
    half = f32(box) * f32(0.5)
    L = f32(box)
    mr = pair_min(f32(r), vr)
    tot = f32(s) * ((f32(r) + vr) - f32(f) * mr)
    t2 = tot * tot
    d2 = np.zeros(len(vx), np.float32)
    for p, v in ((x, vx), (y, vy), (z, vz)):
        a = np.abs(f32(p) - v)
        a = np.where(a > half, L - a, a)
        d2 = d2 + a * a
    return bool(np.any(d2 < t2))


def _hits_old(cen, r, old, tree, max_r, f, s, box, pair_min):
    """Per centre (k, 3) with radius r (k,): collides with any of the voids
    `old`. The pairs come from a periodic KD-tree over the centres, each void
    asking for those within its own largest reach, s (max r + r_j)."""
    hit = np.zeros(len(cen), bool)
    if tree is None or len(cen) == 0:
        return hit
    ctree = cKDTree(cen.astype(np.float64), boxsize=box)
    reach = float(s) * (float(np.max(r)) + old[:, 3].astype(np.float64)) * (1 + 1e-5)
    lists = ctree.query_ball_point(old[:, :3].astype(np.float64), reach, return_sorted=False)
    counts = np.fromiter((len(a) for a in lists), np.int64, len(lists))
    if counts.sum() == 0:
        return hit
    vi = np.repeat(np.arange(len(old)), counts)
    qi = np.fromiter((q for a in lists for q in a), np.int64, int(counts.sum()))
    f32 = np.float32
    half, L = f32(box) * f32(0.5), f32(box)
    rq, rv = r[qi].astype(np.float32), old[vi, 3]
    tot = f32(s) * ((rq + rv) - f32(f) * pair_min(rq, rv))
    d2 = np.zeros(len(qi), np.float32)
    for a in range(3):
        dd = np.abs(cen[qi, a] - old[vi, a])
        dd = np.where(dd > half, L - dd, dd)
        d2 = d2 + dd * dd
    np.logical_or.at(hit, qi, d2 < tot * tot)
    return hit


def greedy_catalogue(case, f, lagrangian, pair_min=np.minimum):
    """The plain greedy pass: per rung (descending), candidates with
    delta <= t in (delta, flat) order; a candidate colliding at the rung
    radius with a void is refused; else it is grown (orc_find_radius, pinned
    by its own brute force) and accepted unless it collides at the grown
    radius. The tests against the voids the rung began with run for the whole
    rung at once, those against the rung's own voids in order."""
    n, box, radii, fields, om = case["n"], case["box"], case["radii"], case["fields"], case["om"]
    l = np.float32(box) / np.float32(n)
    s = np.float32(np.cbrt(1.0 + THRESH)) if lagrangian else np.float32(1.0)
    vol = (np.float32(4.0) / np.float32(3.0)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(case["x"])) / (np.float32(box) * np.float32(box) * np.float32(box)))
    cat = np.zeros((0, 4), np.float32)
    rungs = []
    for i, (R, field) in enumerate(zip(radii, fields)):
        R = np.float32(R)
        rs = R + max((np.float32(1.5) - np.float32(1.0)) * R, np.float32(3.0) * l)
        if i > 0 and radii[i - 1] > rs:
            rs = np.float32(radii[i - 1])
        flat = np.nonzero(field <= np.float32(THRESH))[0]
        flat = flat[np.lexsort((flat, field[flat]))]
        ix, iy, iz = flat // (n * n), (flat // n) % n, flat % n
        cen = np.stack([ix, iy, iz], 1).astype(np.float32) * l
        old = cat
        max_r = float(old[:, 3].max()) if len(old) else 0.0
        tree = cKDTree(old[:, :3].astype(np.float64), boxsize=box) if len(old) else None
        live = ~_hits_old(cen, np.full(len(cen), R, np.float32), old, tree, max_r, f, s, box,
                          pair_min)
        cen, flat = cen[live], flat[live]
        grown = oracle.find_radius(om, cen, R, rs, THRESH, vol)[0] if len(cen) else \
            np.zeros(0, np.float32)
        ok = grown >= 0
        ok[ok] = ~_hits_old(cen[ok], grown[ok], old, tree, max_r, f, s, box, pair_min)
        new = []
        for k in np.nonzero(ok)[0]:
            c, rr = cen[k], grown[k]
            if new:
                nv = np.array(new, np.float32)
                if _collides(c[0], c[1], c[2], R, nv[:, 0], nv[:, 1], nv[:, 2], nv[:, 3], f, s,
                             box, pair_min) or \
                   _collides(c[0], c[1], c[2], rr, nv[:, 0], nv[:, 1], nv[:, 2], nv[:, 3], f, s,
                             box, pair_min):
                    continue
            new.append((c[0], c[1], c[2], rr))
            rungs.append(i)
        if new:
            cat = np.concatenate([cat, np.array(new, np.float32)])
    return cat, np.array(rungs)


@pytest.mark.parametrize("f,lagrangian", [(0.0, False), (0.5, False), (1.0, False),
                                          (0.0, True)])
def test_greedy_pass_reproduces_oracle_commit(config1, f, lagrangian):
    c = config1
    want, nin, rung, _ = oracle.exodus_run(c["fields"], c["n"], c["box"], c["om"], len(c["x"]),
                                           c["radii"], THRESH, overlap=f, lagrangian=lagrangian,
                                           tie_rule=1)
    got, grung = greedy_catalogue(c, f, lagrangian)
    assert len(want) > 1000
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(grung, rung)


def test_greedy_pass_fails_a_planted_mistake(config1):
    """The comparison can tell: overlap fraction 0.5 but the greedy pass taking
    f * max instead of f * min gives another catalogue."""
    c = config1
    want = oracle.exodus_run(c["fields"], c["n"], c["box"], c["om"], len(c["x"]), c["radii"],
                             THRESH, overlap=0.5, tie_rule=1)[0]
    got, _ = greedy_catalogue(c, 0.5, False, pair_min=np.maximum)
    assert not np.array_equal(got, want)


def test_tie_orders_give_one_catalogue(config1):
    """Equal candidate depths occur at this size; the reference's quicksort
    order (sort.h) and the (depth, flat) order give the same catalogue."""
    c = config1
    n_ties = 0
    for fld in c["fields"]:
        sel = fld[fld <= np.float32(THRESH)]
        n_ties += len(sel) - len(np.unique(sel))
    a = oracle.exodus_run(c["fields"], c["n"], c["box"], c["om"], len(c["x"]), c["radii"],
                          THRESH, tie_rule=0)
    b = oracle.exodus_run(c["fields"], c["n"], c["box"], c["om"], len(c["x"]), c["radii"],
                          THRESH, tie_rule=1)
    assert n_ties > 1000
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    np.testing.assert_array_equal(a[1], b[1])


# --- painting -----------------------------------------------------------------

def _bits(words, n):
    flat = np.arange(n ** 3, dtype=np.uint64)
    return ((words[flat >> np.uint64(6)] >> (flat & np.uint64(63))) & np.uint64(1)).astype(bool)


@pytest.mark.parametrize("n,box", [(48, 100.0), (40, 333.3)])
def test_mark_sphere_against_node_brute_force(n, box):
    """Every node whose FP64 minimum-image distance to the centre is below
    r (1 - 1e-6) is painted and every one above r (1 + 1e-6) is not; radii up
    to past half the box, centres off nodes and on them, wrapping every face."""
    rng = np.random.default_rng(n)
    l = np.float32(box) / np.float32(n)
    g = np.arange(n, dtype=np.float64) * float(l)
    k = 60
    cen = rng.uniform(0, box, (k, 3)).astype(np.float32)
    cen[: k // 3] = (rng.integers(0, n, (k // 3, 3)) * l).astype(np.float32)
    cen[k // 3: k // 3 + 6] = np.float32(box) - np.float32(0.25) * l  # on the far faces
    r = rng.uniform(0.2, 0.6 * box, k).astype(np.float32)
    r[:10] = rng.uniform(0.2, 2.0, 10) * l
    ambiguous = 0
    for (cx, cy, cz), rr in zip(cen, r):
        words = np.zeros((n ** 3 + 63) // 64, np.uint64)
        oracle.lib.orc_mark_sphere(C.c_void_p(words.ctypes.data), C.c_float(cx), C.c_float(cy),
                                   C.c_float(cz), C.c_float(rr), C.c_uint32(n), C.c_float(l))
        got = _bits(words, n).reshape(n, n, n)
        d2 = np.zeros((n, n, n))
        for ax, cc in enumerate((cx, cy, cz)):
            a = np.abs(g - float(cc))
            a = np.minimum(a, box - a)
            shape = [1, 1, 1]
            shape[ax] = n
            d2 = d2 + (a ** 2).reshape(shape)
        d = np.sqrt(d2)
        inside = d < float(rr) * (1 - 1e-6)
        outside = d > float(rr) * (1 + 1e-6)
        assert np.all(got[inside]) and not np.any(got[outside])
        ambiguous += int(np.count_nonzero(~inside & ~outside))
    assert ambiguous < 10


def test_overlap_quick_reads_the_six_poles():
    """sif__overlap_quick: the six nodes at +-p along each axis, wrapped; a
    numpy reading of the same mask bits."""
    rng = np.random.default_rng(7)
    n = 36
    words = rng.integers(0, 2 ** 63, (n ** 3 + 63) // 64, dtype=np.uint64)
    words &= rng.integers(0, 2 ** 63, len(words), dtype=np.uint64)  # about a quarter set
    words &= rng.integers(0, 2 ** 63, len(words), dtype=np.uint64)
    bits = _bits(words, n).reshape(n, n, n)
    for _ in range(3000):
        ix, iy, iz = (int(v) for v in rng.integers(0, n, 3))
        p = int(rng.integers(1, 2 * n))
        want = any(bits[(ix + s * p) % n, iy, iz] or bits[ix, (iy + s * p) % n, iz] or
                   bits[ix, iy, (iz + s * p) % n] for s in (1, -1))
        got = oracle.lib.orc_overlap_quick(C.c_void_p(words.ctypes.data), n, ix, iy, iz, p)
        assert bool(got) == want, (ix, iy, iz, p)


# --- the exact overlap test on constructed pairs --------------------------------

def _overlap(voids, max_r, queries, box, f, s):
    v = [np.ascontiguousarray(a, np.float32) for a in voids]
    q = [np.ascontiguousarray(a, np.float32) for a in queries]
    out = np.zeros(len(q[0]), np.uint8)
    p = lambda a: C.c_void_p(a.ctypes.data)
    oracle.lib.orc_overlap_exact_many(p(v[0]), p(v[1]), p(v[2]), p(v[3]), C.c_uint64(len(v[0])),
                                      C.c_float(max_r), p(q[0]), p(q[1]), p(q[2]), p(q[3]),
                                      C.c_uint64(len(q[0])), C.c_float(box), C.c_float(f),
                                      C.c_float(s), p(out))
    return out.astype(bool)


@pytest.mark.parametrize("f,s", [(0.0, 1.0), (0.5, 1.0), (1.0, 1.0),
                                 (0.3, float(np.cbrt(np.float32(0.2))))])
def test_overlap_exact_against_fp64_on_constructed_pairs(f, s):
    """Pairs placed at 0.999 and 1.001 of the collision distance
    s (r_a + r_b - f min(r_a, r_b)), along axes and diagonals, across every
    face of the box and inside it, with unequal radii (f min against f max)
    and the query larger and smaller than the void; the answer is the FP64
    minimum-image comparison."""
    rng = np.random.default_rng(int(f * 10) + 3)
    box = 500.0
    vx, vy, vz, vr, qx, qy, qz, qr, want = ([] for _ in range(9))
    dirs = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 1), (-2, 1, 3)]
    for _ in range(400):
        ra, rb = (float(v) for v in rng.uniform(2.0, 60.0, 2))
        thr = s * (ra + rb - f * min(ra, rb))
        d = np.array(dirs[rng.integers(len(dirs))], float)
        d /= np.linalg.norm(d)
        near_face = rng.random() < 0.5
        c = rng.uniform(0, box, 3)
        if near_face:
            ax = rng.integers(3)
            c[ax] = rng.uniform(0, 0.5 * thr) if rng.random() < 0.5 else box - rng.uniform(0, 0.5 * thr)
        for frac in (0.999, 1.001):
            q = (c + frac * thr * d) % box
            vx.append(c[0]); vy.append(c[1]); vz.append(c[2]); vr.append(rb)
            qx.append(q[0]); qy.append(q[1]); qz.append(q[2]); qr.append(ra)
    V = [np.array(a, np.float32) for a in (vx, vy, vz, vr)]
    Q = [np.array(a, np.float32) for a in (qx, qy, qz, qr)]
    for i in range(len(V[0])):
        # FP64 from the float32 values both sides see
        dd = np.array([float(Q[a][i]) - float(V[a][i]) for a in range(3)])
        dd = np.abs(dd)
        dd = np.minimum(dd, box - dd)
        ra, rb = float(Q[3][i]), float(V[3][i])
        want.append(np.sqrt((dd ** 2).sum()) < s * (ra + rb - f * min(ra, rb)))
    got = np.array([_overlap([V[a][i:i + 1] for a in range(4)], float(V[3][i]),
                             [Q[a][i:i + 1] for a in range(4)], box, f, s)[0]
                    for i in range(len(V[0]))])
    np.testing.assert_array_equal(got, np.array(want))
    assert 0.3 < np.mean(want) < 0.7


def test_overlap_exact_is_strict_at_equal_distance():
    """d^2 equal to the threshold's square, exactly in float: no collision
    (utils.c:850-948 compares with <). Voids on integer nodes, 10 apart, radii
    summing to 10."""
    box = 100.0
    V = [np.array([20.0], np.float32), np.array([30.0], np.float32),
         np.array([40.0], np.float32), np.array([4.0], np.float32)]
    for q, expect in (((30.0, 30.0, 40.0, 6.0), False), ((30.0, 30.0, 40.0, 6.25), True),
                      ((96.0, 30.0, 40.0, 20.0), False),    # 24 across the face, 4 + 20
                      ((96.0, 30.0, 40.0, 20.5), True)):
        Q = [np.array([v], np.float32) for v in q]
        assert _overlap(V, 4.0, Q, box, 0.0, 1.0)[0] == expect, q


def test_overlap_exact_sees_voids_larger_than_the_query():
    """The scanned region follows the catalogue's largest radius: a small query
    inside a large void's reach but far from its cell is still caught."""
    box = 1000.0
    V = [np.array([100.0, 600.0], np.float32), np.array([100.0, 600.0], np.float32),
         np.array([100.0, 600.0], np.float32), np.array([80.0, 3.0], np.float32)]
    Q = [np.array([175.0, 30.0], np.float32), np.array([100.0, 980.0], np.float32),
         np.array([100.0, 100.0], np.float32), np.array([3.0, 3.0], np.float32)]
    got = _overlap(V, 80.0, Q, box, 0.0, 1.0)
    # (175, 100, 100) is 75 from the large void, inside 80 + 3; (30, 980, 100)
    # is 70 away in x and 120 in y through the face, 139 in all: clear
    np.testing.assert_array_equal(got, [True, False])
