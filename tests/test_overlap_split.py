"""The exact overlap test of the commit, decided from overlap states
(sif_amd/csrc/overlap.h, the device index and the host's per-rung list)
instead of one scan of the whole catalogue, against the oracle's
restatement of sif__overlap_exact (src/finder/utils.c:850-948).

The driver splits the catalogue into the voids a rung began with (scanned on
the GPU, for a reference max_r of an earlier moment) and the voids accepted
since (scanned on the host, for the current max_r). The claims pinned here:

- state 1 for a set at some max_r means the reference scan of the whole
  catalogue at any later, larger max_r finds an overlap;
- state 0 means no void of the set can ever collide;
- the prefilter's refusal (state 1 at the rung-start max_r) is the
  reference's own answer at that moment.

Cases are built to sit on the reference's edges: centres on grid nodes,
voids placed at the collision distance along an axis, radii equal to the
catalogue's largest, scans that wrap the box, f in {0, 0.5, 1} and the
Lagrangian scale.
"""
import ctypes as C

import numpy as np
import pytest

import emu_lib
import oracle



@pytest.fixture(scope="module")
def emu():
    lib = emu_lib.load("overlap")
    lib.ovl_states.restype = None
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def states(lib, v, ref_max_r, q, box, f, scale, classes=4):
    out = np.zeros(len(q[0]), np.uint8)
    lib.ovl_states(_p(v[0]), _p(v[1]), _p(v[2]), _p(v[3]), C.c_uint64(len(v[0])),
                   C.c_float(ref_max_r), _p(q[0]), _p(q[1]), _p(q[2]), _p(q[3]),
                   C.c_uint64(len(q[0])), C.c_float(box), C.c_float(f), C.c_float(scale),
                   C.c_int(classes), _p(out))
    return out


def reference(v, max_r, q, box, f, scale):
    out = np.zeros(len(q[0]), np.uint8)
    oracle.lib.orc_overlap_exact_many(
        _p(v[0]), _p(v[1]), _p(v[2]), _p(v[3]), C.c_uint64(len(v[0])), C.c_float(max_r),
        _p(q[0]), _p(q[1]), _p(q[2]), _p(q[3]), C.c_uint64(len(q[0])), C.c_float(box),
        C.c_float(f), C.c_float(scale), _p(out))
    return out


def build_case(seed, box, n_grid, f, scale, on_nodes=True, nv=900):
    rng = np.random.default_rng(seed)
    l = np.float32(box) / np.float32(n_grid)
    if on_nodes:
        ijk = rng.integers(0, n_grid, (nv, 3))
        cen = (ijk.astype(np.float32) * l).astype(np.float32)
    else:
        cen = rng.uniform(0, box, (nv, 3)).astype(np.float32)
    ladder = np.float32(40.0) * np.float32(0.95) ** np.arange(40, dtype=np.float32)
    r = np.where(rng.random(nv) < 0.5, rng.choice(ladder, nv),
                 rng.uniform(3, 45, nv)).astype(np.float32)
    r[rng.random(nv) < 0.05] = r.max()          # ties with the largest radius
    order = np.argsort(-r, kind="stable")       # large voids first, as a ladder does
    cen, r = cen[order], r[order]
    voids = [np.ascontiguousarray(cen[:, 0]), np.ascontiguousarray(cen[:, 1]),
             np.ascontiguousarray(cen[:, 2]), r]

    # queries: random nodes, and nodes at the collision distance of a void
    # along one axis (one node either side of it)
    qs = []
    nq_rand = 3000
    ijk = rng.integers(0, n_grid, (nq_rand, 3))
    qs.append(np.c_[ijk.astype(np.float32) * l, rng.choice(ladder, nq_rand)])
    for _ in range(3000):
        j = rng.integers(0, nv)
        rr = np.float32(rng.choice(ladder) if rng.random() < 0.5 else rng.uniform(3, 60))
        mr = min(rr, r[j])
        tot = np.float32(scale) * (rr + r[j] - np.float32(f) * mr)
        k = int(np.floor(tot / l)) + rng.integers(-1, 2)
        axis = rng.integers(0, 3)
        c = cen[j].copy()
        c[axis] = np.float32((c[axis] + np.float32(k) * l * rng.choice([-1, 1])) % np.float32(box))
        qs.append(np.array([[c[0], c[1], c[2], rr]], np.float32))
    q = np.concatenate(qs).astype(np.float32)
    query = [np.ascontiguousarray(q[:, i]) for i in range(4)]
    return voids, query


CASES = [
    # box, grid, f, scale, centres on nodes, voids
    (1000.0, 512, 0.0, 1.0, True, 900),
    (1000.0, 512, 0.5, 1.0, True, 900),
    (1000.0, 512, 1.0, 1.0, True, 900),
    (250.0, 128, 0.0, float(np.cbrt(np.float32(0.2))), True, 40),
    (100.0, 32, 0.0, 1.0, True, 4),       # scans wider than the box
    (333.3, 100, 0.25, 1.0, False, 40),   # off-node centres (a prior catalogue)
]


@pytest.mark.parametrize("box,n_grid,f,scale,nodes,nv", CASES)
def test_split_decision_matches_reference_scan(emu, box, n_grid, f, scale, nodes, nv):
    voids, query = build_case(int(box) + n_grid, box, n_grid, f, scale, nodes, nv)
    nv = len(voids[0])
    checked = fallback = hits = 0
    for n_old in sorted({0, nv // 4, nv // 2, nv - 1, nv}):
        old = [a[:n_old].copy() for a in voids]
        new = [a[n_old:].copy() for a in voids]
        m_old = float(old[3].max()) if n_old else 0.0
        m_now = float(voids[3].max())
        truth = reference(voids, m_now, query, box, f, scale)
        # the prefilter: the rung-start answer is the reference's at that moment
        if n_old:
            pre = reference(old, m_old, query, box, f, scale)
            s_pre = states(emu, old, m_old, query, box, f, scale)
            np.testing.assert_array_equal(s_pre == 1, pre == 1)
        for m_b in (m_old, 0.5 * (m_old + m_now), m_now):
            m_b = np.float32(max(m_b, m_old))
            s_old = states(emu, old, m_b, query, box, f, scale) if n_old else \
                np.zeros(len(query[0]), np.uint8)
            s_new = states(emu, new, m_now, query, box, f, scale) if n_old < nv else \
                np.zeros(len(query[0]), np.uint8)
            # state 1 at an earlier max_r is an overlap now
            assert np.all(truth[s_old == 1] == 1)
            decided = s_old == 0
            assert np.all(truth[decided & (s_new == 1)] == 1)
            assert np.all(truth[decided & (s_new == 0)] == 0)
            checked += int(decided.sum())
            fallback += int(((s_old == 2) | (decided & (s_new == 2))).sum())
            hits += int(truth.sum())
        # the whole catalogue as one set, at its own max_r, in one list and
        # in radius classes
        for classes in (1, 4):
            s_all = states(emu, voids, m_now, query, box, f, scale, classes)
            assert np.all(truth[s_all == 1] == 1)
            assert np.all(truth[s_all == 0] == 0)
    # the cases reach both answers, and the host's full scan stays rare
    assert hits > 0 and checked > 0
    assert fallback <= 0.01 * checked
