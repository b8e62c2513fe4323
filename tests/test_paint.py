"""The mask of accepted-void nodes: the painting predicate
(sif_amd/csrc/paint.h, used by the device paint kernel and by the host's
probe of a rung's new voids) against the oracle's node-by-node restatement of
sif__mark_sphere (src/finder/utils.c:965-1033), bit for bit; and the host
probe (sif_amd/csrc/rung_voids.h) against the oracle's bits and overlap scan:
its node and pole answers (sif__overlap_quick, utils.c:819-848) exactly, its
overlap state by the rules of overlap.h.

Spheres are centred on nodes (the driver's case) and off them (a prior
catalogue), small and large, wider than half the box, and wrapping the box on
every axis.
"""
import ctypes as C

import numpy as np
import pytest

import emu_lib
import oracle



@pytest.fixture(scope="module")
def emu():
    lib = emu_lib.load("paint")
    lib.paint_nodes.restype = None
    lib.rv_probe_many.restype = None
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def oracle_paint(n, l, spheres):
    words = np.zeros((n ** 3 + 63) // 64, np.uint64)
    for cx, cy, cz, r in spheres:
        oracle.lib.orc_mark_sphere(_p(words), C.c_float(cx), C.c_float(cy), C.c_float(cz),
                                   C.c_float(r), C.c_uint32(n), C.c_float(l))
    return words


def predicate_paint(lib, n, l, spheres, all_nodes):
    s = np.ascontiguousarray(spheres, np.float32)
    cols = [np.ascontiguousarray(s[:, i]) for i in range(4)]
    words = np.zeros((n ** 3 + 63) // 64, np.uint64)
    lib.paint_nodes(C.c_uint32(n), C.c_float(l), C.c_uint64(len(s)), _p(cols[0]), _p(cols[1]),
                    _p(cols[2]), _p(cols[3]), C.c_int(1 if all_nodes else 0), _p(words))
    return words


def spheres_for(rng, n, box, k, nodes):
    l = np.float32(box) / np.float32(n)
    if nodes:
        cen = rng.integers(0, n, (k, 3)).astype(np.float32) * l
    else:
        cen = rng.uniform(0, box, (k, 3)).astype(np.float32)
    r = np.concatenate([rng.uniform(0.3, 3, k // 3) * l, rng.uniform(3, 12, k // 3) * l,
                        rng.uniform(0.2, 0.7, k - 2 * (k // 3)) * box]).astype(np.float32)
    return l, np.c_[cen, r].astype(np.float32)


@pytest.mark.parametrize("n,box,nodes,all_nodes", [(37, 50.0, True, True),
                                                   (64, 100.0, False, True),
                                                   (100, 333.3, True, False),
                                                   (128, 250.0, False, False)])
def test_predicate_matches_reference_bits(emu, n, box, nodes, all_nodes):
    rng = np.random.default_rng(n)
    l, spheres = spheres_for(rng, n, box, 24 if all_nodes else 60, nodes)
    np.testing.assert_array_equal(predicate_paint(emu, n, l, spheres, all_nodes),
                                  oracle_paint(n, l, spheres))


# nv 30: the probe walks the short list directly; nv 1500 (small radii):
# it scans the cells around the candidate
@pytest.mark.parametrize("n,box,f,scale,nv,rfrac", [
    (64, 200.0, 0.0, 1.0, 30, 0.12), (100, 333.3, 0.5, 1.0, 30, 0.12),
    (50, 100.0, 0.0, float(np.cbrt(np.float32(0.2))), 30, 0.12),
    (128, 500.0, 0.0, 1.0, 1500, 0.02), (96, 300.0, 0.5, 1.0, 1500, 0.025)])
def test_rung_probe(emu, n, box, f, scale, nv, rfrac):
    rng = np.random.default_rng(n + 7)
    l = np.float32(box) / np.float32(n)
    ijk = rng.integers(0, n, (nv, 3))
    cen = (ijk.astype(np.float32) * l).astype(np.float32)
    r = rng.uniform(1.0, rfrac * box, nv).astype(np.float32)
    v = [np.ascontiguousarray(cen[:, i]) for i in range(3)] + [r]
    painted = oracle_paint(n, l, np.c_[cen, np.float32(scale) * r].astype(np.float32))
    bits = np.unpackbits(painted.view(np.uint8), bitorder="little")[: n ** 3].reshape(n, n, n)

    # queries: random nodes and nodes near void surfaces, with and without
    # the overlap part, the node test and the poles
    nq = 3000
    q = rng.integers(0, n, (nq, 3)).astype(np.uint32)
    near = rng.integers(0, nv, nq // 2)
    off = rng.normal(0, 1, (nq // 2, 3))
    off = off / np.linalg.norm(off, axis=1)[:, None] * (scale * r[near])[:, None] / l
    q[: nq // 2] = (ijk[near] + np.rint(off)).astype(np.int64) % n
    qr = np.where(rng.random(nq) < 0.8, rng.uniform(1.0, rfrac * box, nq), -1.0).astype(np.float32)
    want_node = (rng.random(nq) < 0.7).astype(np.uint8)
    pole = np.where(rng.random(nq) < 0.8, rng.integers(1, n // 3, nq), 0).astype(np.int32)
    ref_max_r = np.float32(r.max() * 1.1)
    st, nd, pl = (np.zeros(nq, np.uint8) for _ in range(3))
    qix, qiy, qiz = (np.ascontiguousarray(q[:, i]) for i in range(3))
    emu.rv_probe_many(C.c_uint64(nv), _p(v[0]), _p(v[1]), _p(v[2]), _p(v[3]), C.c_float(box),
                      C.c_uint32(n), C.c_float(scale), C.c_float(f), C.c_float(ref_max_r),
                      C.c_uint64(nq), _p(qix), _p(qiy), _p(qiz), _p(qr), _p(want_node),
                      _p(pole), _p(st), _p(nd), _p(pl))
    # The caller refuses on the node, then the poles, then the overlap, and
    # the probe stops once that decision is fixed: the node answer is exact,
    # the pole answer exact where the node does not refuse, the state
    # answered where neither does.
    want_nd = bits[q[:, 0], q[:, 1], q[:, 2]].astype(bool) & (want_node == 1)
    np.testing.assert_array_equal(nd.astype(bool), want_nd)
    pm = pole.astype(np.int64)
    hit = np.zeros(nq, bool)
    for a in range(3):
        for sgn in (1, -1):
            qq = q.astype(np.int64).copy()
            qq[:, a] = (qq[:, a] + sgn * pm) % n
            hit |= bits[qq[:, 0], qq[:, 1], qq[:, 2]].astype(bool)
    want_pl = hit & (pole > 0)
    np.testing.assert_array_equal(pl.astype(bool)[~want_nd], want_pl[~want_nd])
    assert want_nd.sum() > 50 and (want_pl & ~want_nd).sum() > 50
    # overlap state against the reference scan of the same voids
    open_ = ~want_nd & ~want_pl
    sel = (qr >= 0) & open_
    qx = (q[:, 0].astype(np.float32) * l).astype(np.float32)
    qy = (q[:, 1].astype(np.float32) * l).astype(np.float32)
    qz = (q[:, 2].astype(np.float32) * l).astype(np.float32)
    truth = np.zeros(nq, np.uint8)
    oracle.lib.orc_overlap_exact_many(
        _p(v[0]), _p(v[1]), _p(v[2]), _p(v[3]), C.c_uint64(nv), C.c_float(ref_max_r),
        _p(qx), _p(qy), _p(qz), _p(qr), C.c_uint64(nq), C.c_float(box), C.c_float(f),
        C.c_float(scale), _p(truth))
    assert np.all(truth[sel & (st == 1)] == 1)
    assert np.all(truth[sel & (st == 0)] == 0)
    assert np.all(st[qr < 0] == 0)
    assert (sel & (st == 1)).sum() > 50 and (sel & (st == 0)).sum() > 50
