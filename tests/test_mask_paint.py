"""The host mask (sif_amd/csrc/mask.h): sphere painting by z runs against
the oracle's node-by-node restatement of sif__mark_sphere
(src/finder/utils.c:965-1033), bit for bit, and the pole probe of
sif__overlap_quick (utils.c:819-848) against a direct reading of the bits.

Spheres are centred on nodes (the driver's case) and off them (a prior
catalogue), small and large, with spheres wider than half the box (painted
node by node) and spheres that wrap the box on every axis.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
    lib = C.CDLL(os.path.join(EMU_DIR, "libmask_emu.so"))
    lib.mask_paint.restype = None
    lib.mask_quick.restype = C.c_int
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def oracle_paint(n, l, spheres):
    words = np.zeros((n ** 3 + 63) // 64, np.uint64)
    for cx, cy, cz, r in spheres:
        oracle.lib.orc_mark_sphere(_p(words), C.c_float(cx), C.c_float(cy), C.c_float(cz),
                                   C.c_float(r), C.c_uint32(n), C.c_float(l))
    return words


def runs_paint(lib, n, l, spheres):
    s = np.ascontiguousarray(spheres, np.float32)
    words = np.zeros((n ** 3 + 63) // 64, np.uint64)
    cols = [np.ascontiguousarray(s[:, i]) for i in range(4)]
    lib.mask_paint(C.c_uint32(n), C.c_float(l), C.c_uint64(len(s)), _p(cols[0]), _p(cols[1]),
                   _p(cols[2]), _p(cols[3]), _p(words))
    return words


@pytest.mark.parametrize("n,box,nodes", [(64, 100.0, True), (100, 333.3, True),
                                         (128, 250.0, False), (37, 50.0, True)])
def test_painting_matches_reference_bits(emu, n, box, nodes):
    rng = np.random.default_rng(n)
    l = np.float32(box) / np.float32(n)
    k = 60
    if nodes:
        cen = rng.integers(0, n, (k, 3)).astype(np.float32) * l
    else:
        cen = rng.uniform(0, box, (k, 3)).astype(np.float32)
    r = np.concatenate([rng.uniform(0.3, 3, k // 3) * l, rng.uniform(3, 20, k // 3) * l,
                        rng.uniform(0.2, 0.7, k - 2 * (k // 3)) * box]).astype(np.float32)
    spheres = np.c_[cen, r].astype(np.float32)
    # whole catalogue, and one sphere at a time
    np.testing.assert_array_equal(runs_paint(emu, n, l, spheres), oracle_paint(n, l, spheres))
    for sph in spheres[::7]:
        np.testing.assert_array_equal(runs_paint(emu, n, l, sph[None]), oracle_paint(n, l, [sph]))


def test_pole_probe_reads_the_six_axis_nodes(emu):
    n, l = 40, np.float32(2.5)
    rng = np.random.default_rng(3)
    spheres = np.c_[rng.integers(0, n, (12, 3)).astype(np.float32) * l,
                    rng.uniform(1, 8, 12)].astype(np.float32)
    words = oracle_paint(n, l, spheres)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[: n ** 3].reshape(n, n, n)
    for _ in range(400):
        ix, iy, iz = (int(v) for v in rng.integers(0, n, 3))
        r = int(rng.integers(1, n + 5))
        want = any(bits[(ix + s * r) % n, iy, iz] or bits[ix, (iy + s * r) % n, iz] or
                   bits[ix, iy, (iz + s * r) % n] for s in (1, -1))
        assert emu.mask_quick(C.c_uint32(n), _p(words), ix, iy, iz, r) == int(want)
