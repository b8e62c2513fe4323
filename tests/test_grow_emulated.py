"""The sphere-growth kernel's own text (sif_amd/csrc/grow_device.h), run on
the CPU under a 64-lane wavefront emulator (tests/emu/), against the oracle's
restatement of exodus_rescale.c:642-1012 and the brute-force radius of
tests/test_finders.c:433-465. Radii, reasons and enclosed counts bit-exact.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from sif_amd import zeldovich

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
EMU_LIB = os.path.join(EMU_DIR, "libgrow_emu.so")

BOX, MESH = 100.0, 10


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
    lib = C.CDLL(EMU_LIB)
    lib.emu_find_radius.restype = C.c_int
    return lib


@pytest.fixture(scope="module")
def field():
    x, y, z = zeldovich.generate(32, BOX, 11)
    m = oracle.chain_mesh(x, y, z, MESH, BOX)
    return x, y, z, m


def run_emu(lib, m, tpl, centres, rung, rs, t, vol, cache_words=0):
    k = len(centres)
    r = np.zeros(k, np.float32)
    why = np.zeros(k, np.uint32)
    nin = np.zeros(k, np.uint64)
    cx, cy, cz = (np.ascontiguousarray(centres[:, i]) for i in range(3))
    V = C.c_void_p
    st = lib.emu_find_radius(
        V(m["offsets"].ctypes.data), V(m["x"].ctypes.data), V(m["y"].ctypes.data),
        V(m["z"].ctypes.data), C.c_int32(MESH), C.c_float(BOX), V(tpl["dx"].ctypes.data),
        V(tpl["dy"].ctypes.data), V(tpl["dz"].ctypes.data), V(tpl["type"].ctypes.data),
        V(tpl["min_d2"].ctypes.data), V(tpl["max_d2"].ctypes.data),
        C.c_uint32(len(tpl["dx"])), C.c_float(rung), C.c_float(rs), C.c_float(t),
        C.c_float(vol), V(cx.ctypes.data), V(cy.ctypes.data), V(cz.ctypes.data),
        C.c_uint32(k), V(r.ctypes.data), V(why.ctypes.data), V(nin.ctypes.data),
        C.c_uint32(cache_words))
    assert st == 0
    return r, why, nin


def annulus_sizes(x, y, z, centres, rung, rs):
    """Particles with rung^2 < d^2 <= rs^2 (periodic), about what a wave's
    first pass caches."""
    p = np.c_[x, y, z].astype(np.float64)
    out = []
    for c in centres.astype(np.float64):
        d = np.abs(p - c)
        d = np.minimum(d, BOX - d)
        d2 = (d * d).sum(1)
        out.append(int(((d2 > rung * rung) & (d2 <= rs * rs)).sum()))
    return np.array(out)


# cache_words 0: the LDS the library gives the launch (a cache of 1920 here);
# 1024, the least: the annulus is then sorted in windows after it (at most
# 512), sorted whole (512..1024), or too large to cache (walked in passes over
# the mesh), and the case covers all three.
@pytest.mark.parametrize("rung,factor,t,cache_words", [(6.0, 1.5, -0.8, 0),
                                                       (9.0, 2.0, -0.8, 1024),
                                                       (4.0, 1.25, -0.6, 0)])
def test_emulated_kernel_matches_oracle(emu, field, rung, factor, t, cache_words):
    x, y, z, m = field
    rs = rung + max((factor - 1.0) * rung, 3.0 * BOX / 32)
    tpl = oracle.OracleTemplate(MESH, np.float32(BOX) / np.float32(MESH), rung, rs).arrays()
    vol = (np.float32(4) / np.float32(3)) * np.float32(3.14159265358979323846) * \
        (np.float32(len(x)) / (np.float32(BOX) ** 3))
    centres = np.random.default_rng(int(rung * 10)).uniform(0, BOX, (64, 3)).astype(np.float32)
    if cache_words:
        ann = annulus_sizes(x, y, z, centres, rung, rs)
        half = cache_words // 2
        assert (ann <= half - 16).sum() >= 5
        assert ((ann > half + 16) & (ann <= cache_words - 16)).sum() >= 5
        assert (ann > cache_words + 16).sum() >= 5
    r, why, nin = run_emu(emu, m, tpl, centres, rung, rs, np.float32(t), vol, cache_words)
    om = oracle.OracleMesh(m, MESH, BOX)
    orr, owhy, onin = oracle.find_radius(om, centres, rung, rs, np.float32(t), vol)
    np.testing.assert_array_equal(why, owhy)
    np.testing.assert_array_equal(r.view(np.uint32), orr.view(np.uint32))
    np.testing.assert_array_equal(nin[r > 0], onin[orr > 0])
    for i in np.nonzero(r > 0)[0][:10]:
        want = oracle.brute_force_radius(x, y, z, BOX, centres[i], rung, rs, t)
        assert abs(r[i] - want) <= 1e-5 * want
