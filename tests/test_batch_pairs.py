"""The batch's pair effects on the device (sif_amd/csrc/batch_pairs_device.h)
against the host's pair_effects() (sif_amd/csrc/batch_pairs.h), on the CPU.

The commit of src/finder/exodus_finder.c:2174-2329 ORs, for each candidate,
the effects of the batch's earlier candidates it has accepted: paint its node
or a pole node, or an overlap state at the rung radius or at its grown radius.
The device kernel lists them with the host's own functions, one wave per
candidate; its text runs here under the 64-lane wave emulator (tests/emu).
Each candidate's list must hold exactly the host's effects for the same
tests (the device's order is by earlier candidate, the host's by cell), on
batches built so that every kind of effect occurs: clustered nodes, failed
growths, grown states with the pole bit and states 1 and 2, overlap fractions
0 and 0.5, the Lagrangian scale, a catalogue radius that puts pairs on the
reference scan's edge, and candidates near the box faces.
"""
import ctypes as C

import numpy as np
import pytest

from emu_lib import load

P = C.c_void_p
N, BOX = 32, 100.0
L = BOX / N


@pytest.fixture(scope="module")
def lib():
    lib = load("pairs")
    common = [P, P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
              C.c_float, C.c_float, C.c_int32]
    lib.emu_batch_pairs.argtypes = common + [C.c_uint64, P, P, C.c_int]
    lib.emu_batch_pairs.restype = C.c_int
    lib.host_batch_pairs.argtypes = common + [P, P, C.c_int]
    lib.host_batch_pairs.restype = C.c_int
    return lib


def batch(seed, count, near_face):
    """Candidate nodes clustered around a few centres (so that their spheres
    meet), grown radii (a tenth failed), and grown states."""
    rng = np.random.default_rng(seed)
    centres = rng.integers(0, N, (4, 3))
    if near_face:
        centres[0] = (0, N - 1, 1)
    pick = centres[rng.integers(0, len(centres), count)]
    nodes = (pick + rng.integers(-4, 5, (count, 3))) % N
    flat = ((nodes[:, 0] * N + nodes[:, 1]) * N + nodes[:, 2]).astype(np.uint32)
    rad = rng.uniform(2.0, 14.0, count).astype(np.float32)
    rad[rng.random(count) < 0.1] = -1.0
    gs = rng.choice(np.array([0, 0, 0, 1, 2, 16, 17], np.uint32), count)
    return flat, rad, gs


def run(lib, flat, rad, gs, radius, f, scale, cat_max_r, cap=128, capacity=None):
    count = len(flat)
    proxy = int(np.float32(scale) * np.float32(radius) * (np.float32(1) - np.float32(f))
                / np.float32(L)) - 1
    args = (flat.ctypes.data, rad.ctypes.data, gs.ctypes.data, count, N, L, BOX, f, scale,
            radius, cat_max_r, proxy)
    dn = np.zeros(count, np.int32)
    dout = np.zeros(count * cap, np.uint32)
    hn = np.zeros(count, np.int32)
    hout = np.zeros(count * cap, np.uint32)
    cap_total = count * 255 if capacity is None else capacity
    assert lib.emu_batch_pairs(*args, cap_total, dn.ctypes.data, dout.ctypes.data, cap) == 0
    assert lib.host_batch_pairs(*args, hn.ctypes.data, hout.ctypes.data, cap) == 0
    return dn, dout.reshape(count, cap), hn, hout.reshape(count, cap)


@pytest.mark.parametrize("seed,f,lagrangian,radius,cat_max_r,near_face,count", [
    (1, 0.0, False, 5.0, 14.0, False, 96),
    (2, 0.5, False, 4.0, 20.0, True, 96),
    (3, 0.0, True, 6.0, 0.0, False, 96),
    (4, 1.0, False, 3.5, 9.0, True, 96),
    (6, 0.0, False, 10.0, 14.0, False, 96),
    (7, 0.5, True, 12.0, 25.0, True, 96),
    # past one round of the scan's four groups of 64 earlier candidates
    (8, 0.0, False, 5.0, 14.0, True, 300),
])
def test_device_lists_the_hosts_effects(lib, seed, f, lagrangian, radius, cat_max_r,
                                        near_face, count):
    flat, rad, gs = batch(seed, count, near_face)
    scale = float(np.cbrt(np.float64(1.0 - 0.8))) if lagrangian else 1.0
    dn, dout, hn, hout = run(lib, flat, rad, gs, radius, f, scale, cat_max_r)
    kinds = np.zeros(6, np.int64)
    for b in range(len(flat)):
        assert hn[b] >= 0 and dn[b] == hn[b], (b, dn[b], hn[b])
        d = np.sort(dout[b, :dn[b]])
        h = np.sort(hout[b, :hn[b]])
        np.testing.assert_array_equal(d, h, err_msg=f"candidate {b}")
        # the device lists by earlier candidate, ascending
        assert np.all(np.diff(dout[b, :dn[b]] >> 8) > 0)
        assert np.all((d >> 8) < b)
        for code in d:
            kinds += [code & 1, (code >> 1) & 1, (code >> 2) & 1, ((code >> 3) & 3) == 1,
                      ((code >> 3) & 3) == 2, ((code >> 5) & 3) != 0]
    # the batch exercised every kind of effect but the ones its geometry rules
    # out: rung poles need a pole distance of a node or more
    assert kinds[0] > 0 and kinds[3] > 0 and kinds[5] > 0, kinds
    proxy = int(np.float32(scale) * np.float32(radius) * (np.float32(1) - np.float32(f))
                / np.float32(L)) - 1
    assert (kinds[1] > 0) == (proxy > 0), (kinds, proxy)


def test_list_room_and_overflow(lib):
    """A candidate whose effects run past the batch's list, or past the
    per-candidate count the host reads, is marked for the commit's own probe,
    never given a partial list."""
    flat, rad, gs = batch(5, 64, False)
    dn, dout, hn, hout = run(lib, flat, rad, gs, 5.0, 0.0, 1.0, 14.0, capacity=40)
    total = 0
    for b in range(len(flat)):
        if dn[b] < 0:
            assert total + hn[b] > 40
        else:
            assert dn[b] == hn[b]
            np.testing.assert_array_equal(np.sort(dout[b, :dn[b]]), np.sort(hout[b, :hn[b]]))
        total += hn[b]
    assert (dn < 0).any()
