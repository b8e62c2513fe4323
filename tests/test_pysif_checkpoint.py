"""The pysif surface (sif_amd.pysif, over the host-struct C API of include/sif/)
and the exodus checkpoint (exodus_finder.h:409-549).

CPU: checkpoints written here from the documented layout (a 192-byte
header, the ladder, four columns of voids, CRC32 IEEE of the header with its
crc32 field zero followed by the rest -- zlib.crc32 is that CRC) are read
back through sif_finder_exodus_checkpoint_read, with its refusals (checksum,
size, version, precision, byte order) and the upgrade of older versions'
options; the mesh-size and grid-size helpers against their closed forms; the
array views' read-only and ownership rules (py_common.h:59-74).

GPU: the pysif pipeline's catalogue against the device path's on the same
particles; a kept checkpoint holds that catalogue; a run resumed from a
checkpoint saved mid-ladder returns, bit for bit, what the uninterrupted run
returns (exodus_finder.h:41-57); a checkpoint of another run is refused,
naming what differs; a finished run leaves no checkpoint behind.
"""
import gc
import glob
import os
import struct
import zlib

import numpy as np
import pytest

import sif_amd.pysif as pysif
from sif_amd.pysif import finders

HEADER = struct.Struct("<4sIIIIIIIII ddd Q d Q d ddd Q II QQ I 36s")
assert HEADER.size == 192


def write_checkpoint(path, radii, voids, *, version=3, n_prior=0, prior_crc=0, next_rung=None,
                     options=0, grid_n=32, box=100.0, threshold=-0.8, overlap=0.0,
                     n_particles=32 ** 3, total_weight=None, assignment=2, mismatched=0,
                     byte_order=0x01020304, is_double=0, corrupt_crc=False):
    radii = np.asarray(radii, np.float32)
    voids = np.asarray(voids, np.float32).reshape(-1, 4)
    nv = len(voids)
    fields = [b"XCKP", version, byte_order, is_double, 0, 0, options, len(radii),
              len(radii) if next_rung is None else next_rung, grid_n,
              float(np.float32(box)), float(np.float32(threshold)), float(np.float32(overlap)),
              n_particles, float(n_particles if total_weight is None else total_weight), 0, 0.0,
              0.0, 0.0, 0.0, n_prior, prior_crc, 0, nv, mismatched, assignment, b"\0" * 36]
    body = radii.tobytes() + b"".join(np.ascontiguousarray(voids[:, k]).tobytes()
                                      for k in range(4))
    crc = zlib.crc32(HEADER.pack(*fields) + body) & 0xffffffff
    fields[22] = crc ^ (1 if corrupt_crc else 0)
    with open(path, "wb") as f:
        f.write(HEADER.pack(*fields) + body)


def test_read_checkpoint_of_the_documented_layout(tmp_path):
    rng = np.random.default_rng(1)
    radii = np.array([12.0, 10.0, 8.0, 6.0], np.float32)
    voids = rng.uniform(0, 100, (7, 4)).astype(np.float32)
    p = tmp_path / "a.ckpt"
    write_checkpoint(p, radii, voids, next_rung=2, options=(3 << 12) | (1 << 14), n_prior=2,
                     prior_crc=0x1234, mismatched=3)
    cat, info = finders.read_checkpoint(p)
    np.testing.assert_array_equal(cat.centres, voids[:, :3])
    np.testing.assert_array_equal(cat.radii, voids[:, 3])
    assert info["rungs_done"] == 2 and info["n_voids"] == 7 and info["n_prior"] == 2
    np.testing.assert_array_equal(info["radii"], radii)
    assert info["search_factor"] == 2.0 and info["keep_cic_window"] and \
        info["deconvolution"] == "none"
    assert info["mode"] == "box" and info["assignment"] == "cic" and not info["interlaced"]
    assert info["threshold"] == np.float32(-0.8) and info["grid_n_cells"] == 32


def test_checkpoint_refusals(tmp_path):
    radii, voids = [10.0, 8.0], np.ones((3, 4), np.float32)
    cases = {"crc": dict(corrupt_crc=True), "version": dict(version=4),
             "double": dict(is_double=1), "order": dict(byte_order=0x04030201)}
    for name, kw in cases.items():
        p = tmp_path / f"{name}.ckpt"
        write_checkpoint(p, radii, voids, **kw)
        with pytest.raises(OSError):
            finders.read_checkpoint(p)
    p = tmp_path / "short.ckpt"
    write_checkpoint(p, radii, voids)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-4])
    with pytest.raises(OSError):
        finders.read_checkpoint(p)
    open(p, "wb").write(b"NOPE" + data[4:])
    with pytest.raises(OSError):
        finders.read_checkpoint(p)
    with pytest.raises(OSError):
        finders.read_checkpoint(tmp_path / "missing.ckpt")


def test_older_versions_are_upgraded(tmp_path):
    p = tmp_path / "v2.ckpt"
    # version 2: bit 10 was the Gaussian refinement; minima implied it, the
    # extended reach and displacement (exodus_finder.h:463-486)
    write_checkpoint(p, [10.0], np.zeros((0, 4)), version=2, options=(1 << 10))
    _, info = finders.read_checkpoint(p)
    assert info["refine_centres"] == "gaussian" and not info["extend_reach"]
    write_checkpoint(p, [10.0], np.zeros((0, 4)), version=2, options=(1 << 9))
    _, info = finders.read_checkpoint(p)
    assert info["seeds"] == "minima" and info["refine_centres"] == "gaussian"
    assert info["extend_reach"] and info["displace"]
    write_checkpoint(p, [10.0], np.zeros((0, 4)), version=1, assignment=0)
    _, info = finders.read_checkpoint(p)
    assert info["assignment"] == "cic"


def test_sizing_helpers():
    # about 30 tracers a cell, truncated; floored by the search sphere; [8, 1024]
    assert finders.suggest_mesh_cells(512 ** 3, 1000.0) == int(np.cbrt(512 ** 3 / 30.0))
    assert finders.suggest_mesh_cells(100, 10.0) == 8
    assert finders.suggest_mesh_cells(10 ** 12, 1.0) == 1024
    assert finders.suggest_mesh_cells(64, 100.0, 40.0) == 8
    assert finders.suggest_mesh_cells(64, 100.0, 48.0) == int(100.0 / (100.0 - 96.0) + 1.0)
    with pytest.raises(ValueError):
        finders.suggest_mesh_cells(64, 100.0, 50.0)
    # one mean separation a cell: 512 over a 1000 box of 512^3 tracers
    assert finders.grid_cells_from_size(512 ** 3, 1000.0, 1.0) == 512
    assert finders.grid_cells_from_size(0, 1000.0, 10.0, units="physical") == 100
    with pytest.raises(ValueError):
        finders.grid_cells_from_size(100, 1000.0, -1.0)
    assert finders.snap_search_factor(1.6) == 1.5 and finders.snap_search_factor(1.9) == 2.0
    with pytest.raises(ValueError):
        finders.snap_search_factor(0.5)


def test_views_are_read_only_and_keep_their_owner():
    x = np.linspace(0, 90, 50, dtype=np.float32)
    f = pysif.field_from_numpy(x, x[::-1].copy(), x + 1, weights=np.ones(50))
    vx = f.x
    assert not vx.flags.writeable and f.has_weights and not f.has_velocities
    del f
    gc.collect()
    np.testing.assert_array_equal(vx, x)  # the view kept the field
    with pytest.raises(ValueError):
        pysif.field_from_numpy(x, x[:10], x)
    cat = pysif.catalogue_from_numpy([1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [7.0, 8.0])
    assert cat.n_voids == 2 and cat.to_numpy()["r"].tolist() == [7.0, 8.0]


# ---- GPU --------------------------------------------------------------------

N_SIDE, BOX, SEED, GRID, MESH = 32, 100.0, 11, 32, 10


@pytest.fixture(scope="module")
def gpu_inputs(native):
    from sif_amd import zeldovich
    pysif.system_init()
    x, y, z = zeldovich.generate(N_SIDE, BOX, SEED)
    radii = zeldovich.ladder(2.0 * BOX / GRID, 20.0, 0.1)
    return x, y, z, radii


def pysif_run(x, y, z, radii, **kw):
    field = pysif.field_from_numpy(x, y, z)
    grid = pysif.Grid(GRID, BOX)
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(MESH, BOX, field, drop_indices=True)
    cat = finders.exodus_raw(grid, mesh, radii, -0.8, **kw)
    return np.c_[cat.centres, cat.radii]


@pytest.mark.gpu
def test_pysif_catalogue_matches_device_path(gpu_inputs, tmp_path):
    from sif_amd import device as D
    x, y, z, radii = gpu_inputs
    print("[pysif] exodus through the host C API", flush=True)
    got = pysif_run(x, y, z, radii, checkpoint=False)
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    grid = D.cic(dx, dy, dz, len(x), GRID, BOX)
    D.density_contrast(grid, GRID ** 3)
    mesh = D.Mesh(dx, dy, dz, len(x), MESH, BOX, indices=False)
    want, _, _ = D.exodus(grid, GRID, BOX, mesh, radii, -0.8, restore=False)
    assert len(want) > 10
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_checkpoint_kept_resumed_and_refused(gpu_inputs, tmp_path):
    x, y, z, radii = gpu_inputs
    print("[checkpoint] kept, resumed, refused", flush=True)
    path = tmp_path / "run.ckpt"
    full = pysif_run(x, y, z, radii, checkpoint=str(path), keep_checkpoint=True)
    assert path.exists()
    cat, info = finders.read_checkpoint(path)
    np.testing.assert_array_equal(np.c_[cat.centres, cat.radii], full)
    assert info["rungs_done"] == len(radii) and info["n_voids"] == len(full)
    np.testing.assert_array_equal(info["radii"], np.sort(radii)[::-1])
    # a run stopped after k rungs: its catalogue is the first k rungs' one
    k = len(radii) // 2
    part = pysif_run(x, y, z, np.sort(radii)[::-1][:k], checkpoint=False)
    assert 0 < len(part) < len(full)
    np.testing.assert_array_equal(full[:len(part)], part)
    raw = bytearray(open(path, "rb").read())
    h = list(HEADER.unpack(bytes(raw[:192])))
    resume = tmp_path / "resume.ckpt"
    write_checkpoint(resume, np.sort(radii)[::-1], part, next_rung=k, options=h[6],
                     grid_n=h[9], box=h[10], threshold=h[11], overlap=h[12],
                     n_particles=h[13], total_weight=h[14], assignment=h[25])
    resumed = pysif_run(x, y, z, radii, checkpoint=str(resume))
    np.testing.assert_array_equal(resumed, full)
    assert not resume.exists()  # removed once the run completed
    # another run's checkpoint is refused, naming what differs
    other = tmp_path / "other.ckpt"
    write_checkpoint(other, np.sort(radii)[::-1], part, next_rung=k, options=h[6],
                     grid_n=h[9], box=h[10], threshold=-0.7, overlap=h[12],
                     n_particles=h[13], total_weight=h[14], assignment=h[25])
    with pytest.raises(ValueError, match="threshold"):
        pysif_run(x, y, z, radii, checkpoint=str(other))
    # the variant call checkpoints under a name of its own, and removes it
    before = set(glob.glob("exodus-*.ckpt"))
    field = pysif.field_from_numpy(x, y, z)
    grid = pysif.Grid(GRID, BOX)
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(MESH, BOX, field)
    cat = finders.exodus(grid, mesh, radii, -0.8)
    np.testing.assert_array_equal(np.c_[cat.centres, cat.radii], full)
    assert set(glob.glob("exodus-*.ckpt")) == before
