"""The path's file formats through the C ABI (include/sif/io/): .xfield,
.xgrid and the text catalogue. Host code only, so these run without a GPU.

The binary layouts are pinned byte for byte against files assembled here with
numpy and struct from the header definitions (field_io.h:51-84,
grid_io.h:35-84) and zlib's CRC-32, the checksum the reference writes
(include/sif/utils/crc32.h:29-43: IEEE 802.3, the reflected 0xEDB88320
polynomial). The catalogue cases restate
the reference's own known-answer tests (tests/test_io.c:96-452).
"""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

import sif_amd.pysif as pysif
from sif_amd.pysif import io as pio

BOX = 250.0


def _particles(n=1000, seed=3):
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, BOX, n).astype(np.float32) for _ in range(3)]
    vel = [rng.normal(0, 300, n).astype(np.float32) for _ in range(3)]
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return pos, vel, w


def _xfield_bytes(box, pos, vel=None, w=None, version=2, is_double=0):
    blocks = list(pos) + (list(vel) if vel is not None else []) + ([w] if w is not None else [])
    payload = b"".join(np.ascontiguousarray(b, np.float32).tobytes() for b in blocks)
    crc = zlib.crc32(payload) if version >= 2 else 0
    head = struct.pack("<4sIQdIIII24x", b"XFLD", version, len(pos[0]), box,
                       int(w is not None), int(vel is not None), is_double, crc)
    assert len(head) == 64
    return head + payload


def test_header_structs_are_64_bytes():
    assert C.sizeof(pio.XFieldHeader) == 64 and C.sizeof(pio.XGridHeader) == 64


@pytest.mark.parametrize("with_vel,with_w", [(False, False), (True, False), (False, True),
                                             (True, True)])
def test_xfield_bytes_match_the_layout(tmp_path, with_vel, with_w):
    pos, vel, w = _particles()
    path = tmp_path / "f.xfield"
    pio.write_field(path, BOX, *pos, *(vel if with_vel else (None,) * 3),
                    weights=w if with_w else None)
    want = _xfield_bytes(BOX, pos, vel if with_vel else None, w if with_w else None)
    assert path.read_bytes() == want
    h = pio.read_field_header(path)
    assert h == {"n_particles": 1000, "box_length": BOX, "has_weights": with_w,
                 "has_velocities": with_vel, "version": 2}


def test_xfield_round_trip_bit_exact(tmp_path):
    pos, vel, w = _particles(777)
    path = tmp_path / "f.xfield"
    pio.write_field(path, BOX, *pos, *vel, weights=w)
    f = pio.read_field(path)
    assert f.n_particles == 777 and f.has_velocities and f.has_weights
    for got, want in zip((f.x, f.y, f.z, f.vx, f.vy, f.vz, f.weights), pos + vel + [w]):
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_xfield_reads_files_built_elsewhere(tmp_path):
    """A file assembled here, as the reference writes one, reads back; so
    does a version-1 file, whose crc field is 0 and is not checked."""
    pos, vel, w = _particles(64, seed=9)
    for version in (1, 2):
        path = tmp_path / ("v%d.xfield" % version)
        path.write_bytes(_xfield_bytes(100.0, pos, None, w, version=version))
        f = pio.read_field(path)
        assert not f.has_velocities and f.has_weights
        np.testing.assert_array_equal(f.z, pos[2])
        np.testing.assert_array_equal(f.weights, w)


def test_xfield_refusals(tmp_path):
    pos, vel, w = _particles(50)
    good = _xfield_bytes(BOX, pos, vel, w)
    cases = {
        "corrupt": good[:64 + 40] + bytes([good[64 + 40] ^ 0x10]) + good[64 + 41:],
        "magic": b"XFLE" + good[4:],
        "version": good[:4] + struct.pack("<I", 3) + good[8:],
        "double": _xfield_bytes(BOX, pos, vel, w, is_double=1),
        "short": good[:-4],
        "header": good[:40],
    }
    for name, data in cases.items():
        path = tmp_path / (name + ".xfield")
        path.write_bytes(data)
        with pytest.raises(OSError):
            pio.read_field(path)
    with pytest.raises(OSError):
        pio.read_field(tmp_path / "missing.xfield")
    with pytest.raises(ValueError):
        pio.read_field_header(tmp_path / "magic.xfield")
    with pytest.raises(ValueError):
        pio.write_field(tmp_path / "empty.xfield", BOX, [], [], [])


def test_xfield_wrap_folds_into_the_declared_box(tmp_path):
    x = np.array([-1.0, 0.0, 10.0, 100.0, 250.0], np.float32)
    path = tmp_path / "w.xfield"
    pio.write_field(path, 100.0, x, x, x)
    f = pio.read_field(path, wrap=True)
    assert np.all((f.x >= 0) & (f.x < 100.0))
    assert pio.read_field(path).x.tolist() == x.tolist()


def _xgrid_bytes(n, box, values, content, assignment=0, version=2):
    payload = np.ascontiguousarray(values, np.float32).tobytes()
    crc = zlib.crc32(payload) if version >= 2 else 0
    head = struct.pack("<4sIIIQdII16sI4x", b"XGRD", version, n, 0, n ** 3, box, crc, content,
                       b"\0" * 16, assignment)
    assert len(head) == 64
    return head + payload


def test_xgrid_bytes_round_trip_and_refusals(tmp_path):
    n = 8
    g = pysif.Grid(n, BOX)
    vals = (np.arange(n ** 3, dtype=np.float64) * 0.017 - 1.0).astype(np.float32)
    C.memmove(g._p.contents.values, vals.ctypes.data, vals.nbytes)
    g._p.contents.content = 2  # a density contrast
    path = tmp_path / "g.xgrid"
    pio.write_grid(path, g)
    assert path.read_bytes() == _xgrid_bytes(n, BOX, vals, 2)
    back = pio.read_grid(path)
    np.testing.assert_array_equal(back.values.ravel(), vals)
    assert pio.grid_content(back) == "contrast" and back.assignment is None
    # CIC with interlacing comes back as such
    path.write_bytes(_xgrid_bytes(n, BOX, vals, 1, assignment=2 | (1 << 8)))
    back = pio.read_grid(path)
    assert back.assignment == "cic" and back.interlaced and pio.grid_content(back) == "density"
    # a scheme past PCS reads as unknown
    path.write_bytes(_xgrid_bytes(n, BOX, vals, 1, assignment=9))
    assert pio.read_grid(path).assignment is None
    # refused: a flipped payload bit, another geometry, another box
    good = _xgrid_bytes(n, BOX, vals, 2)
    path.write_bytes(good[:64 + 16] + bytes([good[64 + 16] ^ 1]) + good[64 + 17:])
    with pytest.raises(OSError):
        pio.read_grid(path)
    path.write_bytes(good)
    wrong = pysif.Grid(2 * n, BOX)
    with pytest.raises(OSError):
        pysif.N.check(pio.L.sif_grid_read_into(wrong._p, str(path).encode()))
    other_box = pysif.Grid(n, BOX + 1e-3)
    with pytest.raises(OSError):
        pysif.N.check(pio.L.sif_grid_read_into(other_box._p, str(path).encode()))
    # a version-1 grid carries no checksum and still reads
    path.write_bytes(_xgrid_bytes(n, BOX, vals, 2, version=1))
    np.testing.assert_array_equal(pio.read_grid(path).values.ravel(), vals)


# ---- the text catalogue: the reference's known-answer cases ----------------

def _write(path, text):
    path.write_text(text)
    return path


def _read(path, fmt=None):
    return pio.read_catalogue_ascii(path, fmt)


def test_catalogue_round_trip_is_bit_exact(tmp_path):
    cx = np.float32([1234.5678901234, -0.0009765625])
    cy = np.float32([0.1234567890123, 65536.03125])
    cz = np.float32([9.87654321e-7, 1e-8])
    r = np.float32([3.14159265358979, 0.5])
    cat = pysif.catalogue_from_numpy(cx, cy, cz, r)
    path = tmp_path / "c.txt"
    pio.write_catalogue_ascii(path, cat)
    lines = path.read_text().splitlines()
    assert lines[0] == "#n=2" and lines[1] == "#cx cy cz r"
    assert lines[2] == "%.9g %.9g %.9g %.9g" % (cx[0], cy[0], cz[0], r[0])
    back = _read(path)
    assert back.n_voids == 2 and back.footprint is None and back.units == "cartesian"
    np.testing.assert_array_equal(back.centres, np.stack([cx, cy, cz], 1))
    np.testing.assert_array_equal(back.radii, r)
    # numpy reads the rows as they are
    np.testing.assert_array_equal(np.loadtxt(path, dtype=np.float32)[:, 3], r)


def test_catalogue_footprint_round_trip(tmp_path):
    cat = pysif.catalogue_from_numpy([1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [7.0, 8.0],
                                     footprint=[0.123456789, 0.0],
                                     footprint_shell=[1.0, -1.0])
    path = tmp_path / "c.txt"
    pio.write_catalogue_ascii(path, cat)
    assert path.read_text().splitlines()[:2] == ["#n=2", "#cx cy cz r footprint footprint_shell"]
    back = _read(path)
    np.testing.assert_array_equal(back.footprint, np.float32([0.123456789, 0.0]))
    np.testing.assert_array_equal(back.footprint_shell, np.float32([1.0, -1.0]))
    rec = back.to_numpy()
    assert rec.dtype.names == ("cx", "cy", "cz", "r", "footprint", "footprint_shell")


def test_catalogue_legacy_and_blank_lines(tmp_path):
    p = tmp_path / "c.txt"
    with pytest.raises(OSError):
        _read(_write(p, "2\n1 2 3 4 0.5 0.25\n\n5 6 7 8\n"))  # rows of 6 then 4
    back = _read(_write(p, "2\n\n1 2 3 4\n  \n5 6 7 8\n"))
    assert back.n_voids == 2 and back.footprint is None and back.radii[1] == 8
    with pytest.raises(OSError):
        _read(_write(p, "3\n1 2 3 4\n5 6 7 8\n"))  # shorter than its count
    back = _read(_write(p, "1 2 3 4 0.5 0.25\n5 6 7 8 1 0\n"))  # six columns, no header
    assert back.n_voids == 2 and back.footprint.tolist() == [0.5, 1.0]


def test_catalogue_named_columns(tmp_path):
    p = tmp_path / "c.txt"
    back = _read(_write(p, "# made by hand\n#  id  R  CZ cy cx\n1 4 3 2 1\n# a note\n"
                           "2 8 7 6 5\n\n3 12 11 10 9\n"))
    assert back.n_voids == 3 and back.units == "cartesian" and back.footprint is None
    c = back.centres
    assert c[0, 0] == 1 and c[1, 1] == 6 and c[2, 2] == 11 and back.radii[2] == 12
    back = _read(_write(p, "#n = 1\n#ra dec z radius\n10 -5 0.5 20\n"))
    assert back.units == "sky" and back.centres[0, 2] == np.float32(0.5) and back.radii[0] == 20
    assert back.to_numpy().dtype.names == ("ra", "dec", "z", "r")
    # metadata comments are not names and not counts
    back = _read(_write(p, "#finder=\"exodus\"\n#n=1\n#n=5\n#cx cy cz r\n1 2 3 4\n"))
    assert back.n_voids == 1


@pytest.mark.parametrize("text", [
    "#n=1\n#ra cy z r\n1 2 3 4\n",        # sky and Cartesian mixed
    "#n=1\n1 2 3 4\n5 6 7 8\n",           # more rows than the count
    "#n=2\n#cx cy cz r\n1 2 3\n4 5 6\n",  # too few columns
    "#n=1\n#cx cy cz r\n1 2 x 4\n",       # not a number
    "#n=1\n" + " ".join(["1"] * 65) + "\n",  # more than 64 columns
    "#n=1\n1 2 3 " + "4" * 5000 + "\n",    # a line past 4094 characters
])
def test_catalogue_malformed_files_are_refused(tmp_path, text):
    with pytest.raises(OSError):
        _read(_write(tmp_path / "c.txt", text))


def test_catalogue_empty(tmp_path):
    p = tmp_path / "c.txt"
    pio.write_catalogue_ascii(p, pysif.Catalogue(1))
    assert p.read_text() == "#n=0\n#cx cy cz r\n"
    assert _read(p).n_voids == 0


def test_catalogue_formats(tmp_path):
    p = _write(tmp_path / "c.txt", "# ID x_centre y_centre z_centre volume R_eff type\n"
                                   "7, 1.5, 2.5, 3.5, 900, 12, main\n"
                                   "8, 4.5, 5.5, 6.5, 100, 6, sub\n")
    with pytest.raises(OSError):
        _read(p)  # an unnamed foreign layout
    back = _read(p, "* x y z * r")
    assert back.n_voids == 2 and back.units == "cartesian" and back.footprint is None
    assert back.centres[0, 0] == 1.5 and back.centres[1, 2] == 6.5
    assert back.radii.tolist() == [12, 6]
    p = _write(tmp_path / "d.txt", "#n=1\n#cx cy cz r\n20 -5 0.3 11\n")
    back = _read(p, "radecz,r")  # the format wins over the names
    assert back.units == "sky" and back.centres[0, 0] == 20 and back.centres[0, 1] == -5
    assert back.radii[0] == 11
    back = _read(p, "r cx cy cz")
    assert back.radii[0] == 20 and back.centres[0, 0] == -5
    for bad in ("x y z", "x y r", "x y z r r", "ra y z r", "x y z q r", "x y z *8 r"):
        with pytest.raises(ValueError):
            _read(p, bad)
    p = _write(tmp_path / "e.txt", "a 1 2 3 4\n")
    assert _read(p, "* x y z r").radii[0] == 4  # a skipped text column
    with pytest.raises(OSError):
        _read(p, "x y z r")  # a read one


def test_catalogue_messages_say_where(tmp_path):
    p = _write(tmp_path / "c.txt", "#n=2\n#cx cy cz r\n1 2 3 4\n\n1 2 oops 4\n")
    with pytest.raises(OSError, match=r"c\.txt:5: column 3 \(z\): 'oops' is not a number"):
        _read(p)
    with pytest.raises(ValueError, match="position 7: 'q' is not a column name"):
        _read(p, "x y z q r")


def test_catalogue_from_numpy_refusals():
    with pytest.raises(ValueError):
        pysif.catalogue_from_numpy([1.0], [2.0], None, [3.0])
    with pytest.raises(ValueError):
        pysif.catalogue_from_numpy([1.0], [2.0], [3.0], [4.0], ra=[1.0])
    with pytest.raises(ValueError):
        pysif.catalogue_from_numpy([1.0], [2.0], [3.0], [4.0], footprint=[1.0])
    sky = pysif.catalogue_from_numpy(ra=[10.0], dec=[20.0], z=[0.1], r=[5.0])
    assert sky.units == "sky"
