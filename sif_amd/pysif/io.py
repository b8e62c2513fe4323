"""pysif.io for the path's formats (python/pysif/io.pyi of sif v0.5.0): the
.xfield particle binary, the .xgrid grid binary and the text catalogue,
over include/sif/io/ of libsif_amd.so. Files are byte-compatible with the
reference's, both ways.

    pysif.io.write_field("box.xfield", 1000.0, x, y, z, weights=w)
    field = pysif.io.read_field("box.xfield")
    header = pysif.io.read_field_header("box.xfield")   # box_length lives here
    pysif.io.write_catalogue_ascii("voids.txt", cat)
    cat = pysif.io.read_catalogue_ascii("other.txt", format="* x y z * r")
"""
import ctypes as C
import os

import numpy as np

from .. import _native as N
from . import Catalogue, Field, Grid, L, _Catalogue, _Field, _Grid, _check, f64, i32, real


class XFieldHeader(C.Structure):  # field_io.h:51-84
    _fields_ = [("magic", C.c_char * 4), ("version", C.c_uint32), ("n_particles", C.c_uint64),
                ("box_length", C.c_double), ("has_weights", C.c_uint32),
                ("has_velocities", C.c_uint32), ("is_double", C.c_uint32),
                ("crc32", C.c_uint32), ("padding", C.c_char * 24)]


class XGridHeader(C.Structure):  # grid_io.h:35-84
    _fields_ = [("magic", C.c_char * 4), ("version", C.c_uint32), ("n_cells", C.c_uint32),
                ("is_double", C.c_uint32), ("total_cells", C.c_uint64),
                ("box_length", C.c_double), ("crc32", C.c_uint32), ("content", C.c_uint32),
                ("source_key", C.c_uint64 * 2), ("assignment", C.c_uint32),
                ("padding", C.c_char * 4)]


assert C.sizeof(XFieldHeader) == 64 and C.sizeof(XGridHeader) == 64


def _sig(name, res, *args):
    fn = getattr(L, name)
    fn.restype = res
    fn.argtypes = list(args)


_sig("sif_field_write", i32, C.c_char_p, C.POINTER(_Field), f64)
_sig("sif_field_read", C.POINTER(_Field), C.c_char_p, C.POINTER(f64))
_sig("sif_field_read_into", i32, C.POINTER(_Field), C.c_char_p)
_sig("sif_grid_write", i32, C.c_char_p, C.POINTER(_Grid))
_sig("sif_grid_read", C.POINTER(_Grid), C.c_char_p)
_sig("sif_grid_read_into", i32, C.POINTER(_Grid), C.c_char_p)
_sig("sif_catalogue_write_ascii", i32, C.c_char_p, C.POINTER(_Catalogue))
_sig("sif_catalogue_read_ascii", C.POINTER(_Catalogue), C.c_char_p, C.c_char_p)

_CONTENT = {0: "unknown", 1: "density", 2: "contrast"}  # grid.h: sif_grid_content_t


def _path(p):
    return os.fsencode(os.fspath(p))


def _arr(a):
    a = np.ascontiguousarray(a, dtype=real)
    if a.ndim != 1:
        raise ValueError("write_field: arrays must be one-dimensional")
    return a


def write_field(filepath, box_length, x, y, z, vx=None, vy=None, vz=None, weights=None):
    """Writes particle arrays to an .xfield binary: positions, the velocity
    blocks when all three are given, the weights when given, and a checksum
    of the payload in the header (py_io.c:29-112)."""
    cols = [_arr(a) for a in (x, y, z)]
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("write_field: x, y and z differ in length")
    f = _Field()
    f.n_particles = n
    f.x, f.y, f.z = [c.ctypes.data_as(C.POINTER(C.c_float)) for c in cols]
    if vx is not None and vy is not None and vz is not None:
        v = [_arr(a) for a in (vx, vy, vz)]
        if any(len(c) != n for c in v):
            raise ValueError("write_field: velocities differ in length from the positions")
        f.vx, f.vy, f.vz = [c.ctypes.data_as(C.POINTER(C.c_float)) for c in v]
    if weights is not None:
        w = _arr(weights)
        if len(w) != n:
            raise ValueError("write_field: weights differ in length from the positions")
        f.weights = w.ctypes.data_as(C.POINTER(C.c_float))
    _check(L.sif_field_write(_path(filepath), C.byref(f), float(box_length)))


def read_field(filepath, wrap=False):
    """Reads an .xfield binary into a Field. With wrap, every coordinate is
    folded into [0, box_length) of the box the file declares (py_io.c:
    114-189). OSError for a missing, malformed, wrong-precision or corrupt
    file."""
    box = f64()
    ptr = L.sif_field_read(_path(filepath), C.byref(box))
    if not ptr:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_IO)
    field = Field.__new__(Field)
    field._p = ptr
    if wrap:
        nonfinite = C.c_uint64()
        if L.sif_field_wrap_periodic(ptr, box.value, None, None, C.byref(nonfinite)) != 0:
            raise ValueError("wrap=True, but %s declares a box length of %g, which cannot be "
                             "wrapped into" % (os.fspath(filepath), box.value))
    return field


def read_field_header(filepath):
    """The 64-byte .xfield header alone: n_particles, box_length,
    has_weights, has_velocities, version (py_io.c:191-227)."""
    h = XFieldHeader()
    try:
        with open(filepath, "rb") as fh:
            raw = fh.read(C.sizeof(h))
    except OSError as e:
        raise OSError("Failed to open %s" % os.fspath(filepath)) from e
    if len(raw) != C.sizeof(h):
        raise OSError("%s is too short to hold an .xfield header" % os.fspath(filepath))
    C.memmove(C.byref(h), raw, C.sizeof(h))
    if h.magic[:4] != b"XFLD":
        raise ValueError("%s is not an .xfield file" % os.fspath(filepath))
    return {"n_particles": int(h.n_particles), "box_length": float(h.box_length),
            "has_weights": bool(h.has_weights), "has_velocities": bool(h.has_velocities),
            "version": int(h.version)}


def write_grid(filepath, grid):
    """Writes a grid to an .xgrid binary, with what its cells hold and a
    checksum of the values."""
    _check(L.sif_grid_write(_path(filepath), grid._p))


def read_grid(filepath):
    """Reads an .xgrid binary into a Grid tagged with what its cells hold."""
    ptr = L.sif_grid_read(_path(filepath))
    if not ptr:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_IO)
    g = Grid.__new__(Grid)
    g._p = ptr
    return g


def grid_content(grid):
    """What a grid's cells hold: 'density', 'contrast' or 'unknown'."""
    return _CONTENT.get(int(grid._p.contents.content), "unknown")


def write_catalogue_ascii(filepath, catalogue):
    """Writes a catalogue as text: '#n=<count>', a line naming the columns
    ('#cx cy cz r' or '#ra dec z r', plus ' footprint footprint_shell' when
    carried), then one row per void with enough digits to read back exactly."""
    _check(L.sif_catalogue_write_ascii(_path(filepath), catalogue._p))


def read_catalogue_ascii(filepath, format=None):
    """Reads a text catalogue: the columns from the file's header names, or
    from `format` ('x y z r', 'ra dec z r', '*' for a skipped column, blanks
    and commas optional) when given."""
    fmt = format.encode() if format is not None else None
    ptr = L.sif_catalogue_read_ascii(_path(filepath), fmt)
    if not ptr:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_IO)
    return Catalogue(_ptr=ptr)


__all__ = ["write_field", "read_field", "read_field_header", "write_grid", "read_grid",
           "grid_content", "write_catalogue_ascii", "read_catalogue_ascii", "XFieldHeader",
           "XGridHeader"]
