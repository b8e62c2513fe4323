"""The pysif surface for the exodus path (python/pysif/__init__.pyi and
finders.pyi of sif v0.5.0): Field, Grid, ChainMesh and Catalogue over the
host-struct C API of include/sif/ (libsif_amd.so), with pysif's names,
argument meanings and exceptions (py_common.h:95-131: MemoryError for
SIF_ERR_ALLOC, ValueError for INVALID and RANGE, RuntimeError for
UNSUPPORTED and UNINITIALISED, OSError for IO).

Array properties are read-only NumPy views of the library's own buffers,
and keep their owner alive while they exist (py_common.h:59-74). Everything
computes in the HIP library; there is no other implementation behind it.

    import sif_amd.pysif as pysif
    pysif.system_init()
    field = pysif.field_from_numpy(x, y, z)
    grid = pysif.Grid(512, 1000.0)
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(100, 1000.0, field, drop_indices=True)
    voids = pysif.finders.exodus(grid, mesh, radii, -0.8, checkpoint=False)
"""
import ctypes as C

import numpy as np

from .. import _native as N

P = C.c_void_p
f32, f64, u32, u64, i32 = C.c_float, C.c_double, C.c_uint32, C.c_uint64, C.c_int
real = np.float32  # sif_real (macros.h:504-534)


class _Field(C.Structure):  # field.h:90-120
    _fields_ = [("_position_block", P), ("_velocity_block", P),
                ("x", C.POINTER(f32)), ("y", C.POINTER(f32)), ("z", C.POINTER(f32)),
                ("vx", C.POINTER(f32)), ("vy", C.POINTER(f32)), ("vz", C.POINTER(f32)),
                ("weights", C.POINTER(f32)), ("original_indices", C.POINTER(u64)),
                ("min_p", f32 * 3), ("max_p", f32 * 3), ("centre", f32 * 3),
                ("half_span", f32), ("state_flags", u32), ("n_particles", u64),
                ("units", i32)]


class _Grid(C.Structure):  # grid.h:85-110
    _fields_ = [("values", C.POINTER(f32)), ("content", i32), ("assignment", i32),
                ("interlaced", u32), ("n_cells", u32), ("total_cells", u64),
                ("box_length", f32), ("cell_length", f32), ("p2_mask", u32),
                ("p2_shift", u32)]


class _Mesh(C.Structure):  # chain_mesh.h:40-108
    _fields_ = [("n_cells", u32), ("total_cells", u64), ("box_length", f32),
                ("cell_length", f32), ("_position_block", P), ("_velocity_block", P),
                ("x", C.POINTER(f32)), ("y", C.POINTER(f32)), ("z", C.POINTER(f32)),
                ("vx", C.POINTER(f32)), ("vy", C.POINTER(f32)), ("vz", C.POINTER(f32)),
                ("weights", C.POINTER(f32)), ("total_weight", f64),
                ("cell_weights", C.POINTER(f32)), ("original_indices", C.POINTER(u64)),
                ("n_particles", u64), ("cell_offsets", C.POINTER(u64))]


_Catalogue = N.Catalogue

L = N.lib


def _sig(name, res, *args):
    fn = getattr(L, name)
    fn.restype = res
    fn.argtypes = list(args)


_sig("sif_field_alloc", C.POINTER(_Field), u64)
_sig("sif_field_free", None, C.POINTER(_Field))
_sig("sif_field_reserve_positions", i32, C.POINTER(_Field))
_sig("sif_field_reserve_velocities", i32, C.POINTER(_Field))
_sig("sif_field_reserve_weights", i32, C.POINTER(_Field))
_sig("sif_field_assign_positions", i32, C.POINTER(_Field), P, P, P)
_sig("sif_field_assign_velocities", i32, C.POINTER(_Field), P, P, P)
_sig("sif_field_assign_weights", i32, C.POINTER(_Field), P)
_sig("sif_field_wrap_periodic", i32, C.POINTER(_Field), f32, C.POINTER(u64),
     C.POINTER(u64), C.POINTER(u64))
_sig("sif_field_translate", i32, C.POINTER(_Field), C.POINTER(f32))
_sig("sif_grid_alloc", C.POINTER(_Grid), u32, f32)
_sig("sif_grid_free", None, C.POINTER(_Grid))
_sig("sif_grid_assign", i32, C.POINTER(_Grid), C.POINTER(_Field), i32, u32)
_sig("sif_grid_assign_cic", i32, C.POINTER(_Grid), C.POINTER(_Field))
_sig("sif_grid_to_density_contrast", i32, C.POINTER(_Grid))
_sig("sif_chain_mesh_alloc", C.POINTER(_Mesh), u32, f32, C.POINTER(_Field), u32)
_sig("sif_chain_mesh_alloc_consume", C.POINTER(_Mesh), u32, f32, C.POINTER(_Field), u32)
_sig("sif_chain_mesh_free", None, C.POINTER(_Mesh))
_sig("sif_catalogue_alloc", C.POINTER(_Catalogue), u64)
_sig("sif_catalogue_append", i32, C.POINTER(_Catalogue), f32, f32, f32, f32)
_sig("sif_catalogue_trim", i32, C.POINTER(_Catalogue))
_sig("sif_catalogue_reserve_footprint", i32, C.POINTER(_Catalogue))

SIF_COORDINATES_SKY = 1  # macros.h: sif_coordinates_t
SIF_MESH_DROP_INDICES = 1 << 16
SIF_MESH_NO_CANONICAL = 1 << 17
_ASSIGNMENTS = {0: None, 1: "ngp", 2: "cic", 3: "tsc", 4: "pcs"}


def _check(status):
    return N.check(status)


def _null_check(ptr):
    if not ptr:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_INVALID)
    return ptr


class _View:
    """A read-only NumPy view of a library buffer that keeps its owner
    alive (np.asarray of it has this object as its base)."""

    def __init__(self, owner, ptr, shape, dtype):
        self.owner = owner
        addr = C.cast(ptr, C.c_void_p).value or 0
        self.__array_interface__ = {"shape": tuple(shape), "typestr": np.dtype(dtype).str,
                                    "data": (addr, True), "version": 3}


def _view(owner, ptr, shape, dtype=np.float32):
    if not ptr:
        return None
    a = np.asarray(_View(owner, ptr, shape, dtype))
    a.flags.writeable = False
    return a


def system_init(threads=0, skip_tuning=False):
    """Opens the HIP device (system.h). threads and skip_tuning are kept for
    pysif's signature; the device sets its own parallelism."""
    _check(L.sif_system_init(None))


def system_finalise():
    L.sif_system_finalise()


init, finalise = system_init, system_finalise


class Field:
    """Particles: positions, optional velocities and weights (field.h)."""

    def __init__(self, capacity=0):
        self._p = _null_check(L.sif_field_alloc(int(capacity)))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            L.sif_field_free(p)
            self._p = None

    @property
    def _s(self):
        if not self._p:
            raise ValueError("Field: consumed by a ChainMesh")
        return self._p.contents

    @property
    def n_particles(self):
        return int(self._s.n_particles)

    @property
    def has_weights(self):
        return bool(self._s.weights)

    @property
    def has_velocities(self):
        return bool(self._s.vx)

    x = property(lambda self: _view(self, self._s.x, (self.n_particles,)))
    y = property(lambda self: _view(self, self._s.y, (self.n_particles,)))
    z = property(lambda self: _view(self, self._s.z, (self.n_particles,)))
    vx = property(lambda self: _view(self, self._s.vx, (self.n_particles,)))
    vy = property(lambda self: _view(self, self._s.vy, (self.n_particles,)))
    vz = property(lambda self: _view(self, self._s.vz, (self.n_particles,)))
    weights = property(lambda self: _view(self, self._s.weights, (self.n_particles,)))

    def to_numpy(self):
        """The positions as an (n, 3) array (a copy)."""
        return np.stack([self.x, self.y, self.z], axis=1)

    def wrap(self, box_length):
        """Folds positions into [0, box) (field.h:230-260)."""
        b, w, nf = u64(), u64(), u64()
        _check(L.sif_field_wrap_periodic(self._p, float(box_length), C.byref(b), C.byref(w),
                                         C.byref(nf)))
        return {"boundary": b.value, "wrapped": w.value, "nonfinite": nf.value}

    def translate(self, offset):
        off = (f32 * 3)(*[float(v) for v in offset])
        _check(L.sif_field_translate(self._p, off))


def _col(a, n=None):
    a = np.ascontiguousarray(a, dtype=real)
    if a.ndim != 1 or (n is not None and len(a) != n):
        raise ValueError("arrays must be one-dimensional and of one length")
    return a


def field_from_numpy(x=None, y=None, z=None, *, ra=None, dec=None, vx=None, vy=None, vz=None,
                     weights=None):
    """A new Field with its particles copied out of NumPy arrays."""
    if ra is not None or dec is not None:
        raise RuntimeError("field_from_numpy: sky coordinates are outside the exodus box path")
    if x is None or y is None or z is None:
        raise ValueError("field_from_numpy: x, y and z are required")
    x = _col(x)
    y, z = _col(y, len(x)), _col(z, len(x))
    vel = [v is not None for v in (vx, vy, vz)]
    if any(vel) and not all(vel):
        raise ValueError("field_from_numpy: velocities are given all three or none")
    f = Field(len(x))
    _check(L.sif_field_reserve_positions(f._p))
    _check(L.sif_field_assign_positions(f._p, x.ctypes.data, y.ctypes.data, z.ctypes.data))
    if all(vel):
        v = [_col(a, len(x)) for a in (vx, vy, vz)]
        _check(L.sif_field_reserve_velocities(f._p))
        _check(L.sif_field_assign_velocities(f._p, *[a.ctypes.data for a in v]))
    if weights is not None:
        w = _col(weights, len(x))
        _check(L.sif_field_reserve_weights(f._p))
        _check(L.sif_field_assign_weights(f._p, w.ctypes.data))
    return f


class Grid:
    """A cubic grid over a periodic box (grid.h)."""

    def __init__(self, n_cells, box_length):
        self._p = _null_check(L.sif_grid_alloc(int(n_cells), float(box_length)))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            L.sif_grid_free(p)
            self._p = None

    n_cells = property(lambda self: int(self._p.contents.n_cells))
    box_length = property(lambda self: float(self._p.contents.box_length))
    cell_length = property(lambda self: float(self._p.contents.cell_length))
    interlaced = property(lambda self: bool(self._p.contents.interlaced))

    @property
    def assignment(self):
        return _ASSIGNMENTS.get(int(self._p.contents.assignment))

    @property
    def values(self):
        n = self.n_cells
        return _view(self, self._p.contents.values, (n, n, n))

    def assign_cic(self, field):
        _check(L.sif_grid_assign_cic(self._p, field._p))

    def assign(self, field, scheme="cic", interlace=False):
        if scheme != "cic" or interlace:
            raise RuntimeError("Grid.assign: only CIC without interlacing is on this path")
        self.assign_cic(field)

    def to_density_contrast(self):
        _check(L.sif_grid_to_density_contrast(self._p))


class ChainMesh:
    """Particles sorted into cubic cells, in the canonical order
    (chain_mesh.h)."""

    def __init__(self, n_cells, box_length, field, drop_indices=False, consume_field=False,
                 canonical=True):
        opt = (SIF_MESH_DROP_INDICES if drop_indices else 0) | \
              (0 if canonical else SIF_MESH_NO_CANONICAL)
        if consume_field:
            self._p = _null_check(L.sif_chain_mesh_alloc_consume(int(n_cells), float(box_length),
                                                                 field._p, opt))
            field._p = None
        else:
            self._p = _null_check(L.sif_chain_mesh_alloc(int(n_cells), float(box_length),
                                                         field._p, opt))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            L.sif_chain_mesh_free(p)
            self._p = None

    _s = property(lambda self: self._p.contents)
    n_cells = property(lambda self: int(self._s.n_cells))
    total_cells = property(lambda self: int(self._s.total_cells))
    box_length = property(lambda self: float(self._s.box_length))
    cell_length = property(lambda self: float(self._s.cell_length))
    n_particles = property(lambda self: int(self._s.n_particles))
    x = property(lambda self: _view(self, self._s.x, (self.n_particles,)))
    y = property(lambda self: _view(self, self._s.y, (self.n_particles,)))
    z = property(lambda self: _view(self, self._s.z, (self.n_particles,)))
    vx = property(lambda self: _view(self, self._s.vx, (self.n_particles,)))
    vy = property(lambda self: _view(self, self._s.vy, (self.n_particles,)))
    vz = property(lambda self: _view(self, self._s.vz, (self.n_particles,)))
    weights = property(lambda self: _view(self, self._s.weights, (self.n_particles,)))
    original_indices = property(lambda self: _view(self, self._s.original_indices,
                                                   (self.n_particles,), np.uint64))
    cell_offsets = property(lambda self: _view(self, self._s.cell_offsets,
                                               (self.total_cells + 1,), np.uint64))


class Catalogue:
    """Voids: centres and radii (catalogue.h)."""

    def __init__(self, capacity=0, _ptr=None):
        self._p = _ptr if _ptr else _null_check(L.sif_catalogue_alloc(int(capacity)))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            N.lib.sif_catalogue_free(p)
            self._p = None

    @property
    def n_voids(self):
        return int(self._p.contents.n_voids)

    @property
    def centres(self):
        """(n, 3) array of the centres (a copy: the columns are stored apart)."""
        c = self._p.contents
        n = self.n_voids
        if n == 0:
            return np.zeros((0, 3), real)
        return np.stack([np.ctypeslib.as_array(p, (n,)) for p in (c.cx, c.cy, c.cz)], axis=1)

    @property
    def radii(self):
        return _view(self, self._p.contents.radii, (self.n_voids,))

    @property
    def units(self):
        """'cartesian', or 'sky' for centres as ra, dec (degrees) and redshift."""
        return "sky" if self._p.contents.units == SIF_COORDINATES_SKY else "cartesian"

    @property
    def footprint(self):
        """The fraction of each void's sphere inside a survey footprint, a
        read-only view, or None for a catalogue with none (-1: not measured)."""
        return _view(self, C.cast(self._p.contents.footprint, C.POINTER(f32)), (self.n_voids,))

    @property
    def footprint_shell(self):
        """The same fraction over the shell of one to two radii, or None."""
        return _view(self, C.cast(self._p.contents.footprint_shell, C.POINTER(f32)),
                     (self.n_voids,))

    def to_numpy(self):
        """A structured array, a copy, with the fields named as the file
        headers name them: cx, cy, cz, r (ra, dec, z, r on the sky), and
        footprint, footprint_shell when carried."""
        names = ["ra", "dec", "z"] if self.units == "sky" else ["cx", "cy", "cz"]
        extra = ["footprint", "footprint_shell"] if self._p.contents.footprint else []
        out = np.zeros(self.n_voids, dtype=[(k, real) for k in names + ["r"] + extra])
        cen = self.centres
        for k, name in enumerate(names):
            out[name] = cen[:, k]
        out["r"] = self.radii if self.n_voids else []
        for name in extra:
            out[name] = getattr(self, name) if self.n_voids else []
        return out

    def _append(self, x, y, z, r):
        _check(L.sif_catalogue_append(self._p, float(x), float(y), float(z), float(r)))


def catalogue_from_numpy(cx=None, cy=None, cz=None, r=None, *, ra=None, dec=None, z=None,
                         footprint=None, footprint_shell=None, radii=None):
    """A new Catalogue with its voids copied out of NumPy arrays: cx, cy, cz
    for Cartesian centres or ra, dec, z for a catalogue on the sky; r (or
    radii) the radii; footprint and footprint_shell both or neither."""
    if r is None:
        r = radii
    sky = [a is not None for a in (ra, dec, z)]
    box = [a is not None for a in (cx, cy, cz)]
    if any(sky) and any(box):
        raise ValueError("catalogue_from_numpy: Cartesian and sky names are mixed")
    centre = (ra, dec, z) if any(sky) else (cx, cy, cz)
    if any(a is None for a in centre) or r is None:
        raise ValueError("catalogue_from_numpy: the three centre columns and r are required")
    if (footprint is None) != (footprint_shell is None):
        raise ValueError("catalogue_from_numpy: footprint and footprint_shell go together")
    cols = [_col(centre[0])]
    cols += [_col(a, len(cols[0])) for a in (centre[1], centre[2], r)]
    cat = Catalogue(len(cols[0]))
    for i in range(len(cols[0])):
        cat._append(cols[0][i], cols[1][i], cols[2][i], cols[3][i])
    if any(sky):
        cat._p.contents.units = SIF_COORDINATES_SKY
    if footprint is not None:
        fp = [_col(a, len(cols[0])) for a in (footprint, footprint_shell)]
        _check(L.sif_catalogue_reserve_footprint(cat._p))
        n = len(cols[0])
        for k, name in enumerate(("footprint", "footprint_shell")):
            C.memmove(getattr(cat._p.contents, name), fp[k].ctypes.data, n * 4)
    return cat


from . import finders, io  # noqa: E402  (use the types above)

__all__ = ["system_init", "system_finalise", "init", "finalise", "Field", "field_from_numpy",
           "Grid", "ChainMesh", "Catalogue", "catalogue_from_numpy", "finders", "io", "real"]
