"""pysif.measure for the path (python/pysif/measure.pyi of sif v0.5.0):
profiles() and profiles_suggest_mesh_cells(), over sif_profiles
(include/sif/measure/profiles.h). The binning runs on the GPU.
"""
import ctypes as C

import numpy as np

from .. import _native as N
from . import _Catalogue, _Mesh, _check, _view, real

f32, u32, u64, i32 = C.c_float, C.c_uint32, C.c_uint64, C.c_int
L = N.lib

PBC_OPEN = 1 << 0
PROFILES_DIFFERENTIAL = 1 << 8
PROFILES_VELOCITY_WEIGHTED = 1 << 9


class _Density(C.Structure):  # profiles.h:30-46
    _fields_ = [("n_voids", u64), ("n_bins", u32), ("ext", f32), ("differential", C.c_bool),
                ("r_edges", C.POINTER(f32)), ("profiles", C.POINTER(f32))]


class _Velocity(C.Structure):  # profiles.h:71-83
    _fields_ = [("n_voids", u64), ("n_bins", u32), ("ext", f32),
                ("r_edges", C.POINTER(f32)), ("v_rad", C.POINTER(f32))]


L.sif_profiles.restype = i32
L.sif_profiles.argtypes = [C.POINTER(_Catalogue), C.POINTER(_Mesh), f32, u32, u32,
                           C.POINTER(C.POINTER(_Density)), C.POINTER(C.POINTER(_Velocity))]
L.sif_density_profiles_free.restype = None
L.sif_density_profiles_free.argtypes = [C.POINTER(_Density)]
L.sif_velocity_profiles_free.restype = None
L.sif_velocity_profiles_free.argtypes = [C.POINTER(_Velocity)]


class Profiles:
    """Stacked radial profiles of voids: density and radial velocity, one row
    a void, binned in radius over each void's own radius
    (measure.pyi:250-290). Returned by profiles(); not constructed directly."""

    def __init__(self, _dens=None, _vel=None):
        self._d = _dens
        self._v = _vel

    def __del__(self):
        if getattr(self, "_d", None):
            L.sif_density_profiles_free(self._d)
            self._d = None
        if getattr(self, "_v", None):
            L.sif_velocity_profiles_free(self._v)
            self._v = None

    def _any(self):
        return (self._d or self._v).contents

    has_density = property(lambda self: bool(self._d))
    has_velocity = property(lambda self: bool(self._v))
    n_voids = property(lambda self: int(self._any().n_voids))
    n_bins = property(lambda self: int(self._any().n_bins))
    ext = property(lambda self: float(self._any().ext))

    @property
    def differential(self):
        return bool(self._d.contents.differential) if self._d else None

    @property
    def r_edges(self):
        s = self._any()
        return _view(self, s.r_edges, (s.n_bins + 1,))

    @property
    def r_bins(self):
        """profiles.h:48-69: the outer edge of a cumulative bin, the middle
        of a differential one; None without densities."""
        if not self._d:
            return None
        e = self.r_edges
        if self.differential:
            return ((e[:-1] + e[1:]) * real(0.5)).astype(real)
        return e[1:].copy()

    @property
    def density(self):
        if not self._d:
            return None
        s = self._d.contents
        return _view(self, s.profiles, (s.n_voids, s.n_bins))

    @property
    def velocity(self):
        if not self._v:
            return None
        s = self._v.contents
        return _view(self, s.v_rad, (s.n_voids, s.n_bins))


def profiles(catalogue, mesh, n_bins, ext=5.0, compute_velocity=False, use_pbc=True,
             differential=False, weighted_velocity=False):
    """Stack radial density (and velocity) profiles around voids
    (measure.pyi:112-147)."""
    if int(n_bins) <= 0:
        raise ValueError("n_bins must be positive")
    opt = (0 if use_pbc else PBC_OPEN) | (PROFILES_DIFFERENTIAL if differential else 0) | \
        (PROFILES_VELOCITY_WEIGHTED if weighted_velocity else 0)
    d = C.POINTER(_Density)()
    v = C.POINTER(_Velocity)()
    _check(L.sif_profiles(catalogue._p, mesh._p, float(ext), int(n_bins), opt, C.byref(d),
                          C.byref(v) if compute_velocity else None))
    return Profiles(d, v if compute_velocity else None)


def profiles_suggest_mesh_cells(n_particles):
    """profiles.h:228-241: about 36 tracers a cell, at most 256 a side."""
    n = int(n_particles)
    if n <= 0:
        return 1
    c = float(np.cbrt(n / 36.0))
    if c >= 256:
        return 256
    return int(c) if c >= 1.0 else 1
