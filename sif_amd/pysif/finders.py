"""pysif.finders for the exodus path (python/pysif/finders.pyi of sif
v0.5.0): exodus(), exodus_raw(), exodus_survey(), exodus_survey_raw(),
survey_box(), read_checkpoint(), suggest_mesh_cells() and
grid_cells_from_size(), over sif_finder_exodus(_raw),
sif_finder_exodus_survey(_raw), sif_finder_exodus_survey_box,
sif_finder_exodus_checkpoint_read and sif_finder_grid_cells_from_size.
"""
import ctypes as C
import math
import os

import numpy as np

from .. import _native as N
from . import Catalogue, ChainMesh, Grid, _Field, _Grid, _Mesh, _Catalogue, _check, real

P = C.c_void_p
f32, f64, u32, u64, i32 = C.c_float, C.c_double, C.c_uint32, C.c_uint64, C.c_int
L = N.lib

# macros.h finder options
SEARCH = {1.5: 0 << 12, 1.25: 1 << 12, 1.75: 2 << 12, 2.0: 3 << 12}
ALIASING_WINDOW = 1 << 8
SEEDS_MINIMA = 1 << 9
NO_CHECKPOINT = 1 << 11
KEEP_CIC_WINDOW = 1 << 14
CONSUME_GRID = 1 << 15
LAGRANGIAN_OVERLAP = 1 << 18
REFINE = {"none": 0 << 19, None: 0 << 19, "gaussian": 1 << 19, "tracers": 2 << 19,
          "hill": 3 << 19}
EXTEND_REACH = 1 << 21
DISPLACE = 1 << 22
VARIANTS = {"excursion-set": 0, "excursion-peak": 1, "lagrangian-overlap": 2}


class _Extras(C.Structure):  # exodus_finder.h:40-106
    _fields_ = [("checkpoint", C.c_char_p), ("prior", C.POINTER(_Catalogue)),
                ("keep_checkpoint", i32), ("save_memory", i32)]


class _Checkpoint(C.Structure):  # exodus_finder.h:427-531
    _fields_ = [("magic", C.c_char * 4), ("version", u32), ("byte_order", u32),
                ("is_double", u32), ("is_survey", u32), ("is_weighted", u32),
                ("options", u32), ("n_radii", u32), ("next_rung", u32),
                ("grid_n_cells", u32), ("box_length", f64), ("threshold", f64),
                ("overlap_fraction", f64), ("n_particles", u64), ("total_weight", f64),
                ("random_n_particles", u64), ("random_total_weight", f64),
                ("offset", f64 * 3), ("n_prior", u64), ("prior_crc32", u32),
                ("crc32", u32), ("n_voids", u64), ("n_mismatched", u64),
                ("assignment", u32), ("reserved", C.c_char * 36)]


assert C.sizeof(_Checkpoint) == 192

L.sif_finder_exodus.restype = C.POINTER(_Catalogue)
L.sif_finder_exodus.argtypes = [C.POINTER(_Grid), C.POINTER(_Mesh), C.POINTER(f32), u32, f32,
                                i32, C.POINTER(_Extras)]
L.sif_finder_exodus_raw.restype = C.POINTER(_Catalogue)
L.sif_finder_exodus_raw.argtypes = [C.POINTER(_Grid), C.POINTER(_Mesh), C.POINTER(f32), u32,
                                    f32, f32, u32, C.POINTER(_Extras)]
L.sif_finder_exodus_checkpoint_read.restype = C.POINTER(_Catalogue)
L.sif_finder_exodus_checkpoint_read.argtypes = [C.c_char_p, C.POINTER(_Checkpoint)]
L.sif_finder_grid_cells_from_size.restype = u32
L.sif_finder_grid_cells_from_size.argtypes = [u64, f32, f32, i32, f32]


def _path(p):
    return os.fsencode(p) if p is not None else None


def _extras(checkpoint, prior, keep_checkpoint, save_memory):
    e = _Extras()
    e.checkpoint = _path(checkpoint) if checkpoint not in (None, False) else None
    e.prior = prior._p if prior is not None else None
    e.keep_checkpoint = 1 if keep_checkpoint else 0
    e.save_memory = 1 if save_memory else 0
    return e


def _radii(radii):
    r = np.ascontiguousarray(radii, dtype=real)
    if r.ndim != 1 or len(r) == 0:
        raise ValueError("radii: a non-empty one-dimensional ladder is required")
    return r


def _catalogue(ptr):
    if not ptr:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_INVALID)
    return Catalogue(_ptr=ptr)


def snap_search_factor(search_factor):
    """py_finders.c:163-185: the nearest of 1.25, 1.5, 1.75, 2.0."""
    f = float(search_factor)
    if not math.isfinite(f) or f <= 1.0:
        raise ValueError(f"search_factor: {search_factor!r}, expected a finite value above 1")
    return min(SEARCH, key=lambda s: (abs(s - f), s))


def exodus(grid, mesh, radii, threshold, variant="excursion-set", *, checkpoint=None,
           prior=None, keep_checkpoint=False, save_memory=False):
    """Find voids, growing each to the radius its tracers support."""
    if variant not in VARIANTS:
        raise ValueError(f"variant: {variant!r}, expected one of {sorted(VARIANTS)}")
    r = _radii(radii)
    ex = _extras(checkpoint, prior, keep_checkpoint, save_memory)
    ptr = L.sif_finder_exodus(grid._p, mesh._p, r.ctypes.data_as(C.POINTER(f32)), len(r),
                              float(threshold), VARIANTS[variant], C.byref(ex))
    return _catalogue(ptr)


def exodus_raw(grid, mesh, radii, threshold, overlap_fraction=0.0, *, overlap_type="eulerian",
               seeds="cells", refine_centres="none", extend_reach=False, displace=False,
               search_factor=1.5, deconvolution="full", consume_grid=False, checkpoint=None,
               prior=None, keep_checkpoint=False, save_memory=False):
    """exodus(), with every setting a variant fixes taken one by one."""
    if overlap_type not in ("eulerian", "lagrangian"):
        raise ValueError(f"overlap_type: {overlap_type!r}")
    if seeds not in ("cells", "minima"):
        raise ValueError(f"seeds: {seeds!r}")
    if refine_centres not in REFINE:
        raise ValueError(f"refine_centres: {refine_centres!r}")
    if deconvolution not in ("full", "aliasing", "none"):
        raise ValueError(f"deconvolution: {deconvolution!r}")
    opt = SEARCH[snap_search_factor(search_factor)]
    opt |= LAGRANGIAN_OVERLAP if overlap_type == "lagrangian" else 0
    opt |= SEEDS_MINIMA if seeds == "minima" else 0
    opt |= REFINE[refine_centres]
    opt |= EXTEND_REACH if extend_reach else 0
    opt |= DISPLACE if displace else 0
    opt |= {"full": 0, "aliasing": ALIASING_WINDOW, "none": KEEP_CIC_WINDOW}[deconvolution]
    opt |= CONSUME_GRID if consume_grid else 0
    opt |= NO_CHECKPOINT if checkpoint is False else 0
    r = _radii(radii)
    ex = _extras(checkpoint, prior, keep_checkpoint, save_memory)
    ptr = L.sif_finder_exodus_raw(grid._p, mesh._p, r.ctypes.data_as(C.POINTER(f32)), len(r),
                                  float(threshold), float(overlap_fraction), opt, C.byref(ex))
    return _catalogue(ptr)


SURVEY_VARIANTS = ("excursion-set", "lagrangian-overlap")
_RP = C.POINTER(f32)
L.sif_finder_exodus_survey.restype = C.POINTER(_Catalogue)
L.sif_finder_exodus_survey.argtypes = [C.POINTER(_Grid), C.POINTER(_Grid), C.POINTER(_Mesh),
                                       C.POINTER(_Mesh), _RP, _RP, u32, f32, i32,
                                       C.POINTER(_Extras)]
L.sif_finder_exodus_survey_raw.restype = C.POINTER(_Catalogue)
L.sif_finder_exodus_survey_raw.argtypes = [C.POINTER(_Grid), C.POINTER(_Grid),
                                           C.POINTER(_Mesh), C.POINTER(_Mesh), _RP, _RP, u32,
                                           f32, f32, u32, C.POINTER(_Extras)]
L.sif_finder_exodus_survey_box.restype = i32
L.sif_finder_exodus_survey_box.argtypes = [C.POINTER(_Field), C.POINTER(_Field), _RP, u32,
                                           u32, _RP, _RP]


def _offset(offset):
    o = [float(v) for v in offset]
    if len(o) != 3:
        raise ValueError(f"offset: {offset!r}, expected 3 numbers")
    return (f32 * 3)(*o)


def exodus_survey(data_grid, random_grid, data_mesh, random_mesh, offset, radii, threshold,
                  variant="excursion-set", *, checkpoint=None, prior=None,
                  keep_checkpoint=False, save_memory=False):
    """Find voids in a survey whose geometry a random catalogue describes;
    the voids come back in the caller's frame, with their footprint."""
    if variant not in SURVEY_VARIANTS:
        raise ValueError(f"variant: {variant!r}, expected one of {list(SURVEY_VARIANTS)}")
    r = _radii(radii)
    ex = _extras(checkpoint, prior, keep_checkpoint, save_memory)
    ptr = L.sif_finder_exodus_survey(data_grid._p, random_grid._p, data_mesh._p,
                                     random_mesh._p, _offset(offset),
                                     r.ctypes.data_as(_RP), len(r), float(threshold),
                                     VARIANTS[variant], C.byref(ex))
    return _catalogue(ptr)


def exodus_survey_raw(data_grid, random_grid, data_mesh, random_mesh, offset, radii, threshold,
                      overlap_fraction=0.0, *, overlap_type="eulerian", seeds="cells",
                      search_factor=1.5, deconvolution="full", consume_grid=False,
                      checkpoint=None, prior=None, keep_checkpoint=False, save_memory=False):
    """exodus_survey(), every setting taken one by one; seeds='cells' only."""
    if overlap_type not in ("eulerian", "lagrangian"):
        raise ValueError(f"overlap_type: {overlap_type!r}")
    if seeds not in ("cells", "minima"):
        raise ValueError(f"seeds: {seeds!r}")
    if deconvolution not in ("full", "aliasing", "none"):
        raise ValueError(f"deconvolution: {deconvolution!r}")
    opt = SEARCH[snap_search_factor(search_factor)]
    opt |= LAGRANGIAN_OVERLAP if overlap_type == "lagrangian" else 0
    opt |= SEEDS_MINIMA if seeds == "minima" else 0
    opt |= {"full": 0, "aliasing": ALIASING_WINDOW, "none": KEEP_CIC_WINDOW}[deconvolution]
    opt |= CONSUME_GRID if consume_grid else 0
    opt |= NO_CHECKPOINT if checkpoint is False else 0
    r = _radii(radii)
    ex = _extras(checkpoint, prior, keep_checkpoint, save_memory)
    ptr = L.sif_finder_exodus_survey_raw(data_grid._p, random_grid._p, data_mesh._p,
                                         random_mesh._p, _offset(offset),
                                         r.ctypes.data_as(_RP), len(r), float(threshold),
                                         float(overlap_fraction), opt, C.byref(ex))
    return _catalogue(ptr)


def survey_box(data, randoms, radii, search_factor=1.5):
    """(offset, box_length): the box a survey has to be searched in and the
    offset that centres it there."""
    r = _radii(radii)
    off = (f32 * 3)()
    box = f32()
    _check(L.sif_finder_exodus_survey_box(data._p, randoms._p, r.ctypes.data_as(_RP), len(r),
                                          SEARCH[snap_search_factor(search_factor)], off,
                                          C.byref(box)))
    return np.array(list(off), dtype=real), float(box.value)


def read_checkpoint(path):
    """(Catalogue, dict) of an exodus checkpoint (finders.pyi:315-345)."""
    h = _Checkpoint()
    ptr = L.sif_finder_exodus_checkpoint_read(_path(path), C.byref(h))
    cat = _catalogue(ptr)
    with open(path, "rb") as f:  # the ladder follows the header
        f.seek(C.sizeof(_Checkpoint))
        radii = np.frombuffer(f.read(4 * h.n_radii), dtype=real).copy()
    search = {v: k for k, v in SEARCH.items()}[h.options & (3 << 12)]
    refine = {v: k for k, v in REFINE.items() if k is not None}[h.options & (3 << 19)]
    info = {
        "mode": "survey" if h.is_survey else "box", "weighted": bool(h.is_weighted),
        "radii": radii, "rungs_done": int(h.next_rung), "n_voids": int(h.n_voids),
        "n_prior": int(h.n_prior), "threshold": float(np.float32(h.threshold)),
        "overlap_fraction": float(np.float32(h.overlap_fraction)),
        "overlap_type": "lagrangian" if h.options & LAGRANGIAN_OVERLAP else "eulerian",
        "search_factor": search, "keep_cic_window": bool(h.options & KEEP_CIC_WINDOW),
        "deconvolution": ("none" if h.options & KEEP_CIC_WINDOW else
                          "aliasing" if h.options & ALIASING_WINDOW else "full"),
        "seeds": "minima" if h.options & SEEDS_MINIMA else "cells", "refine_centres": refine,
        "extend_reach": bool(h.options & EXTEND_REACH),
        "displace": bool(h.options & DISPLACE), "grid_n_cells": int(h.grid_n_cells),
        "box_length": float(np.float32(h.box_length)),
        "assignment": {0: None, 1: "ngp", 2: "cic", 3: "tsc", 4: "pcs"}.get(
            h.assignment & 0xff), "interlaced": bool(h.assignment >> 8),
        "n_particles": int(h.n_particles), "total_weight": float(h.total_weight),
        "n_randoms": int(h.random_n_particles),
        "random_total_weight": float(h.random_total_weight),
        "offset": tuple(float(np.float32(v)) for v in h.offset),
    }
    return cat, info


def suggest_mesh_cells(n_particles, box_length, max_radius=0.0, volume=None):
    """exodus_finder.h:654-685: about 30 tracers a cell, floored so the widest
    search sphere fits, in [8, 1024], truncated to an integer."""
    vol = float(np.float32(0.0 if volume is None else volume))
    side = float(np.float32(box_length))
    if n_particles <= 0 or not (side > 0) or not math.isfinite(side) or not (vol >= 0) or \
            not math.isfinite(vol):
        raise ValueError("suggest_mesh_cells: tracers, box and volume must be positive and "
                         "finite")
    as_box = n_particles * side * side * side / vol if vol > 0 else float(n_particles)
    cells = float(np.cbrt(as_box / 30.0))
    r = float(np.float32(max_radius))
    if r > 0:
        reach = 2.0 * r
        if reach >= side:
            raise ValueError("suggest_mesh_cells: the search sphere is wider than the box")
        cells = max(cells, side / (side - reach) + 1.0)
    return int(min(max(cells, 8.0), 1024.0))


def grid_cells_from_size(n_particles, box_length, cell_size, units="mps", volume=None):
    """sif_finder_grid_cells_from_size (exodus_finder.c:2504-2560)."""
    u = {"physical": 0, "mps": 1}.get(units)  # sif_finder_units_t
    if u is None:
        raise ValueError(f"units: {units!r}, expected 'mps' or 'physical'")
    n = L.sif_finder_grid_cells_from_size(int(n_particles), float(box_length), float(cell_size),
                                          u, 0.0 if volume is None else float(volume))
    if n == 0:
        raise N.error_for(L.sif_error_status() or N.SIF_ERR_INVALID)
    return int(n)
