"""ctypes binding of libsif_amd.so (the C ABI of include/sif/ and
include/sif_amd.h) and a small device-array helper.

The library is loaded from sif_amd/lib/ in this tree; there is no fallback:
if it is missing, importing this module raises, so nothing can silently run
a different implementation.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsif_amd.so")

SIF_OK = 0
SIF_ERR_ALLOC = -1
SIF_ERR_INVALID = -2
SIF_ERR_RANGE = -3
SIF_ERR_IO = -4
SIF_ERR_UNSUPPORTED = -5
SIF_ERR_UNINITIALISED = -6

FILTER_NONE, FILTER_TOP_HAT, FILTER_GAUSSIAN = 0, 1, 2
N_STAGES = 15
STAGE_NAMES = ("cic", "contrast", "mesh", "fft_forward", "smooth", "scan",
               "prefilter", "grow", "commit", "retest", "sort", "host", "profiles",
               "commit_order", "commit_index")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m sif_amd.build` "
        "(the HIP library is the only implementation of this package)")

# torch ships its own HIP runtime under the same soname as the system's. A
# library loaded after torch binds to torch's runtime, and the two share the
# device; torch imported after the library would load a second runtime, which
# then sees no device. sif_amd.autograd checks this.
TORCH_FIRST = "torch" in sys.modules

lib = C.CDLL(LIB_PATH)

P = C.c_void_p
u32, u64, f32, f64, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_int


class Catalogue(C.Structure):
    _fields_ = [("_block", P), ("cx", C.POINTER(f32)), ("cy", C.POINTER(f32)),
                ("cz", C.POINTER(f32)), ("radii", C.POINTER(f32)),
                ("_footprint_block", P), ("footprint", P), ("footprint_shell", P),
                ("units", i32), ("_meta", P), ("_n_meta", u32),
                ("n_voids", u64), ("capacity", u64)]


class RungStats(C.Structure):
    _fields_ = [("radius", f32), ("candidates", u64), ("masked", u64), ("proxy", u64),
                ("overlap", u64), ("rescale_fail", u64), ("accepted", u64),
                ("evaluated", u64)]


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("sif_error_message", C.c_char_p)
_sig("sif_error_status", i32)
_sig("sif_error_clear", None)
_sig("sif_system_init", i32, P)
_sig("sif_system_finalise", None)
_sig("sif_system_is_initialised", i32)
_sig("sif_amd_device_count", i32)
_sig("sif_amd_malloc", i32, C.POINTER(P), u64)
_sig("sif_amd_free", i32, P)
_sig("sif_amd_memcpy_h2d", i32, P, P, u64)
_sig("sif_amd_memcpy_d2h", i32, P, P, u64)
_sig("sif_amd_memset", i32, P, i32, u64)
_sig("sif_amd_synchronize", i32)
_sig("sif_amd_device_memory", i32, C.POINTER(u64), C.POINTER(u64))
_sig("sif_amd_copy", i32, P, P, u64)
_sig("sif_amd_timing_enable", i32, i32)
_sig("sif_amd_timing_reset", None)
_sig("sif_amd_timing_read", i32, C.POINTER(f64), C.POINTER(u64), i32)
_sig("sif_amd_timing_bytes", i32, C.POINTER(f64), i32)
_sig("sif_amd_check_in_box", i32, P, P, P, u64, f32,
     C.POINTER(u64), C.POINTER(u64), C.POINTER(u64))
_sig("sif_amd_cic_work_bytes", u64, u32)
_sig("sif_amd_grid_assign_cic", i32, P, P, P, P, u64, u32, f32, P, P)
_sig("sif_amd_grid_mean", i32, P, u64, C.POINTER(f64))
_sig("sif_amd_grid_to_density_contrast", i32, P, u64, C.POINTER(f64))
_sig("sif_amd_chain_mesh_build", i32, P, P, P, P, u64, u32, f32, P, P, P, P, P, P, P,
     C.POINTER(f64))
_sig("sif_amd_spectrum_bytes", u64, u32)
_sig("sif_amd_spectrum_pitch", u32, u32)
_sig("sif_amd_fft_r2c", i32, P, u32, P, P)
_sig("sif_amd_fft_filter_c2r", i32, P, u32, f32, i32, f32, f32, i32, i32, P, P)
_sig("sif_amd_scan_candidates", i32, P, P, u64, f32, P, P, u64, C.POINTER(u64))
_sig("sif_amd_scan_minima", i32, P, P, u32, f32, P, P, u64, C.POINTER(u64))
_sig("sif_amd_mark_spheres", i32, P, u32, f32, P, P, P, P, u32, f32)
_sig("sif_amd_grid_profiles", i32, P, u32, f32, P, P, P, P, u32, f32, u32, P)
_sig("sif_amd_grid_profiles_adjoint", i32, P, u32, f32, P, P, P, P, u32, f32, u32, P)
_sig("sif_amd_cic_adjoint", i32, P, P, P, u64, u32, f32, P, P, P, P)
_sig("sif_amd_squared_residual", i32, P, P, u64, f32, P, C.POINTER(f64))
_sig("sif_amd_find_radius", i32, P, P, P, P, u32, f32, P, P, P, u32, f32, f32, f32, f32,
     P, P, P)
_sig("sif_amd_find_radius_weighted", i32, P, P, P, P, P, P, u32, f32, P, P, P, u32, f32, f32,
     f32, f32, f32, P, P, P)
_sig("sif_amd_exodus_device", C.POINTER(Catalogue), P, u32, f32, P, P, P, P, u32, u64,
     C.POINTER(f32), u32, f32, f32, u32, i32, C.POINTER(C.POINTER(u64)),
     C.POINTER(RungStats))
_sig("sif_amd_exodus_device_weighted", C.POINTER(Catalogue), P, u32, f32, P, P, P, P, P, P,
     f64, u32, u64, C.POINTER(f32), u32, f32, f32, u32, i32, C.POINTER(C.POINTER(u64)),
     C.POINTER(RungStats))
_sig("sif_amd_exodus_fields", i32, P, u32, f32, C.POINTER(f32), u32, u32, P)


class MeshDesc(C.Structure):  # sif_amd_mesh_t
    _fields_ = [("offsets", P), ("x", P), ("y", P), ("z", P), ("weights", P),
                ("cell_weights", P), ("total_weight", f64), ("n_cells", u32),
                ("n_particles", u64)]


_sig("sif_amd_survey_occupancy", i32, P, P, P, u64, u32, f32, P, C.POINTER(C.c_int32),
     C.POINTER(f64), C.POINTER(u64), C.POINTER(u64))
_sig("sif_amd_survey_contrast", i32, P, P, P, f32, u64)
_sig("sif_amd_survey_footprint", i32, P, u32, f32, P, P, P, P, u64, P, P)
_sig("sif_amd_find_radius_survey", i32, C.POINTER(MeshDesc), C.POINTER(MeshDesc), f32, P, P, P,
     u32, f32, f32, f32, P, P, P)
_sig("sif_amd_exodus_survey_device", C.POINTER(Catalogue), P, P, u32, f32, C.POINTER(MeshDesc),
     C.POINTER(MeshDesc), C.POINTER(f32), u32, f32, f32, u32, i32, C.POINTER(C.POINTER(u64)),
     C.POINTER(RungStats))
_sig("sif_amd_set_batch_window", i32, u32, u32)
_sig("sif_amd_set_cell_lists", i32, i32)
_sig("sif_amd_set_void_room", i32, u32)


class Box(C.Structure):  # sif_amd_box_t
    _fields_ = [("d_x", P), ("d_y", P), ("d_z", P), ("n_particles", u64), ("n_cells", u32),
                ("mesh_cells", u32), ("box", f32), ("radii", C.POINTER(f32)), ("n_radii", u32)]


class Tuning(C.Structure):  # sif_amd_tuning_t
    _fields_ = [("workers", u32), ("cic_tile", u32), ("split_waves", u32),
                ("batch_floor", u32), ("batch_ceiling", u32)]


_sig("sif_amd_exodus_batch", i32, C.POINTER(Box), u32, f32, f32, u32, C.POINTER(Tuning),
     C.POINTER(C.POINTER(Catalogue)))
_sig("sif_amd_batch_autotune", i32, C.POINTER(Box), u32, f32, f32, u32, u32,
     C.POINTER(Tuning))
_sig("sif_amd_tracer_profiles", i32, P, P, P, P, P, P, P, P, u32, f32, f64, P, P, P, P, u32,
     f32, u32, u32, P, P)
_sig("sif_catalogue_free", None, C.POINTER(Catalogue))

libc = C.CDLL(None)
libc.free.argtypes = [P]
libc.free.restype = None


class SifError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


def error_for(status):
    msg = (lib.sif_error_message() or b"").decode(errors="replace")
    if status == SIF_ERR_ALLOC:
        return MemoryError(msg)
    if status in (SIF_ERR_INVALID, SIF_ERR_RANGE):
        return ValueError(msg)
    if status == SIF_ERR_IO:
        return OSError(msg)
    return RuntimeError(msg)


def check(status):
    if status != SIF_OK:
        raise error_for(status)
    return status


def has_gpu():
    return lib.sif_amd_device_count() > 0


class DeviceArray:
    """A device buffer of a numpy dtype and shape, freed with the object."""

    def __init__(self, shape, dtype=np.float32, zero=False):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        ptr = P()
        check(lib.sif_amd_malloc(C.byref(ptr), max(self.nbytes, 1)))
        self.ptr = ptr.value
        if zero:
            check(lib.sif_amd_memset(self.ptr, 0, self.nbytes))

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib.sif_amd_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def get(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib.sif_amd_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.sif_amd_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib.sif_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def catalogue_to_numpy(cat_ptr, footprint=False):
    """Copies a sif_catalogue_t* into an (n, 4) array [cx, cy, cz, r] and frees
    it; with `footprint`, also its (n, 2) footprint columns (inside, shell)."""
    if not cat_ptr:
        raise error_for(lib.sif_error_status() or SIF_ERR_INVALID)
    c = cat_ptr.contents
    n = int(c.n_voids)
    out = np.zeros((n, 4), np.float32)
    fp = np.zeros((n, 2), np.float32)
    if n:
        out[:, 0] = np.ctypeslib.as_array(c.cx, (n,))
        out[:, 1] = np.ctypeslib.as_array(c.cy, (n,))
        out[:, 2] = np.ctypeslib.as_array(c.cz, (n,))
        out[:, 3] = np.ctypeslib.as_array(c.radii, (n,))
        if footprint and c.footprint:
            fp[:, 0] = np.ctypeslib.as_array(C.cast(c.footprint, C.POINTER(f32)), (n,))
            fp[:, 1] = np.ctypeslib.as_array(C.cast(c.footprint_shell, C.POINTER(f32)), (n,))
    lib.sif_catalogue_free(cat_ptr)
    return (out, fp) if footprint else out


def timing_read():
    """{stage: (ms, launches, algorithmic bytes)} since the last reset."""
    ms = (f64 * N_STAGES)()
    cnt = (u64 * N_STAGES)()
    by = (f64 * N_STAGES)()
    lib.sif_amd_timing_read(ms, cnt, N_STAGES)
    lib.sif_amd_timing_bytes(by, N_STAGES)
    return {STAGE_NAMES[i]: (ms[i], int(cnt[i]), by[i]) for i in range(N_STAGES)}
