"""TEST INFRASTRUCTURE ONLY -- ctypes binding of the oracle (liboracle.so), the
CPU restatement of the sif v0.5.0 exodus path, and of the reference's own
quicksort harness (_ref/libref_sort.so) when it has been built.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker or the timed CPU baseline. The product (sif_amd/) never imports
it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_SORT_PATH = os.path.join(HERE, "_ref", "libref_sort.so")


def build():
    """Compiles the oracle (and the reference sort harness when the
    reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()

lib = C.CDLL(LIB_PATH)
P = C.c_void_p
u32, u64, f32, f64, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_int


class Candidate(C.Structure):
    _fields_ = [("delta", f32), ("refined", u32), ("flat_idx", u64)]


class Template(C.Structure):
    _fields_ = [("dx", P), ("dy", P), ("dz", P), ("type", P), ("min_d2", P),
                ("max_d2", P), ("count", u32)]


class Mesh(C.Structure):
    _fields_ = [("n_cells", u32), ("box_length", f32), ("cell_length", f32),
                ("offsets", P), ("x", P), ("y", P), ("z", P), ("w", P), ("cell_w", P),
                ("total_weight", f64), ("vx", P), ("vy", P), ("vz", P)]


class Catalogue(C.Structure):
    _fields_ = [("cx", C.POINTER(f32)), ("cy", C.POINTER(f32)), ("cz", C.POINTER(f32)),
                ("r", C.POINTER(f32)), ("n_in", C.POINTER(u64)), ("rung", C.POINTER(u32)),
                ("n", u64), ("capacity", u64)]


class RungStats(C.Structure):
    _fields_ = [("candidates", u64), ("masked", u64), ("proxy", u64), ("overlap", u64),
                ("witness", u64), ("rescale_fail", u64), ("accepted", u64),
                ("displaced", u64)]


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("orc_box_fold", f32, f32, f32)
_sig("orc_check_in_box", u64, P, P, P, u64, f32, C.POINTER(u64), C.POINTER(u64))
_sig("orc_cic_float", None, P, P, P, P, u64, u32, f32, u64, P)
_sig("orc_cic_unit_exponent", i32, u64, f32)
_sig("orc_cic_fixed", i32, P, P, P, P, u64, u32, f32, P)
_sig("orc_cic_fixed_density", None, P, u32, f32, i32, P, C.POINTER(f64))
_sig("orc_cic_brute_double", None, P, P, P, P, u64, u32, f32, P)
_sig("orc_density_contrast", i32, P, u64, C.POINTER(f64))
_sig("orc_density_contrast_with_mean", None, P, u64, f64)
_sig("orc_chain_mesh", None, P, P, P, P, u64, u32, f32, P, P, P, P, P, P, P,
     C.POINTER(f64))
_sig("orc_filter_lut_size", u32, u32)
_sig("orc_filter_lut", None, i32, f32, f32, f32, u32, P)
_sig("orc_window_axis", None, u32, i32, i32, P)
_sig("orc_window_corner_gain", f64, i32, i32, i32)
_sig("orc_mode_factor", f32, P, P, u32, u32, u32, u32)
_sig("orc_fft_r2c", i32, P, u32, P)
_sig("orc_fft_c2r", i32, P, u32, P)
_sig("orc_scan_candidates", u64, P, P, u64, f32, P)
_sig("orc_scan_minima", u64, P, u32, P, f32, P)
_sig("orc_sort_candidates_ref", None, P, u64)
_sig("orc_sort_candidates_flat", None, P, u64)
_sig("orc_template_build", i32, C.POINTER(Template), u32, f32, f32, f32)
_sig("orc_template_free", None, C.POINTER(Template))
_sig("orc_find_radius", f32, C.POINTER(Mesh), f32, f32, f32, f32, f32, f32, f32,
     C.POINTER(Template), u32, C.POINTER(C.c_uint8), C.POINTER(u64))
_sig("orc_find_radius_weighted", f32, C.POINTER(Mesh), f32, f32, f32, f32, f32, f32, f32,
     f32, C.POINTER(Template), u32, C.POINTER(C.c_uint8), C.POINTER(u64))
_sig("orc_witness", i32, C.POINTER(Mesh), f32, f32, f32, f32, f32, f32, f32,
     C.POINTER(Template), i32)
_sig("orc_brute_force_radius", f64, P, P, P, u64, f32, f64, f64, f64, f64, f64, f64)
_sig("orc_overlap_quick", i32, P, u32, u32, u32, u32, u32)
_sig("orc_mark_sphere", None, P, f32, f32, f32, f32, u32, f32)
_sig("orc_exodus_run", i32, C.POINTER(P), u32, f32, C.POINTER(Mesh), u64, P, u32, f32,
     f32, i32, u32, i32, i32, u32, C.POINTER(Catalogue), C.POINTER(RungStats))
_sig("orc_catalogue_free", None, C.POINTER(Catalogue))
_sig("orc_exodus_pipeline", i32, P, P, P, u64, u32, f32, u32, P, u32, f32, u32, i32,
     C.POINTER(Catalogue), P)
_sig("orc_fftf_r2c", i32, P, u32, P)
_sig("orc_fftf_filter_c2r", i32, P, u32, P, P, f32, P, P)
_sig("orc_num_threads", i32)


def _p(a):
    return None if a is None else a.ctypes.data


def f32a(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def check_in_box(x, y, z, box):
    x, y, z = f32a(x), f32a(y), f32a(z)
    edge, first = C.c_uint64(), C.c_uint64()
    bad = lib.orc_check_in_box(_p(x), _p(y), _p(z), len(x), box, C.byref(edge), C.byref(first))
    return bad, edge.value, first.value


def cic_float(x, y, z, n_cells, box, w=None, tile=0):
    x, y, z = f32a(x), f32a(y), f32a(z)
    w = None if w is None else f32a(w)
    out = np.empty(n_cells ** 3, np.float32)
    lib.orc_cic_float(_p(x), _p(y), _p(z), _p(w), len(x), n_cells, box, tile, _p(out))
    return out


def cic_fixed(x, y, z, n_cells, box, w=None, with_exponent=False):
    """(density float32, fixed-point int64 accumulators in 2^-s units,
    fixed-point mean[, s])."""
    x, y, z = f32a(x), f32a(y), f32a(z)
    w = None if w is None else f32a(w)
    acc = np.empty(n_cells ** 3, np.int64)
    s = lib.orc_cic_fixed(_p(x), _p(y), _p(z), _p(w), len(x), n_cells, box, _p(acc))
    out = np.empty(n_cells ** 3, np.float32)
    mean = C.c_double()
    lib.orc_cic_fixed_density(_p(acc), n_cells, box, s, _p(out), C.byref(mean))
    return (out, acc, mean.value, s) if with_exponent else (out, acc, mean.value)


def cic_unit_exponent(n, wmax):
    return lib.orc_cic_unit_exponent(n, wmax)


def cic_brute(x, y, z, n_cells, box, w=None):
    x, y, z = f32a(x), f32a(y), f32a(z)
    w = None if w is None else f32a(w)
    out = np.empty(n_cells ** 3, np.float64)
    lib.orc_cic_brute_double(_p(x), _p(y), _p(z), _p(w), len(x), n_cells, box, _p(out))
    return out


def density_contrast(values):
    v = np.array(values, dtype=np.float32, copy=True)
    mean = C.c_double()
    st = lib.orc_density_contrast(_p(v), v.size, C.byref(mean))
    if st != 0:
        raise ValueError(f"mean density {mean.value} not positive")
    return v, mean.value


def density_contrast_with_mean(values, mean):
    v = np.array(values, dtype=np.float32, copy=True)
    lib.orc_density_contrast_with_mean(_p(v), v.size, mean)
    return v


def chain_mesh(x, y, z, n_cells, box, w=None):
    x, y, z = f32a(x), f32a(y), f32a(z)
    w = None if w is None else f32a(w)
    n = len(x)
    cells = n_cells ** 3
    offsets = np.empty(cells + 1, np.uint64)
    mx, my, mz = (np.empty(n, np.float32) for _ in range(3))
    mw = np.empty(n, np.float32) if w is not None else None
    cw = np.empty(cells, np.float32) if w is not None else None
    orig = np.empty(n, np.uint64)
    tw = C.c_double()
    lib.orc_chain_mesh(_p(x), _p(y), _p(z), _p(w), n, n_cells, box, _p(offsets), _p(mx),
                       _p(my), _p(mz), _p(mw), _p(orig), _p(cw), C.byref(tw))
    return dict(offsets=offsets, x=mx, y=my, z=mz, w=mw, orig=orig, cell_weights=cw,
                total_weight=tw.value)


def filter_lut(filt, r, sigma, box, n):
    lut = np.empty(lib.orc_filter_lut_size(n), np.float32)
    lib.orc_filter_lut(filt, r, sigma, box, n, _p(lut))
    return lut


def window_axis(n, order=2, mode=0):
    a = np.empty(n, np.float64)
    lib.orc_window_axis(n, order, mode, _p(a))
    return a


def corner_gain(order, mode, interlaced=0):
    return lib.orc_window_corner_gain(order, mode, interlaced)


def mode_factors(lut, axis, n):
    """The per-mode smoothing factor over the half spectrum, as the reference
    applies it (fft.c:753-780), vectorised here in the same float/double
    steps: (float)((double)lut[k2] * (axis[ix] * axis[iy]) * axis[iz])."""
    h = n // 2
    idx = np.arange(n)
    k = np.where(idx > h, idx - n, idx)
    kz = np.arange(h + 1)
    k2 = (k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    s = lut[k2].astype(np.float32)
    if axis is None:
        return s
    cxy = axis[:, None] * axis[None, :]
    return (s.astype(np.float64) * cxy[:, :, None] * axis[None, None, :h + 1]).astype(np.float32)


def fft_r2c(values, n):
    v = np.ascontiguousarray(values, dtype=np.float64)
    spec = np.empty((n, n, n // 2 + 1, 2), np.float64)
    if lib.orc_fft_r2c(_p(v), n, _p(spec)) != 0:
        raise ValueError("n must be a power of two")
    return spec[..., 0] + 1j * spec[..., 1]


def fft_c2r(spec, n):
    s = np.empty((n, n, n // 2 + 1, 2), np.float64)
    s[..., 0] = spec.real
    s[..., 1] = spec.imag
    out = np.empty(n ** 3, np.float64)
    if lib.orc_fft_c2r(_p(s), n, _p(out)) != 0:
        raise ValueError("n must be a power of two")
    return out


def fftf_r2c(values, n):
    """Float forward transform (orc_fftf_r2c): the half spectrum as complex64."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    spec = np.empty((n, n, n // 2 + 1, 2), np.float32)
    if lib.orc_fftf_r2c(_p(v), n, _p(spec)) != 0:
        raise ValueError("n must be a power of two of at least 4")
    return spec[..., 0] + 1j * spec[..., 1]


def fftf_filter_c2r(spec, n, lut=None, axis=None, norm=1.0):
    """Float inverse transform (orc_fftf_filter_c2r) of spec times the mode
    factors of lut and axis, times norm."""
    s = np.empty((n, n, n // 2 + 1, 2), np.float32)
    s[..., 0] = spec.real
    s[..., 1] = spec.imag
    work = np.empty_like(s)
    out = np.empty(n ** 3, np.float32)
    lut = None if lut is None else np.ascontiguousarray(lut, np.float32)
    axis = None if axis is None else np.ascontiguousarray(axis, np.float64)
    if lib.orc_fftf_filter_c2r(_p(s), n, _p(lut), _p(axis), norm, _p(work), _p(out)) != 0:
        raise ValueError("n must be a power of two of at least 4")
    return out


def scan_candidates(delta, threshold, mask=None, tie_rule=0):
    d = f32a(delta)
    buf = (Candidate * max(d.size, 1))()
    k = lib.orc_scan_candidates(_p(d), _p(mask), d.size, threshold, buf)
    if tie_rule == 0:
        lib.orc_sort_candidates_ref(buf, k)
    else:
        lib.orc_sort_candidates_flat(buf, k)
    arr = np.ctypeslib.as_array(buf)[:k]
    return arr["delta"].copy(), arr["flat_idx"].copy()


def scan_minima(delta, n, threshold, mask=None, tie_rule=1):
    """The SIF_FINDER_SEEDS_MINIMA candidates of an n^3 field, sorted."""
    d = f32a(delta)
    assert d.size == n ** 3
    buf = (Candidate * max(d.size, 1))()
    k = lib.orc_scan_minima(_p(d), n, _p(mask), threshold, buf)
    if tie_rule == 0:
        lib.orc_sort_candidates_ref(buf, k)
    else:
        lib.orc_sort_candidates_flat(buf, k)
    arr = np.ctypeslib.as_array(buf)[:k]
    return arr["delta"].copy(), arr["flat_idx"].copy()


def sort_candidates_ref(delta, flat):
    buf = (Candidate * max(len(delta), 1))()
    arr = np.ctypeslib.as_array(buf)
    arr["delta"][:len(delta)] = delta
    arr["flat_idx"][:len(delta)] = flat
    lib.orc_sort_candidates_ref(buf, len(delta))
    return arr["delta"][:len(delta)].copy(), arr["flat_idx"][:len(delta)].copy()


def ref_sort_available():
    return os.path.exists(REF_SORT_PATH)


def sort_candidates_reference_build(delta, flat):
    """The same sort through the reference's own sort.h (oracle/_ref)."""
    ref = C.CDLL(REF_SORT_PATH)
    ref.ref_sort_candidates.argtypes = [P, u64]
    ref.ref_sort_candidates.restype = None
    buf = (Candidate * max(len(delta), 1))()
    arr = np.ctypeslib.as_array(buf)
    arr["delta"][:len(delta)] = delta
    arr["flat_idx"][:len(delta)] = flat
    ref.ref_sort_candidates(buf, len(delta))
    return arr["delta"][:len(delta)].copy(), arr["flat_idx"][:len(delta)].copy()


class OracleMesh:
    def __init__(self, m, n_cells, box):
        self.keep = m
        w = m.get("w")
        cw = m.get("cell_weights")
        v = [m.get(a) for a in ("vx", "vy", "vz")]
        self.s = Mesh(n_cells, box, np.float32(box) / np.float32(n_cells),
                      m["offsets"].ctypes.data, m["x"].ctypes.data, m["y"].ctypes.data,
                      m["z"].ctypes.data, None if w is None else w.ctypes.data,
                      None if cw is None else cw.ctypes.data,
                      float(m.get("total_weight", len(m["x"]))),
                      *[None if a is None else a.ctypes.data for a in v])


class OracleTemplate:
    def __init__(self, mesh_cells, cell_length, rmin, r_search):
        self.t = Template()
        st = lib.orc_template_build(C.byref(self.t), mesh_cells, cell_length, rmin, r_search)
        if st != 0:
            raise ValueError("search sphere wider than the mesh")

    def arrays(self):
        n = self.t.count
        g = lambda p, ct: np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), (n,)).copy()
        return dict(dx=g(self.t.dx, C.c_int32), dy=g(self.t.dy, C.c_int32),
                    dz=g(self.t.dz, C.c_int32), type=g(self.t.type, C.c_uint8),
                    min_d2=g(self.t.min_d2, C.c_float), max_d2=g(self.t.max_d2, C.c_float))

    def __del__(self):
        lib.orc_template_free(C.byref(self.t))


def find_radius(mesh, centres, rmin, r_search, threshold, vol_factor, window=8,
                mass_factor=None):
    """Rescaled radii of `centres`; weighted (against mass_factor) when the
    mesh carries weights."""
    tpl = OracleTemplate(mesh.s.n_cells, mesh.s.cell_length, rmin, r_search)
    out = np.empty(len(centres), np.float32)
    reason = np.empty(len(centres), np.uint8)
    nin = np.zeros(len(centres), np.uint64)
    weighted = bool(mesh.s.w)
    if weighted and mass_factor is None:
        raise ValueError("a weighted mesh needs its mass_factor")
    for i, (x, y, z) in enumerate(np.asarray(centres, np.float32)):
        rs = C.c_uint8()
        ni = C.c_uint64()
        if weighted:
            out[i] = lib.orc_find_radius_weighted(C.byref(mesh.s), x, y, z, r_search,
                                                  threshold, vol_factor, mass_factor, rmin,
                                                  C.byref(tpl.t), window, C.byref(rs),
                                                  C.byref(ni))
        else:
            out[i] = lib.orc_find_radius(C.byref(mesh.s), x, y, z, r_search, threshold,
                                         vol_factor, rmin, C.byref(tpl.t), window,
                                         C.byref(rs), C.byref(ni))
        reason[i] = rs.value
        nin[i] = ni.value
    return out, reason, nin


_sig("orc_profiles", i32, C.POINTER(Mesh), P, P, P, P, u64, f32, u32, u32, P, P, P)


def profiles(mesh, voids, ext, n_bins, opt=0, dens=True, vel=False):
    """sif_profiles restated (profiles.c:354-673): (dens, vel, count), each
    (n_voids, n_bins) or None; voids (n, 4) centres and radii."""
    v = np.ascontiguousarray(voids, dtype=np.float32).reshape(-1, 4)
    cols = [np.ascontiguousarray(v[:, k]) for k in range(4)]
    nv = len(v)
    d = np.zeros((nv, n_bins), np.float32) if dens else None
    u = np.zeros((nv, n_bins), np.float32) if vel else None
    cnt = np.zeros((nv, n_bins), np.uint64)
    st = lib.orc_profiles(C.byref(mesh.s), *[c.ctypes.data for c in cols], nv, ext, n_bins, opt,
                          None if d is None else d.ctypes.data,
                          None if u is None else u.ctypes.data, cnt.ctypes.data)
    if st != 0:
        raise ValueError("orc_profiles: bad arguments")
    return d, u, cnt


def brute_force_radius(x, y, z, box, centre, rung, r_search, threshold):
    x, y, z = f32a(x), f32a(y), f32a(z)
    return lib.orc_brute_force_radius(_p(x), _p(y), _p(z), len(x), box, float(centre[0]),
                                      float(centre[1]), float(centre[2]), rung, r_search,
                                      threshold)


def _catalogue(cat):
    n = cat.n
    if n == 0:
        return np.zeros((0, 4), np.float32), np.zeros(0, np.uint64), np.zeros(0, np.uint32)
    g = lambda p: np.ctypeslib.as_array(p, (n,)).copy()
    out = np.stack([g(cat.cx), g(cat.cy), g(cat.cz), g(cat.r)], axis=1)
    return out, g(cat.n_in), g(cat.rung)


def exodus_run(fields, n_cells, box, mesh, n_particles, radii_desc, threshold,
               overlap=0.0, lagrangian=False, search_opt=0, tie_rule=0, witness=True,
               window=8):
    """Sequential exodus commit on given per-rung smoothed fields."""
    fields = [np.ascontiguousarray(f, dtype=np.float32) for f in fields]
    ptrs = (P * len(fields))(*[f.ctypes.data for f in fields])
    radii = np.ascontiguousarray(radii_desc, dtype=np.float32)
    cat = Catalogue()
    stats = (RungStats * len(radii))()
    st = lib.orc_exodus_run(ptrs, n_cells, box, C.byref(mesh.s), n_particles,
                            radii.ctypes.data, len(radii), threshold, overlap,
                            1 if lagrangian else 0, search_opt, tie_rule,
                            1 if witness else 0, window, C.byref(cat), stats)
    if st != 0:
        lib.orc_catalogue_free(C.byref(cat))
        raise ValueError(f"oracle exodus failed ({st})")
    out = _catalogue(cat)
    lib.orc_catalogue_free(C.byref(cat))
    return out + ([dict((k, getattr(s, k)) for k, _ in RungStats._fields_) for s in stats],)


def exodus_pipeline(x, y, z, n_cells, box, mesh_cells, radii, threshold, max_rungs=0,
                    fp32=False):
    """The whole CPU path from particles (orc_exodus_pipeline): transforms in
    double (each field rounded to float) or, with fp32, in float as the
    reference's FFTW plans run. Returns (catalogue, counts, stage seconds)."""
    x, y, z = f32a(x), f32a(y), f32a(z)
    radii = np.ascontiguousarray(radii, dtype=np.float32)
    cat = Catalogue()
    stage = np.zeros(6, np.float64)
    st = lib.orc_exodus_pipeline(_p(x), _p(y), _p(z), len(x), n_cells, box, mesh_cells,
                                 _p(radii), len(radii), threshold, max_rungs, int(fp32),
                                 C.byref(cat), _p(stage))
    if st != 0:
        lib.orc_catalogue_free(C.byref(cat))
        raise ValueError(f"oracle pipeline failed ({st})")
    out = _catalogue(cat)
    lib.orc_catalogue_free(C.byref(cat))
    return out[0], out[1], stage



# --- survey (src/finder/exodus_survey.c, exodus_rescale.c:1909-2317) -------------

_sig("orc_survey_node", u32, f32, f32, u32)
_sig("orc_survey_randoms_per_node", f64, f64)
_sig("orc_survey_close", None, P, u32, C.c_int32)
_sig("orc_survey_occupancy", i32, P, P, P, u64, u32, f32, P, C.POINTER(f64), C.POINTER(u64),
     C.POINTER(u64))
_sig("orc_survey_contrast", None, P, P, P, f32, u64)
_sig("orc_survey_void_footprint", None, P, u32, f32, f32, f32, f32, f32, f32,
     C.POINTER(f32), C.POINTER(f32))
_sig("orc_find_radius_survey", f32, C.POINTER(Mesh), C.POINTER(Mesh), C.POINTER(Template),
     C.POINTER(Template), f32, f32, f32, f32, f32, f32, u32, C.POINTER(C.c_uint8),
     C.POINTER(u64))
_sig("orc_exodus_run_survey", i32, C.POINTER(P), u32, f32, C.POINTER(Mesh), C.POINTER(Mesh),
     f32, P, u32, f32, f32, i32, u32, i32, u32, C.POINTER(Catalogue), C.POINTER(RungStats))


def survey_randoms_per_node(m):
    return lib.orc_survey_randoms_per_node(float(m))


def survey_close(occ, n, k):
    """The (2k + 1)^3 closing of a 0/1 byte mask of n^3 nodes, in a copy."""
    o = np.ascontiguousarray(occ, dtype=np.uint8).copy()
    lib.orc_survey_close(_p(o), n, k)
    return o


def survey_occupancy(rx, ry, rz, n, box):
    """(mask of n^3 bytes, k, lambda, marked, observed) for the randoms on a
    grid of n cells over `box` (cell length box / n in float)."""
    rx, ry, rz = f32a(rx), f32a(ry), f32a(rz)
    occ = np.empty(n ** 3, np.uint8)
    lam, marked, observed = C.c_double(), C.c_uint64(), C.c_uint64()
    k = lib.orc_survey_occupancy(_p(rx), _p(ry), _p(rz), len(rx), n,
                                 np.float32(box) / np.float32(n), _p(occ), C.byref(lam),
                                 C.byref(marked), C.byref(observed))
    return occ, k, lam.value, marked.value, observed.value


def survey_contrast(data, randoms, occ, alpha):
    d = np.ascontiguousarray(data, dtype=np.float32).copy()
    r = f32a(randoms)
    o = np.ascontiguousarray(occ, dtype=np.uint8)
    lib.orc_survey_contrast(_p(d), _p(r), _p(o), alpha, d.size)
    return d


def survey_footprint(occ, n, box, voids):
    """(inside, shell) observed fractions of voids (k, 4)."""
    o = np.ascontiguousarray(occ, dtype=np.uint8)
    v = np.asarray(voids, np.float32).reshape(-1, 4)
    l = np.float32(box) / np.float32(n)
    fin = np.empty(len(v), np.float32)
    fsh = np.empty(len(v), np.float32)
    a, b = C.c_float(), C.c_float()
    for i, (x, y, z, r) in enumerate(v):
        lib.orc_survey_void_footprint(_p(o), n, l, box, x, y, z, r, C.byref(a), C.byref(b))
        fin[i], fsh[i] = a.value, b.value
    return fin, fsh


def find_radius_survey(dmesh, rmesh, centres, rmin, r_search, K, window=8):
    """Survey radii of `centres` (reason, enclosed data tracers beside)."""
    dt = OracleTemplate(dmesh.s.n_cells, dmesh.s.cell_length, rmin, r_search)
    rt = OracleTemplate(rmesh.s.n_cells, rmesh.s.cell_length, rmin, r_search)
    c = np.asarray(centres, np.float32).reshape(-1, 3)
    out = np.empty(len(c), np.float32)
    reason = np.empty(len(c), np.uint8)
    nin = np.zeros(len(c), np.uint64)
    for i, (x, y, z) in enumerate(c):
        rs, ni = C.c_uint8(), C.c_uint64()  # n_in is written only when a radius is found
        out[i] = lib.orc_find_radius_survey(C.byref(dmesh.s), C.byref(rmesh.s), C.byref(dt.t),
                                            C.byref(rt.t), x, y, z, r_search, K, rmin, window,
                                            C.byref(rs), C.byref(ni))
        reason[i] = rs.value
        nin[i] = ni.value
    return out, reason, nin


def exodus_run_survey(fields, n_cells, box, dmesh, rmesh, survey_K, radii_desc, threshold,
                      overlap=0.0, lagrangian=False, search_opt=0, tie_rule=0, window=8):
    """Sequential survey commit on given per-rung survey contrasts."""
    fields = [np.ascontiguousarray(f, dtype=np.float32) for f in fields]
    ptrs = (P * len(fields))(*[f.ctypes.data for f in fields])
    radii = np.ascontiguousarray(radii_desc, dtype=np.float32)
    cat = Catalogue()
    stats = (RungStats * len(radii))()
    st = lib.orc_exodus_run_survey(ptrs, n_cells, box, C.byref(dmesh.s), C.byref(rmesh.s),
                                   survey_K, radii.ctypes.data, len(radii), threshold, overlap,
                                   1 if lagrangian else 0, search_opt, tie_rule, window,
                                   C.byref(cat), stats)
    if st != 0:
        lib.orc_catalogue_free(C.byref(cat))
        raise ValueError(f"oracle survey exodus failed ({st})")
    out = _catalogue(cat)
    lib.orc_catalogue_free(C.byref(cat))
    return out + ([dict((k, getattr(s, k)) for k, _ in RungStats._fields_) for s in stats],)
