"""Python entry points over the device-level C ABI (include/sif_amd.h): the
stages of the exodus path on device-resident arrays. Each function calls the
HIP library; there is no other implementation behind them.
"""
import ctypes as C

import numpy as np

from . import _native as N
from ._native import DeviceArray, check, lib


def to_device(a, dtype=np.float32):
    return DeviceArray.from_host(np.ascontiguousarray(a, dtype=dtype))


def check_in_box(dx, dy, dz, n, box):
    bad, edge, first = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib.sif_amd_check_in_box(dx.ptr, dy.ptr, dz.ptr, n, box, C.byref(bad),
                                   C.byref(edge), C.byref(first)))
    return bad.value, edge.value, first.value


def cic(dx, dy, dz, n, n_cells, box, dw=None, out=None, work=None):
    """CIC density per unit volume (grid.c:679-746) as a device array."""
    total = n_cells ** 3
    out = out or DeviceArray((total,), np.float32)
    own_work = work is None
    work = work or DeviceArray((total,), np.int64)
    check(lib.sif_amd_grid_assign_cic(dx.ptr, dy.ptr, dz.ptr, dw.ptr if dw else None, n,
                                      n_cells, box, out.ptr, work.ptr))
    if own_work:
        work.free()
    return out


def density_contrast(dvalues, total_cells):
    mean = C.c_double()
    check(lib.sif_amd_grid_to_density_contrast(dvalues.ptr, total_cells, C.byref(mean)))
    return mean.value


def grid_mean(dvalues, total_cells):
    mean = C.c_double()
    check(lib.sif_amd_grid_mean(dvalues.ptr, total_cells, C.byref(mean)))
    return mean.value


class Mesh:
    """A chain mesh on the device (chain_mesh.h:40-108 layout, SoA arrays)."""

    def __init__(self, dx, dy, dz, n, n_cells, box, dw=None, indices=True):
        cells = n_cells ** 3
        self.n_cells, self.box, self.n = n_cells, float(box), n
        self.offsets = DeviceArray((cells + 1,), np.uint64)
        self.x = DeviceArray((n,), np.float32)
        self.y = DeviceArray((n,), np.float32)
        self.z = DeviceArray((n,), np.float32)
        self.w = DeviceArray((n,), np.float32) if dw else None
        self.cell_weights = DeviceArray((cells,), np.float32) if dw else None
        self.orig = DeviceArray((n,), np.uint64) if indices else None
        tw = C.c_double()
        check(lib.sif_amd_chain_mesh_build(
            dx.ptr, dy.ptr, dz.ptr, dw.ptr if dw else None, n, n_cells, box,
            self.offsets.ptr, self.x.ptr, self.y.ptr, self.z.ptr,
            self.w.ptr if self.w else None, self.orig.ptr if self.orig else None,
            self.cell_weights.ptr if self.cell_weights else None, C.byref(tw)))
        self.total_weight = tw.value


def spectrum_pitch(n):
    """Complex values per row of a device spectrum (n/2 + 1 modes, padded)."""
    return int(lib.sif_amd_spectrum_pitch(n))


def spectrum_plane(n):
    """Complex values per x-plane of a device spectrum (n padded rows and a
    plane pad)."""
    return int(lib.sif_amd_spectrum_bytes(n)) // (8 * n)


def spectrum_buffer(n):
    return DeviceArray((n, spectrum_plane(n), 2), np.float32)


def spectrum(dvalues, n):
    """The r2c transform of an n^3 device field as a device array
    (n, plane, 2): x-plane i holds rows of spectrum_pitch(n) complex values,
    the first n/2 + 1 of each row the modes (spectrum_modes reads them)."""
    spec = spectrum_buffer(n)
    work = spectrum_buffer(n)
    check(lib.sif_amd_fft_r2c(dvalues.ptr, n, spec.ptr, work.ptr))
    work.free()
    return spec


def spectrum_modes(dspec, n):
    """A device spectrum on the host as float32 (n, n, n/2 + 1, 2), FFTW's
    r2c layout (real, imaginary)."""
    p = spectrum_pitch(n)
    s = dspec.get()[:, :n * p].reshape(n, n, p, 2)
    return s[:, :, :n // 2 + 1]


def smooth(dspec, n, box, filt, r, sigma=0.0, window_order=2, window_mode=0, out=None,
           work=None):
    out = out or DeviceArray((n ** 3,), np.float32)
    own = work is None
    work = work or spectrum_buffer(n)
    check(lib.sif_amd_fft_filter_c2r(dspec.ptr, n, box, filt, r, sigma, window_order,
                                     window_mode, out.ptr, work.ptr))
    if own:
        work.free()
    return out


def scan_candidates(ddelta, total, threshold, dmask=None, capacity=None):
    cap = capacity or total
    d_delta = DeviceArray((max(cap, 1),), np.float32)
    d_flat = DeviceArray((max(cap, 1),), np.uint32)
    count = C.c_uint64()
    check(lib.sif_amd_scan_candidates(ddelta.ptr, dmask.ptr if dmask else None, total,
                                      threshold, d_delta.ptr, d_flat.ptr, cap,
                                      C.byref(count)))
    k = count.value
    return d_delta.get()[:k], d_flat.get()[:k]


def scan_minima(ddelta, n, threshold, dmask=None, capacity=None):
    """The SIF_FINDER_SEEDS_MINIMA candidates of an n^3 device field:
    (depths, flat indices), sorted (depth, flat)."""
    total = n ** 3
    cap = capacity or total
    d_delta = DeviceArray((max(cap, 1),), np.float32)
    d_flat = DeviceArray((max(cap, 1),), np.uint32)
    count = C.c_uint64()
    check(lib.sif_amd_scan_minima(ddelta.ptr, dmask.ptr if dmask else None, n, threshold,
                                  d_delta.ptr, d_flat.ptr, cap, C.byref(count)))
    k = count.value
    return d_delta.get()[:k], d_flat.get()[:k]


def find_radius(mesh, centres, rmin, r_search, threshold, vol_factor, mass_factor=None):
    """Grows spheres at (k, 3) float32 centres; returns radius, reason, n_in.
    On a weighted mesh the density tests use mass_factor."""
    c = np.ascontiguousarray(centres, dtype=np.float32)
    k = c.shape[0]
    dcx, dcy, dcz = (to_device(c[:, a]) for a in range(3))
    dr = DeviceArray((max(k, 1),), np.float32)
    drs = DeviceArray((max(k, 1),), np.uint32)
    dn = DeviceArray((max(k, 1),), np.uint64)
    if mesh.w is not None:
        if mass_factor is None:
            raise ValueError("a weighted mesh needs its mass_factor")
        check(lib.sif_amd_find_radius_weighted(
            mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr, mesh.w.ptr,
            mesh.cell_weights.ptr, mesh.n_cells, mesh.box, dcx.ptr, dcy.ptr, dcz.ptr, k, rmin,
            r_search, threshold, vol_factor, mass_factor, dr.ptr, drs.ptr, dn.ptr))
    else:
        check(lib.sif_amd_find_radius(mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr,
                                      mesh.n_cells, mesh.box, dcx.ptr, dcy.ptr, dcz.ptr, k,
                                      rmin, r_search, threshold, vol_factor, dr.ptr, drs.ptr,
                                      dn.ptr))
    return dr.get()[:k], drs.get()[:k], dn.get()[:k]


def tracer_profiles(mesh, voids, ext, n_bins, options=0, dens=True, velocities=None):
    """sif_profiles on device arrays: (dens, vel), each (n_voids, n_bins) or
    None. voids: (n, 4) centres and radii (host); velocities: (vx, vy, vz)
    device arrays in mesh order, or None for no velocity profile."""
    v = np.ascontiguousarray(voids, dtype=np.float32).reshape(-1, 4)
    nv = len(v)
    cols = [to_device(np.ascontiguousarray(v[:, k])) for k in range(4)]
    dd = DeviceArray((max(nv * n_bins, 1),), np.float32) if dens else None
    dv = DeviceArray((max(nv * n_bins, 1),), np.float32) if velocities is not None else None
    for a in (dd, dv):
        if a is not None:
            check(lib.sif_amd_memset(a.ptr, 0, a.nbytes))
    vx, vy, vz = velocities if velocities is not None else (None, None, None)
    check(lib.sif_amd_tracer_profiles(
        mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr,
        mesh.w.ptr if mesh.w is not None else None, vx.ptr if vx else None,
        vy.ptr if vy else None, vz.ptr if vz else None, mesh.n_cells, mesh.box,
        mesh.total_weight, cols[0].ptr, cols[1].ptr, cols[2].ptr, cols[3].ptr, nv, ext, n_bins,
        options, dd.ptr if dd else None, dv.ptr if dv else None))
    out = [a.get()[:nv * n_bins].reshape(nv, n_bins) if a is not None else None for a in (dd, dv)]
    return out[0], out[1]


def exodus(ddelta, n_cells, box, mesh, radii, threshold, overlap=0.0, options=0,
           restore=True):
    """The exodus finder on device data: (catalogue (n, 4), n_in, stats); on
    the mesh's weights when it carries them."""
    radii = np.ascontiguousarray(radii, dtype=np.float32)
    nr = len(radii)
    stats = (N.RungStats * nr)()
    nin = C.POINTER(C.c_uint64)()
    rp = radii.ctypes.data_as(C.POINTER(C.c_float))
    if mesh.w is not None:
        cat = lib.sif_amd_exodus_device_weighted(
            ddelta.ptr, n_cells, box, mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr,
            mesh.w.ptr, mesh.cell_weights.ptr, mesh.total_weight, mesh.n_cells, mesh.n, rp,
            nr, threshold, overlap, options, 1 if restore else 0, C.byref(nin), stats)
    else:
        cat = lib.sif_amd_exodus_device(
            ddelta.ptr, n_cells, box, mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr,
            mesh.n_cells, mesh.n, rp, nr, threshold, overlap, options, 1 if restore else 0,
            C.byref(nin), stats)
    if not cat:
        raise N.error_for(lib.sif_error_status())
    out = N.catalogue_to_numpy(cat)
    n_in = np.ctypeslib.as_array(nin, (max(len(out), 1),))[:len(out)].copy()
    N.libc.free(C.cast(nin, C.c_void_p))
    st = [dict(radius=s.radius, candidates=s.candidates, masked=s.masked, proxy=s.proxy,
               overlap=s.overlap, rescale_fail=s.rescale_fail, accepted=s.accepted,
               evaluated=s.evaluated) for s in stats]
    return out, n_in, st


# --- survey (src/finder/exodus_survey.c) ---


def mesh_desc(mesh):
    """The sif_amd_mesh_t of a device Mesh."""
    w = mesh.w is not None
    return N.MeshDesc(mesh.offsets.ptr, mesh.x.ptr, mesh.y.ptr, mesh.z.ptr,
                      mesh.w.ptr if w else None, mesh.cell_weights.ptr if w else None,
                      mesh.total_weight if w else float(mesh.n), mesh.n_cells, mesh.n)


def survey_occupancy(dx, dy, dz, n, n_cells, box):
    """(occupied nodes as an n_cells^3 uint8 device array, k, lambda, marked,
    observed) for positions on the device."""
    occ = DeviceArray((n_cells ** 3,), np.uint8)
    k, lam, marked, observed = C.c_int32(), C.c_double(), C.c_uint64(), C.c_uint64()
    check(lib.sif_amd_survey_occupancy(dx.ptr, dy.ptr, dz.ptr, n, n_cells, box, occ.ptr,
                                       C.byref(k), C.byref(lam), C.byref(marked),
                                       C.byref(observed)))
    return occ, k.value, lam.value, marked.value, observed.value


def survey_contrast(ddata, drandoms, docc, alpha):
    """In place on ddata: data / (alpha randoms) - 1 on observed nodes, 1e30 elsewhere."""
    check(lib.sif_amd_survey_contrast(ddata.ptr, drandoms.ptr, docc.ptr, alpha,
                                      int(np.prod(ddata.shape))))


def survey_footprint(docc, n_cells, box, voids):
    """(inside, shell) observed fractions of voids (k, 4) (host arrays)."""
    v = np.ascontiguousarray(np.asarray(voids, np.float32).reshape(-1, 4).T)
    k = v.shape[1]
    fin = np.empty(k, np.float32)
    fsh = np.empty(k, np.float32)
    check(lib.sif_amd_survey_footprint(docc.ptr, n_cells, box, v[0].ctypes.data,
                                       v[1].ctypes.data, v[2].ctypes.data, v[3].ctypes.data, k,
                                       fin.ctypes.data, fsh.ctypes.data))
    return fin, fsh


def find_radius_survey(dmesh, rmesh, centres, rmin, r_search, K):
    """Survey radii of host `centres` (k, 3): (radius, reason, n_in)."""
    c = np.ascontiguousarray(np.asarray(centres, np.float32).reshape(-1, 3).T)
    k = c.shape[1]
    dc = [DeviceArray.from_host(c[a]) for a in range(3)]
    rad = DeviceArray((max(k, 1),), np.float32)
    rea = DeviceArray((max(k, 1),), np.uint32)
    nin = DeviceArray((max(k, 1),), np.uint64)
    dd, rd = mesh_desc(dmesh), mesh_desc(rmesh)
    check(lib.sif_amd_find_radius_survey(C.byref(dd), C.byref(rd), dmesh.box, dc[0].ptr,
                                         dc[1].ptr, dc[2].ptr, k, rmin, r_search, K, rad.ptr,
                                         rea.ptr, nin.ptr))
    return rad.get()[:k], rea.get()[:k], nin.get()[:k]


def exodus_survey(ddata, drandoms, n_cells, box, dmesh, rmesh, radii, threshold, overlap=0.0,
                  options=0, restore=True):
    """The survey finder on device data: (catalogue (n, 4), footprint (n, 2),
    n_in, stats). ddata holds the data's CIC densities, drandoms the randoms'."""
    radii = np.ascontiguousarray(radii, dtype=np.float32)
    nr = len(radii)
    stats = (N.RungStats * nr)()
    nin = C.POINTER(C.c_uint64)()
    dd, rd = mesh_desc(dmesh), mesh_desc(rmesh)
    cat = lib.sif_amd_exodus_survey_device(
        ddata.ptr, drandoms.ptr, n_cells, box, C.byref(dd), C.byref(rd),
        radii.ctypes.data_as(C.POINTER(C.c_float)), nr, threshold, overlap, options,
        1 if restore else 0, C.byref(nin), stats)
    if not cat:
        raise N.error_for(lib.sif_error_status())
    out, fp = N.catalogue_to_numpy(cat, footprint=True)
    n_in = np.ctypeslib.as_array(nin, (max(len(out), 1),))[:len(out)].copy()
    N.libc.free(C.cast(nin, C.c_void_p))
    st = [dict(radius=s.radius, candidates=s.candidates, masked=s.masked, proxy=s.proxy,
               overlap=s.overlap, rescale_fail=s.rescale_fail, accepted=s.accepted,
               evaluated=s.evaluated) for s in stats]
    return out, fp, n_in, st


def set_cell_lists(on=True):
    """Whether later exodus runs list each rung's candidate cells in the
    smoothing (the default) or write the smoothed field; the catalogue is
    the same."""
    check(lib.sif_amd_set_cell_lists(1 if on else 0))


def set_void_room(voids=0):
    """Frees the device copy of the catalogue that exodus runs keep and sets
    the voids it is next allocated for (0: the default); the catalogue does
    not depend on it."""
    check(lib.sif_amd_set_void_room(int(voids)))


def set_batch_window(floor=0, ceiling=0):
    """The growth batch window of later exodus runs in this process; (0, 0)
    restores the defaults."""
    check(lib.sif_amd_set_batch_window(floor, ceiling))


def _boxes(boxes):
    """ctypes boxes from dicts: d (dx, dy, dz device arrays), n, n_cells,
    mesh_cells, box, radii; the ladders are kept alive with the array."""
    arr = (N.Box * len(boxes))()
    keep = []
    for b, bx in enumerate(boxes):
        r = np.ascontiguousarray(bx["radii"], dtype=np.float32)
        keep.append(r)
        dx, dy, dz = bx["d"]
        arr[b] = N.Box(dx.ptr, dy.ptr, dz.ptr, bx["n"], bx["n_cells"], bx["mesh_cells"],
                       bx["box"], r.ctypes.data_as(C.POINTER(C.c_float)), len(r))
    return arr, keep


TUNING_FIELDS = ("workers", "cic_tile", "split_waves", "batch_floor", "batch_ceiling")


def exodus_batch(boxes, threshold, overlap=0.0, options=0, tuning=None):
    """Every box's exodus catalogue (sif_amd_exodus_batch): a list of (n, 4)
    arrays in box order. tuning: a dict of TUNING_FIELDS (missing: defaults)."""
    arr, keep = _boxes(boxes)
    t = N.Tuning(*[int((tuning or {}).get(k, 0)) for k in TUNING_FIELDS])
    cats = (C.POINTER(N.Catalogue) * len(boxes))()
    check(lib.sif_amd_exodus_batch(arr, len(boxes), threshold, overlap, options, C.byref(t),
                                   cats))
    return [N.catalogue_to_numpy(c) for c in cats]


def batch_autotune(boxes, threshold, overlap=0.0, options=0, probe=8):
    """The fastest tuning on `probe` of the boxes (sif_amd_batch_autotune)."""
    arr, keep = _boxes(boxes)
    t = N.Tuning()
    check(lib.sif_amd_batch_autotune(arr, len(boxes), threshold, overlap, options, probe,
                                     C.byref(t)))
    return {k: getattr(t, k) for k in TUNING_FIELDS}


def exodus_fields(ddelta, n_cells, box, radii, options=0):
    radii = np.ascontiguousarray(radii, dtype=np.float32)
    total = n_cells ** 3
    d = DeviceArray((len(radii), total), np.float32)
    check(lib.sif_amd_exodus_fields(ddelta.ptr, n_cells, box,
                                    radii.ctypes.data_as(C.POINTER(C.c_float)),
                                    len(radii), options, d.ptr))
    return d.get()
