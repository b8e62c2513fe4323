"""exodus particles/s on one MI355X, with the HBM roofline of the dominant
kernel and the oracle's CPU path timed beside it.

A step is one full exodus catalogue of a seeded Zel'dovich particle set whose
positions are already in HBM: CIC deposit, density contrast, chain mesh,
forward FFT, then every rung of the radius ladder (top-hat smoothing with the
CIC window divided out, candidate scan and sort, prefilter, sphere growth,
ordered commit). Default workload: 512^3 particles in a 1000 Mpc/h box, 512^3
grid, mesh cell 10 Mpc/h, threshold -0.8, a geometric 5% ladder from two grid
cells to 40 Mpc/h.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config plain512|plain128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: particles per side, box, grid cells per side, mesh cell, rmax, seed
    "plain512": dict(n_side=512, box=1000.0, n_cells=512, mesh_cell=10.0, rmax=40.0, seed=2),
    "plain128": dict(n_side=128, box=250.0, n_cells=128, mesh_cell=5.0, rmax=40.0, seed=1),
    # one eighth of plain512's volume at the same particle density, grid
    # spacing, mesh cell and ladder
    "sample256": dict(n_side=256, box=500.0, n_cells=256, mesh_cell=10.0, rmax=40.0, seed=2),
    # eight times the volume at the same density and grid spacing, a 15 Mpc/h
    # mesh cell (133^3) and radii to 60 Mpc/h
    "plain1024": dict(n_side=1024, box=2000.0, n_cells=1024, mesh_cell=15.0, rmax=60.0, seed=5),
}
THRESHOLD = -0.8
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(cfg):
    from sif_amd import zeldovich
    c = CONFIGS[cfg]
    t0 = time.time()
    x, y, z = zeldovich.generate(c["n_side"], c["box"], c["seed"],
                                 progress=lambda m: log(f"[bench] {m}, {time.time() - t0:.1f}s"))
    log(f"[bench] generated {c['n_side']}^3 Zel'dovich particles in {time.time() - t0:.1f}s, "
        f"crc32 {zeldovich.crc32(x, y, z):08x}")
    cell = c["box"] / c["n_cells"]
    radii = zeldovich.ladder(2.0 * cell, c["rmax"], 0.05)
    mesh_cells = int(round(c["box"] / c["mesh_cell"]))
    return c, x, y, z, radii, mesh_cells


class Step:
    """Device buffers allocated once; run() is one catalogue."""

    def __init__(self, c, x, y, z, radii, mesh_cells):
        from sif_amd import device as D
        from sif_amd._native import DeviceArray
        self.D = D
        self.c, self.radii, self.mesh_cells = c, radii, mesh_cells
        self.n = len(x)
        self.dx, self.dy, self.dz = D.to_device(x), D.to_device(y), D.to_device(z)
        n3 = c["n_cells"] ** 3
        self.grid = DeviceArray((n3,), np.float32)
        self.work = DeviceArray((n3,), np.int64)

    def run(self):
        D, c = self.D, self.c
        D.cic(self.dx, self.dy, self.dz, self.n, c["n_cells"], c["box"], out=self.grid,
              work=self.work)
        D.density_contrast(self.grid, c["n_cells"] ** 3)
        mesh = D.Mesh(self.dx, self.dy, self.dz, self.n, self.mesh_cells, c["box"],
                      indices=False)
        cat, nin, stats = D.exodus(self.grid, c["n_cells"], c["box"], mesh, self.radii,
                                   THRESHOLD, restore=False)
        return cat, stats


# stage -> (the kernels one timed launch of it runs, how many)
STAGE_KERNELS = {
    "grow": ("grow_kernel (+ grow_bin_kernel on launches of under 2048 candidates)", 1),
    "smooth": ("cols_kernel x2 + c2r_rows_kernel (one filtered c2r; its rows pass lists "
               "the cells at or below the threshold instead of writing the field)", 3),
    "cic": ("cic_count/bin/tile_kernel + cic_finish_kernel (tile binning, LDS deposit, flush)", 4),
    "mesh": ("mesh_key_kernel + radix sort + mesh_gather_kernel", 17),
    "scan": ("list_count_kernel + tile scan + list_write_kernel (the smoothing's cell lists "
             "less the masked cells)", 3),
    "fft_forward": ("r2c_rows_kernel + cols_kernel x2", 3),
}


# HBM traffic per launch of a stage, measured with rocprof's FETCH_SIZE and
# WRITE_SIZE counters scaled against a copy kernel of known size; the summary
# that holds it names the commands and carries a line
# "traffic_bytes_per_launch[<stage>]: <bytes>" per stage measured
TRAFFIC = {"plain512": "profiles/hbm_traffic_plain512.md"}
# The kernels' own durations per launch of a stage from the rocprofv3 kernel
# trace of the same command ("kernel_ms_per_launch[<stage>]: <ms>"): the HIP
# events here bracket a stage's launches on its stream, which also holds the
# waits of its kernels for CUs the other stream's kernels occupy.
TRACE = {"plain512": "profiles/exodus_plain512_kernel_trace.md"}


def summary_value(table, config, key, cast):
    """(value, the summary under profiles/ that holds it) or (None, None)."""
    path = table.get(config)
    if not path or not os.path.exists(os.path.join(ROOT, path)):
        return None, None
    for line in open(os.path.join(ROOT, path)):
        if line.startswith(key):
            return cast(line.split(":")[1]), path
    return None, None


def measured_traffic(config, stage):
    """(bytes per launch, the summary under profiles/ that holds it) or
    (None, None)."""
    return summary_value(TRAFFIC, config, f"traffic_bytes_per_launch[{stage}]:", int)


def roofline(timing, config):
    """Achieved HBM rate of each timed stage whose algorithmic bytes are
    known (bytes / event-timed duration over the timed region), and that of
    the dominant one: the stage with the most device time per step (a stage
    of several kernels is timed as one region; the smoothing's runs on a
    second stream beside the commit). Its measured traffic is read from the
    rocprof counter summary named in `traffic_source`, and its kernels' own
    durations from the kernel trace named in `profiled_source` (earlier
    profiled runs of this workload, not this run)."""
    stages = {}
    for name, (ms, launches, by) in timing.items():
        if name not in STAGE_KERNELS or launches == 0 or by <= 0 or ms <= 0:
            continue
        gbs = by / (ms * 1e-3) / 1e9
        stages[name] = {"kernel": STAGE_KERNELS[name][0], "launches": launches,
                        "avg_launch_ms": round(ms / launches, 4),
                        "bytes_per_launch": round(by / launches),
                        "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    top = max(stages, key=lambda k: timing[k][0])
    d = stages[top]
    traffic, source = measured_traffic(config, top)
    kernel_ms, trace = summary_value(TRACE, config, f"kernel_ms_per_launch[{top}]:", float)
    return {"bound": "hbm", "kernel": d["kernel"], "stage": top, "achieved": d["achieved"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
            "traffic": traffic, "traffic_source": source,
            "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_launch_ms"],
            # the kernels' own durations a launch in the rocprof trace of
            # this command, and the rate the algorithmic bytes make over them
            "profiled_kernel_ms": kernel_ms, "profiled_source": trace,
            "profiled_achieved": (round(d["bytes_per_launch"] / (kernel_ms * 1e-3) / 1e9, 1)
                                  if kernel_ms else None),
            "stages": stages}


def cpu_baseline(c, x, y, z, radii, mesh_cells, gpu_cat):
    """The oracle's CPU path -- the reference algorithm restated in C with
    OpenMP: float CIC in the reference's order, contrast, chain mesh, a
    single-precision FFT of its own with the reference's plane and pencil
    structure (FFTW, which the reference links, is absent), each rung
    smoothed as the finder reaches it, the sequential finder -- on the same
    particles as the timed GPU steps: the whole catalogue, on the host's
    cores. Also says how its catalogue compares with the GPU's: the voids
    found at the same node by both, and of those the ones of equal radius
    (the FP32 fields of the two differ in their last bits, which can move a
    decision between candidates of nearly equal depth)."""
    import oracle
    t0 = time.time()
    cat, _, stage = oracle.exodus_pipeline(x, y, z, c["n_cells"], c["box"], mesh_cells, radii,
                                           THRESHOLD, fp32=True)
    dt = time.time() - t0
    # voids at the same node (centres are node index times the cell length,
    # bit-identical when the node is), and of those the radii bit-equal
    key = {tuple(v.view(np.uint32)[:3]): i for i, v in enumerate(cat)}
    pairs = [(key[tuple(v.view(np.uint32)[:3])], j) for j, v in enumerate(gpu_cat)
             if tuple(v.view(np.uint32)[:3]) in key]
    same_r = sum(1 for i, j in pairs if cat[i, 3] == gpu_cat[j, 3])
    return {"value": round(len(x) / dt, 1), "unit": "particles/s",
            "cores": oracle.lib.orc_num_threads(), "kind": "port",
            "sample": f"the whole workload: {c['n_side']}^3 particles, L={c['box']:g}, "
                      f"{c['n_cells']}^3 grid, mesh {mesh_cells}^3, {len(radii)} rungs, one "
                      f"catalogue ({len(cat)} voids) in {dt:.1f}s; stages "
                      f"cic/contrast/mesh/fft/smooth/finder "
                      f"{'/'.join(f'{s:.2f}' for s in stage)} s; float FFT of the oracle's own",
            "against_gpu": {"voids_cpu": int(len(cat)), "voids_gpu": int(len(gpu_cat)),
                            "same_node": len(pairs), "same_node_and_radius": same_r,
                            "identical": bool(len(pairs) == len(cat) == len(gpu_cat) ==
                                              same_r)}}


# the differentiable profile chain: 256^3 Zel'dovich particles in a
# 500 Mpc/h box, 256^3 grid, Gaussian R = 3 Mpc/h, 1e5 centres uniform in the
# box with radii uniform in 5-40 Mpc/h, 16 cumulative bins to 3R, loss the sum
# of squared residuals against a seeded perturbation of the forward profile
ADJOINT = dict(n_side=256, box=500.0, n_cells=256, r_smooth=3.0, n_centres=100_000,
               r_lo=5.0, r_hi=40.0, ext=3.0, n_bins=16, seed=4, centre_seed=41,
               radius_seed=42, target_seed=43)
# the same chain on the 1024^3 set of plain1024, Gaussian R = 6 Mpc/h, at the
# 2e5 largest voids of its catalogue
ADJOINT1024 = dict(n_side=1024, box=2000.0, n_cells=1024, r_smooth=6.0, n_centres=200_000,
                   ext=3.0, n_bins=16, seed=5, target_seed=53)


class AdjointChain:
    """The chain on device-resident arrays through the C ABI: forward
    (deposit, contrast, smoothing, profiles, loss) and adjoint (the loss's
    gradient, the profiles' transpose, the same smoothing, the deposit's
    transpose). The adjoint stores no field of the forward."""

    def __init__(self, a=ADJOINT, positions=None, voids=None):
        """positions: the particles (x, y, z) when given, else a's seeded set;
        voids: (n_centres, 4) centres and radii when given, else a's seeded
        uniform centres and radii."""
        import ctypes as C
        from sif_amd import _native as N, zeldovich
        from sif_amd._native import DeviceArray
        self.C, self.N, self.a = C, N, a
        if positions is None:
            x, y, z = zeldovich.generate(a["n_side"], a["box"], a["seed"])
        else:
            x, y, z = positions
        self.n = len(x)
        if voids is None:
            rng = np.random.default_rng(a["centre_seed"])
            cen = rng.uniform(0, a["box"], (a["n_centres"], 3)).astype(np.float32)
            rv = np.random.default_rng(a["radius_seed"]).uniform(
                a["r_lo"], a["r_hi"], a["n_centres"]).astype(np.float32)
        else:
            cen, rv = voids[:, :3].astype(np.float32), voids[:, 3].astype(np.float32)
        up = lambda v: DeviceArray.from_host(np.ascontiguousarray(v, np.float32))
        self.x, self.y, self.z = up(x), up(y), up(z)
        self.cx, self.cy, self.cz, self.rv = up(cen[:, 0]), up(cen[:, 1]), up(cen[:, 2]), up(rv)
        nc3 = a["n_cells"] ** 3
        spec = int(N.lib.sif_amd_spectrum_bytes(a["n_cells"]))
        work = max(spec, int(N.lib.sif_amd_cic_work_bytes(a["n_cells"])))
        self.grid = DeviceArray((nc3,), np.float32)
        self.field = DeviceArray((nc3,), np.float32)
        self.spec = DeviceArray((spec,), np.uint8)
        self.work = DeviceArray((work,), np.uint8)
        npr = a["n_centres"] * a["n_bins"]
        self.p = DeviceArray((npr,), np.float32)
        self.gp = DeviceArray((npr,), np.float32)
        self.gx, self.gy, self.gz = (DeviceArray((self.n,), np.float32) for _ in range(3))
        self.target = self.p  # the first forward's profiles, perturbed below
        self.forward()
        p0 = self.p.get()
        noise = np.random.default_rng(a["target_seed"]).normal(0.0, 0.1, p0.shape)
        self.target = DeviceArray.from_host((p0 + noise).astype(np.float32))

    def _smooth(self, src, dst):
        a, N = self.a, self.N
        N.check(N.lib.sif_amd_fft_r2c(src.ptr, a["n_cells"], self.spec.ptr, self.work.ptr))
        N.check(N.lib.sif_amd_fft_filter_c2r(self.spec.ptr, a["n_cells"], a["box"],
                                             N.FILTER_GAUSSIAN, a["r_smooth"], 0.0, 0, 0,
                                             dst.ptr, self.work.ptr))

    def forward(self):
        a, N, C = self.a, self.N, self.C
        N.check(N.lib.sif_amd_grid_assign_cic(self.x.ptr, self.y.ptr, self.z.ptr, None, self.n,
                                              a["n_cells"], a["box"], self.grid.ptr,
                                              self.work.ptr))
        mean = C.c_double()
        N.check(N.lib.sif_amd_grid_to_density_contrast(self.grid.ptr, a["n_cells"] ** 3,
                                                       C.byref(mean)))
        self._smooth(self.grid, self.field)
        N.check(N.lib.sif_amd_grid_profiles(self.field.ptr, a["n_cells"], a["box"], self.cx.ptr,
                                            self.cy.ptr, self.cz.ptr, self.rv.ptr,
                                            a["n_centres"], a["ext"], a["n_bins"], self.p.ptr))
        self.mean = mean.value
        return self.loss(None)

    def loss(self, grad):
        N, C = self.N, self.C
        loss = C.c_double()
        N.check(N.lib.sif_amd_squared_residual(self.p.ptr, self.target.ptr, self.p.shape[0],
                                               1.0 / self.mean, grad.ptr if grad else None,
                                               C.byref(loss)))
        return loss.value

    def forward_adjoint(self):
        a, N = self.a, self.N
        self.forward()
        loss = self.loss(self.gp)  # dL/dp over the contrast's mean
        N.check(N.lib.sif_amd_grid_profiles_adjoint(self.gp.ptr, a["n_cells"], a["box"],
                                                    self.cx.ptr, self.cy.ptr, self.cz.ptr,
                                                    self.rv.ptr, a["n_centres"], a["ext"],
                                                    a["n_bins"], self.field.ptr))
        self._smooth(self.field, self.grid)
        N.check(N.lib.sif_amd_cic_adjoint(self.x.ptr, self.y.ptr, self.z.ptr, self.n,
                                          a["n_cells"], a["box"], self.grid.ptr, self.gx.ptr,
                                          self.gy.ptr, self.gz.ptr))
        return loss


def adjoint_form(steps, warmup, a=ADJOINT, positions=None, voids=None, what=None):
    """Forward and forward+adjoint times of the profile chain, positions
    resident in HBM, each step synchronised."""
    from sif_amd import _native as N
    t0 = time.time()
    ch = AdjointChain(a, positions, voids)
    log(f"[bench] adjoint chain set up in {time.time() - t0:.1f}s")

    def timed(fn):
        for _ in range(warmup):
            fn()
        N.check(N.lib.sif_amd_synchronize())
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        N.check(N.lib.sif_amd_synchronize())
        return (time.perf_counter() - t) / steps, out

    t_fwd, loss = timed(ch.forward)
    t_both, _ = timed(ch.forward_adjoint)
    g = np.concatenate([v.get()[:4] for v in (ch.gx, ch.gy, ch.gz)])
    centres = what or (f"{a['n_centres']} centres radii {a['r_lo']:g}-{a['r_hi']:g}")
    log(f"[bench] adjoint chain: forward {t_fwd * 1e3:.2f} ms, forward+adjoint "
        f"{t_both * 1e3:.2f} ms, loss {loss:.6g}")
    return {"metric": "profile-chain particles/s, forward + adjoint",
            "value": round(ch.n / t_both, 1), "unit": "particles/s",
            "ms_forward": round(t_fwd * 1e3, 3), "ms_forward_adjoint": round(t_both * 1e3, 3),
            "ratio_forward_adjoint_to_forward": round(t_both / t_fwd, 3),
            "loss": loss, "grad_sample": [float(v) for v in g],
            "config": {"workload": f"{a['n_side']}^3 Zel'dovich particles, L={a['box']:g}, "
                                   f"grid {a['n_cells']}^3, Gaussian R={a['r_smooth']:g}, "
                                   f"{centres}, "
                                   f"{a['n_bins']} cumulative bins to {a['ext']:g}R"}}


def dropin_form(c, x, y, z, radii, mesh_cells, want_voids):
    """The drop-in path through the host structs of include/sif/ as a caller
    of sif v0.5.0 uses it (sif_field_*, sif_grid_assign_cic,
    sif_grid_to_density_contrast, sif_chain_mesh_alloc, sif_finder_exodus_raw),
    from host arrays to a host catalogue: every call moves its inputs to the
    device and its results back (PCIe included), one run."""
    import sif_amd.pysif as pysif
    from sif_amd.pysif import finders
    pysif.system_init()
    t = time.perf_counter()
    field = pysif.field_from_numpy(x, y, z)
    grid = pysif.Grid(c["n_cells"], c["box"])
    grid.assign_cic(field)
    grid.to_density_contrast()
    mesh = pysif.ChainMesh(mesh_cells, c["box"], field, drop_indices=True)
    cat = finders.exodus_raw(grid, mesh, radii, THRESHOLD, 0.0, checkpoint=False,
                             consume_grid=True)
    dt = time.perf_counter() - t
    log(f"[bench] drop-in (host structs): {dt:.2f}s, {cat.n_voids} voids")
    if cat.n_voids != want_voids:
        raise RuntimeError(f"drop-in catalogue has {cat.n_voids} voids, the device path "
                           f"{want_voids}")
    return {"metric": "exodus particles/s through the host-struct C API (PCIe included)",
            "value": round(len(x) / dt, 1), "unit": "particles/s", "seconds": round(dt, 3),
            "voids": int(cat.n_voids),
            "config": {"workload": "the same particles from host arrays: sif_field, "
                                   "sif_grid_assign_cic, sif_grid_to_density_contrast, "
                                   "sif_chain_mesh_alloc, sif_finder_exodus_raw, one run"}}


def profiles_form(step, cat, steps, warmup):
    """sif_profiles over the plain catalogue: every void, the reference's
    default reach of 5 void radii, 16 bins, density and radial velocity
    (seeded velocities), over the same 512^3 tracers on their mesh."""
    from sif_amd import _native as N
    D = step.D
    c = step.c
    mesh = D.Mesh(step.dx, step.dy, step.dz, step.n, step.mesh_cells, c["box"])
    o = mesh.orig.get().astype(np.int64)
    rng = np.random.default_rng(77)
    vel = tuple(D.to_device(rng.normal(0.0, 300.0, step.n).astype(np.float32)[o])
                for _ in range(3))
    del o

    def run():
        return D.tracer_profiles(mesh, cat, 5.0, 16, 0, velocities=vel)

    for _ in range(warmup):
        run()
    N.check(N.lib.sif_amd_synchronize())
    t = time.perf_counter()
    for _ in range(steps):
        dens, v = run()
    N.check(N.lib.sif_amd_synchronize())
    dt = (time.perf_counter() - t) / steps
    log(f"[bench] profiles of {len(cat)} voids: {dt * 1e3:.1f} ms")
    return {"metric": "stacked profiles of the plain catalogue, voids/s", "unit": "voids/s",
            "value": round(len(cat) / dt, 1), "ms": round(dt * 1e3, 2),
            "config": {"workload": f"{len(cat)} voids, {step.n} tracers on the {step.mesh_cells}^3 "
                                   f"mesh, ext 5, 16 bins, density and radial velocity"}}


# many boxes of uneven size: 96 seeded Zel'dovich sets, N^3 particles with N
# log-uniform in 64-256 (a multiple of 8), boxes at the fixed mean separation
# 100/64 Mpc/h (100-400 Mpc/h a side), each on the grid matching its particle
# lattice (N^3 cells), mesh cell 10 Mpc/h, radii from 3 Mpc/h to
# min(60, L/4) Mpc/h (the growth template must stay inside the mesh),
# threshold -0.8; box b uses seed 3000 + b
BATCHED = dict(n_boxes=96, side_lo=64, side_hi=256, spacing=100.0 / 64.0, mesh_cell=10.0,
               r_lo=3.0, r_hi=60.0, seed0=3000, size_seed=3)


def batched_boxes():
    a = BATCHED
    rng = np.random.default_rng(a["size_seed"])
    sides = np.exp(rng.uniform(np.log(a["side_lo"]), np.log(a["side_hi"]), a["n_boxes"]))
    out = []
    for b, side in enumerate(sides):
        n_side = int(min(max(round(side / 8.0) * 8, a["side_lo"]), a["side_hi"]))
        box = n_side * a["spacing"]
        mesh_cells = int(round(box / a["mesh_cell"]))
        r_hi = min(a["r_hi"], box / 4.0)
        out.append(dict(seed=a["seed0"] + b, n_side=n_side, box=box, n_cells=n_side,
                        mesh_cells=mesh_cells, r_hi=r_hi))
    return out


def batched_form():
    """Sum of particles over the 96 boxes / the time of all their catalogues
    (deposit, contrast, mesh, every rung), positions resident in HBM, through
    sif_amd_exodus_batch: the boxes side by side under the tuning the
    autotuner picked (before the timed region); the same boxes one after
    another on one worker, timed beside it."""
    from sif_amd import device as D, zeldovich
    t0 = time.time()
    boxes = batched_boxes()
    for bx in boxes:
        x, y, z = zeldovich.generate(bx["n_side"], bx["box"], bx["seed"])
        bx["n"] = len(x)
        bx["d"] = (D.to_device(x), D.to_device(y), D.to_device(z))
        bx["radii"] = zeldovich.ladder(BATCHED["r_lo"], bx["r_hi"], 0.05)
    total = sum(bx["n"] for bx in boxes)
    log(f"[bench] batched: {len(boxes)} boxes, {total} particles "
        f"({min(bx['n_side'] for bx in boxes)}^3..{max(bx['n_side'] for bx in boxes)}^3, "
        f"grids matching), set up in {time.time() - t0:.1f}s")
    t1 = time.perf_counter()
    tuning = D.batch_autotune(boxes, THRESHOLD, probe=12)
    t_tune = time.perf_counter() - t1
    log(f"[bench] batched: autotuned in {t_tune:.1f}s: {tuning}")
    D.exodus_batch(boxes, THRESHOLD, tuning=tuning)  # every worker's buffers at full size
    t = time.perf_counter()
    cats = D.exodus_batch(boxes, THRESHOLD, tuning=tuning)
    dt = time.perf_counter() - t
    voids = sum(len(c) for c in cats)
    log(f"[bench] batched: {dt:.2f}s for {len(boxes)} catalogues side by side, {voids} voids")
    serial = dict(tuning, workers=1)
    D.exodus_batch(boxes[:1], THRESHOLD, tuning=serial)
    t = time.perf_counter()
    cats1 = D.exodus_batch(boxes, THRESHOLD, tuning=serial)
    dt1 = time.perf_counter() - t
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(cats, cats1))
    log(f"[bench] batched: {dt1:.2f}s one after another, catalogues identical: {same}")
    if not same:
        raise RuntimeError("the batched catalogues differ from the boxes run one by one")
    return {"metric": "exodus particles/s over 96 boxes of uneven size",
            "value": round(total / dt, 1), "unit": "particles/s", "seconds": round(dt, 3),
            "seconds_one_after_another": round(dt1, 3), "voids": int(voids),
            "particles": int(total), "tuning": tuning, "autotune_seconds": round(t_tune, 2),
            "config": {"workload": "96 Zel'dovich boxes, N^3 log-uniform in 64^3-256^3 "
                                   "(N a multiple of 8), L = N * 100/64 Mpc/h, grid N^3 (the "
                                   "particle lattice), mesh cell 10 Mpc/h, radii "
                                   "3-min(60, L/4) Mpc/h, t=-0.8, sif_amd_exodus_batch"}}


SURVEY = dict(box=1000.0, n_cells=256, centre=500.0, foot_r=380.0, n_data=2_000_000,
              n_randoms=20_000_000, n_holes=300, hole_lo=10.0, hole_hi=35.0, kept=0.03,
              r_lo=8.0, r_hi=40.0, seed=17)


def survey_tracers(a, rng, n, carve, holes):
    """n points uniform over a ball footprint; with `carve`, all but a few
    per cent of what falls in a hole dropped."""
    out, got = [], 0
    while got < n:
        m = 2 * (n - got) + 65536
        p = a["centre"] + (2.0 * rng.random((m, 3)) - 1.0) * a["foot_r"]
        keep = ((p - a["centre"]) ** 2).sum(axis=1) <= a["foot_r"] ** 2
        p = p[keep]
        if carve:
            from scipy.spatial import cKDTree
            hit = np.zeros(len(p), bool)
            tree = cKDTree(p)
            for h in holes:
                hit[tree.query_ball_point(h[:3], h[3])] = True
            p = p[~hit | (rng.random(len(p)) < a["kept"])]
        p = p[: n - got]
        out.append(p.astype(np.float32))
        got += len(p)
    p = np.vstack(out)
    return [np.ascontiguousarray(p[:, k]) for k in range(3)]


def survey_form(steps, warmup):
    """The survey finder (sif_amd_exodus_survey_device) on a ball footprint
    of 2 M galaxies and 20 M randoms in a 1000 Mpc/h box, 300 carved holes:
    both CIC grids, both meshes and the run (occupancy, both fields smoothed
    each rung, the two-mesh growth, footprints) per step, positions resident
    in HBM."""
    from sif_amd import _native as N, device as D, zeldovich
    a = SURVEY
    rng = np.random.default_rng(a["seed"])
    t0 = time.time()
    r = a["foot_r"] - a["hole_hi"] - 20.0
    hc = a["centre"] + rng.uniform(-1.0, 1.0, (4 * a["n_holes"], 3)) * r
    hc = hc[((hc - a["centre"]) ** 2).sum(axis=1) <= r * r][: a["n_holes"]]
    holes = np.c_[hc, rng.uniform(a["hole_lo"], a["hole_hi"], len(hc))]
    data = survey_tracers(a, rng, a["n_data"], True, holes)
    randoms = survey_tracers(a, rng, a["n_randoms"], False, holes)
    radii = zeldovich.ladder(a["r_lo"], a["r_hi"], 0.05)
    n, box = a["n_cells"], a["box"]
    dd = [D.to_device(v) for v in data]
    rd = [D.to_device(v) for v in randoms]
    md = max(8, int(min((a["n_data"] / 30.0) ** (1.0 / 3.0), 1024)))
    mr = max(8, int(min((a["n_randoms"] / 30.0) ** (1.0 / 3.0), 1024)))
    log(f"[bench] survey: {a['n_data']} galaxies, {a['n_randoms']} randoms, {len(holes)} holes, "
        f"set up in {time.time() - t0:.1f}s")

    def run():
        gd = D.cic(*dd, a["n_data"], n, box)
        gr = D.cic(*rd, a["n_randoms"], n, box)
        mesh_d = D.Mesh(*dd, a["n_data"], md, box, indices=False)
        mesh_r = D.Mesh(*rd, a["n_randoms"], mr, box, indices=False)
        cat, fp, _, _ = D.exodus_survey(gd, gr, n, box, mesh_d, mesh_r, radii, THRESHOLD,
                                        restore=False)
        return cat, fp

    for _ in range(warmup):
        cat, fp = run()
    N.check(N.lib.sif_amd_synchronize())
    t = time.perf_counter()
    for _ in range(steps):
        cat, fp = run()
    N.check(N.lib.sif_amd_synchronize())
    dt = (time.perf_counter() - t) / steps
    found = 0
    for h in holes:
        d = np.sqrt(((cat[:, :3] - h[:3]) ** 2).sum(axis=1)) if len(cat) else np.array([np.inf])
        found += d.min() <= 0.5 * h[3]
    log(f"[bench] survey: {dt * 1e3:.1f} ms a catalogue, {len(cat)} voids, {found} of "
        f"{len(holes)} holes within half their radius, mean footprint {fp[:, 0].mean():.3f}")
    return {"metric": "survey exodus galaxies/s", "value": round(a["n_data"] / dt, 1),
            "unit": "galaxies/s", "seconds": round(dt, 4), "voids": int(len(cat)),
            "holes_recovered": int(found), "holes": int(len(holes)),
            "config": {"workload": f"ball footprint r={a['foot_r']:g} in L={box:g}, "
                                   f"{a['n_data']} galaxies (300 carved holes of r "
                                   f"{a['hole_lo']:g}-{a['hole_hi']:g}), {a['n_randoms']} "
                                   f"randoms, grid {n}^3, meshes {md}^3/{mr}^3, "
                                   f"{len(radii)} rungs {a['r_lo']:g}-{a['r_hi']:g}, "
                                   f"t={THRESHOLD}, sif_amd_exodus_survey_device"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="plain512", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable profile chain's leg")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 96-box leg")
    ap.add_argument("--no-survey", action="store_true",
                    help="skip the survey finder's leg")
    ap.add_argument("--no-forms", action="store_true",
                    help="time the plain catalogue only (for profiling it)")
    ap.add_argument("--forms-only", action="store_true",
                    help="run only the adjoint and batched legs (for profiling them)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device, 4 times: the known traffic that "
                         "scales rocprof's FETCH_SIZE and WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU figures are not part of "
              "it; one box would shard by x-slab (CIC, scan) and pencil (FFT)", file=sys.stderr)
        sys.exit(2)

    from sif_amd import _native as N
    if not N.has_gpu():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        from sif_amd._native import DeviceArray
        src = DeviceArray((1 << 28,), np.float32)
        dst = DeviceArray((1 << 28,), np.float32)
        N.check(N.lib.sif_amd_memset(src.ptr, 1, 1 << 30))
        for _ in range(4):
            N.check(N.lib.sif_amd_copy(dst.ptr, src.ptr, 1 << 30))
        log("[bench] copy probe: 4 launches of copy_kernel, 1073741824 B read and written each")
        return

    if args.forms_only:
        forms = {}
        if not args.no_adjoint:
            forms["adjoint"] = adjoint_form(args.steps, args.warmup)
        if not args.no_batched:
            forms["batched"] = batched_form()
        if not args.no_survey:
            forms["survey"] = survey_form(args.steps, args.warmup)
        print(json.dumps({"forms": forms}), flush=True)
        return

    c, x, y, z, radii, mesh_cells = workload(args.config)
    step = Step(c, x, y, z, radii, mesh_cells)
    log(f"[bench] ladder {len(radii)} rungs {radii[0]:.2f}..{radii[-1]:.2f}, mesh {mesh_cells}^3")
    for w in range(args.warmup):
        t0 = time.time()
        cat, stats = step.run()
        log(f"[bench] warmup {w}: {time.time() - t0:.2f}s, {len(cat)} voids")
    # the timed region, with the stage events on (recorded without
    # synchronising; read after the region)
    N.check(N.lib.sif_amd_timing_enable(1))
    N.lib.sif_amd_timing_reset()
    N.check(N.lib.sif_amd_synchronize())
    t0 = time.perf_counter()
    for s in range(args.steps):
        cat, stats = step.run()
        log(f"[bench] step {s} done")
    N.check(N.lib.sif_amd_synchronize())
    dt = time.perf_counter() - t0
    value = step.n * args.steps / dt
    timing = N.timing_read()
    N.check(N.lib.sif_amd_timing_enable(0))
    tot = {k: sum(s[k] for s in stats) for k in stats[0] if k != "radius"}
    log("[bench] candidates over all rungs: " + ", ".join(f"{k} {v}" for k, v in tot.items()))
    log("[bench] stage times over the timed steps (ms, launches): " +
        ", ".join(f"{k} {v[0]:.1f}/{v[1]}" for k, v in timing.items()))
    roof = roofline(timing, args.config)
    import ctypes as C
    used, total = C.c_uint64(), C.c_uint64()
    N.check(N.lib.sif_amd_device_memory(C.byref(used), C.byref(total)))
    log(f"[bench] device memory in use after the steps: {used.value / 1e9:.1f} GB of "
        f"{total.value / 1e9:.1f} GB")

    forms = {}
    if args.config == "plain1024":
        # the chain on this set: the catalogue's largest voids as centres
        a = ADJOINT1024
        big = cat[np.argsort(-cat[:, 3], kind="stable")[:a["n_centres"]]]
        a = dict(a, n_centres=len(big))
        del step
        if not args.no_adjoint:
            forms["adjoint"] = adjoint_form(
                args.steps, args.warmup, a, (x, y, z), big,
                f"the {len(big)} largest voids of the plain catalogue "
                f"(radii {big[-1, 3]:.2f}-{big[0, 3]:.2f})")
    elif not args.no_forms:
        forms["profiles"] = profiles_form(step, cat, args.steps, args.warmup)
        del step
        forms["dropin"] = dropin_form(c, x, y, z, radii, mesh_cells, len(cat))
        if not args.no_adjoint:
            forms["adjoint"] = adjoint_form(args.steps, args.warmup)
        if not args.no_batched:
            forms["batched"] = batched_form()
        if not args.no_survey:
            forms["survey"] = survey_form(args.steps, args.warmup)

    cpu = None
    if not args.no_cpu:
        t1 = time.time()
        cpu = cpu_baseline(c, x, y, z, radii, mesh_cells, cat)
        log(f"[bench] cpu baseline {cpu['value']:.3g} particles/s in {time.time() - t1:.1f}s, "
            f"against the GPU's catalogue: {cpu['against_gpu']}")

    out = {
        "metric": "exodus particles/sec at 512^3 (plain vs batched/adjoint) + % HBM roofline",
        "value": round(value, 1), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Zel'dovich/BBKS LCDM particle set)",
        "config": {"workload": f"plain exodus, {c['n_side']}^3 particles, L={c['box']:g}, "
                               f"grid {c['n_cells']}^3, mesh {mesh_cells}^3, t={THRESHOLD}, "
                               f"{len(radii)} rungs {radii[-1]:.3f}-{radii[0]:.2f}",
                   "voids": int(len(cat)), "parallelism": "single GPU"},
        "stages_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in timing.items()},
        "hbm_footprint_gb": round(used.value / 1e9, 2),
        "roofline": roof,
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
