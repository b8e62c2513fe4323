"""The differentiable profile chain on the GPU, as a torch.autograd.Function:

    positions -> CIC density (sif_amd_grid_assign_cic)
              -> density contrast (sif_amd_grid_to_density_contrast)
              -> Gaussian smoothing (sif_amd_fft_r2c + sif_amd_fft_filter_c2r)
              -> cumulative spherically averaged density profiles at fixed
                 centres (sif_amd_grid_profiles; sif_profiles,
                 src/measure/profiles.c:354-673, on the smoothed grid)

and back through hand-written adjoints: sif_amd_grid_profiles_adjoint, the
same smoothing (a real, even filter is its own transpose), and
sif_amd_cic_adjoint. The backward pass stores no field and recomputes none:
after the deposit the chain is linear, so it needs only the positions and
the contrast's mean.

torch is the plumbing (device memory, the autograd tape). Every stage runs in
the HIP library, on the library's own stream; each call waits for torch's
stream first, and torch's next use waits for the library's stream.
"""
import ctypes as C

import torch

from . import _native as N

if not N.TORCH_FIRST:
    raise ImportError(
        "sif_amd.autograd: import torch (or sif_amd.autograd) before any other sif_amd "
        "module that loads the HIP library: torch brings its own HIP runtime, which the "
        "library must share")

FILTER_GAUSSIAN = 2


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class ProfileChain:
    """Fixed geometry of the chain: an n_cells^3 grid over a periodic box,
    Gaussian smoothing of radius r_smooth, centres (cx, cy, cz) of radius rv
    (float32 device tensors), profiles of n_bins cumulative bins out to
    ext * rv. Scratch is allocated once, on torch's allocator."""

    def __init__(self, n_cells, box, r_smooth, cx, cy, cz, rv, ext=3.0, n_bins=16):
        if not N.has_gpu():
            raise RuntimeError("sif_amd.autograd: no HIP device; the chain has no CPU path")
        self.n_cells, self.box, self.r = int(n_cells), float(box), float(r_smooth)
        self.ext, self.n_bins = float(ext), int(n_bins)
        dev = cx.device
        self.cx, self.cy, self.cz, self.rv = (t.detach().to(torch.float32).contiguous()
                                              for t in (cx, cy, cz, rv))
        self.n_centres = self.cx.numel()
        total = self.n_cells ** 3
        spec = int(N.lib.sif_amd_spectrum_bytes(self.n_cells))
        work = max(spec, int(N.lib.sif_amd_cic_work_bytes(self.n_cells)))
        self.grid = torch.empty(total, dtype=torch.float32, device=dev)
        self.field = torch.empty(total, dtype=torch.float32, device=dev)
        self.spec = torch.empty(spec, dtype=torch.uint8, device=dev)
        self.work = torch.empty(work, dtype=torch.uint8, device=dev)

    def _centres(self):
        return _ptr(self.cx), _ptr(self.cy), _ptr(self.cz), _ptr(self.rv)

    def _smooth(self, src, dst):
        N.check(N.lib.sif_amd_fft_r2c(_ptr(src), self.n_cells, _ptr(self.spec), _ptr(self.work)))
        N.check(N.lib.sif_amd_fft_filter_c2r(_ptr(self.spec), self.n_cells, self.box,
                                             FILTER_GAUSSIAN, self.r, 0.0, 0, 0, _ptr(dst),
                                             _ptr(self.work)))

    def smoothed_contrast(self, x, y, z):
        """The smoothed density contrast (in self.field) and the mean density."""
        torch.cuda.current_stream().synchronize()
        N.check(N.lib.sif_amd_grid_assign_cic(_ptr(x), _ptr(y), _ptr(z), None, x.numel(),
                                              self.n_cells, self.box, _ptr(self.grid),
                                              _ptr(self.work)))
        mean = C.c_double()
        N.check(N.lib.sif_amd_grid_to_density_contrast(_ptr(self.grid), self.grid.numel(),
                                                       C.byref(mean)))
        self._smooth(self.grid, self.field)
        N.check(N.lib.sif_amd_synchronize())
        return self.field, mean.value

    def forward(self, x, y, z):
        """Profiles [n_centres, n_bins] and the mean density."""
        _, mean = self.smoothed_contrast(x, y, z)
        out = torch.empty((self.n_centres, self.n_bins), dtype=torch.float32, device=x.device)
        N.check(N.lib.sif_amd_grid_profiles(_ptr(self.field), self.n_cells, self.box,
                                            *self._centres(), self.n_centres, self.ext,
                                            self.n_bins, _ptr(out)))
        N.check(N.lib.sif_amd_synchronize())
        return out, mean

    def adjoint(self, x, y, z, grad_p, mean):
        """dL/d(x, y, z) from dL/d(profiles)."""
        # the contrast divides the density by its mean, which the deposit's
        # mass conservation holds fixed: fold 1/mean into the profiles' side
        g = (grad_p.to(torch.float32) / mean).contiguous()
        torch.cuda.current_stream().synchronize()
        N.check(N.lib.sif_amd_grid_profiles_adjoint(_ptr(g), self.n_cells, self.box,
                                                    *self._centres(), self.n_centres, self.ext,
                                                    self.n_bins, _ptr(self.field)))
        self._smooth(self.field, self.grid)
        gx, gy, gz = (torch.empty_like(x) for _ in range(3))
        N.check(N.lib.sif_amd_cic_adjoint(_ptr(x), _ptr(y), _ptr(z), x.numel(), self.n_cells,
                                          self.box, _ptr(self.grid), _ptr(gx), _ptr(gy),
                                          _ptr(gz)))
        N.check(N.lib.sif_amd_synchronize())
        return gx, gy, gz


class VoidProfiles(torch.autograd.Function):
    """profiles = VoidProfiles.apply(x, y, z, chain): the chain's profiles of
    the particles at (x, y, z) (float32 device tensors), differentiable in
    the positions."""

    @staticmethod
    def forward(ctx, x, y, z, chain):
        for t in (x, y, z):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("VoidProfiles: positions must be contiguous float32 "
                                 "device tensors")
        p, mean = chain.forward(x, y, z)
        ctx.save_for_backward(x, y, z)
        ctx.chain, ctx.mean = chain, mean
        return p

    @staticmethod
    def backward(ctx, grad_p):
        x, y, z = ctx.saved_tensors
        gx, gy, gz = ctx.chain.adjoint(x, y, z, grad_p, ctx.mean)
        return gx, gy, gz, None


def void_profiles(x, y, z, chain):
    return VoidProfiles.apply(x, y, z, chain)
