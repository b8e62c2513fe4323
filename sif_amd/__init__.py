"""sif-amd: the exodus void-finder hot path of sif, MI355X-native.

The C ABI lives in sif_amd/lib/libsif_amd.so (built from sif_amd/csrc/*.hip
by `python -m sif_amd.build`). Python entry points:

* sif_amd.device   -- the stages on device-resident arrays (include/sif_amd.h),
                      weighted meshes included
* sif_amd.pysif    -- the pysif surface for the path (Field, Grid, ChainMesh,
                      finders.exodus/exodus_raw/read_checkpoint) over the
                      host-struct C API
* sif_amd.autograd -- the void-profile chain as a torch.autograd.Function
                      (import torch first: the library then shares torch's
                      HIP runtime)
* sif_amd.zeldovich -- the seeded synthetic particle sets
"""
import os

__version__ = "0.1.0"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib",
                        "libsif_amd.so")
