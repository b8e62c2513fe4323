"""The density contrast through the entry point that allocates nothing
(sif_amd_grid_to_density_contrast_noalloc) against the one that does
(sif_amd_grid_to_density_contrast), through the C ABI: two calls in a row,
the second on a smaller grid, each bit for bit (mean and contrast); sizes
past the fixed scratch are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(native):
    lib = native.lib
    fn = lib.sif_amd_grid_to_density_contrast_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]
    native.check(lib.sif_system_init(None))
    return native


def _grid(n, seed):
    from sif_amd import device as D
    from sif_amd import zeldovich
    x, y, z = zeldovich.generate(n, 100.0, seed)
    dx, dy, dz = D.to_device(x), D.to_device(y), D.to_device(z)
    return D.cic(dx, dy, dz, len(x), n, 100.0).get()


def test_noalloc_density_contrast_equals_allocating_one(lib):
    from sif_amd import device as D
    for n, seed in ((96, 3), (40, 4)):  # the second call smaller than the first
        v = _grid(n, seed)
        a, b = D.to_device(v), D.to_device(v)
        want_mean = D.density_contrast(a, n ** 3)
        mean = C.c_double()
        lib.check(lib.lib.sif_amd_grid_to_density_contrast_noalloc(b.ptr, n ** 3,
                                                                   C.byref(mean)))
        assert mean.value == want_mean
        np.testing.assert_array_equal(b.get().view(np.uint32), a.get().view(np.uint32))


def test_noalloc_refuses_a_grid_past_its_scratch(lib):
    from sif_amd._native import DeviceArray
    d = DeviceArray((1024,), np.float32, zero=True)
    mean = C.c_double()
    too_many = 32768 * 65536 + 1  # refused before any kernel reads the grid
    st = lib.lib.sif_amd_grid_to_density_contrast_noalloc(d.ptr, too_many, C.byref(mean))
    assert st == lib.SIF_ERR_RANGE
    assert b"without allocation" in lib.lib.sif_error_message()
