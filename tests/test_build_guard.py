"""The build's check of the device assembly (sif_amd/build.py): a device
function that jumps through s[30:31], its return-address pair, loops on
return. Constructed listings, and the library as built."""
import glob
import os

from sif_amd import build

HAZARD = """
\t.type\t_Z3fooPf,@function
_Z3fooPf:                               ; @_Z3fooPf
\ts_getpc_b64 s[30:31]
\ts_add_u32 s30, s30, (.LBB0_9-.Lpost_getpc0)&4294967295
\ts_setpc_b64 s[30:31]
.LBB0_9:
\ts_setpc_b64 s[30:31]
"""
CLEAN_FUNCTION = """
\t.type\t_Z3barPf,@function
_Z3barPf:                               ; @_Z3barPf
\ts_getpc_b64 s[4:5]
\ts_setpc_b64 s[4:5]
\ts_setpc_b64 s[30:31]
"""
KERNEL = """
\t.type\t_Z6kernelPf,@function
_Z6kernelPf:                            ; @_Z6kernelPf
\ts_getpc_b64 s[30:31]
\ts_setpc_b64 s[30:31]
\ts_endpgm
\t.amdhsa_kernel _Z6kernelPf
"""


def test_guard_finds_a_long_branch_through_the_return_address():
    assert build.return_address_hazards(HAZARD) == ["_Z3fooPf"]
    assert build.return_address_hazards(CLEAN_FUNCTION) == []
    assert build.return_address_hazards(KERNEL + CLEAN_FUNCTION) == []
    assert build.return_address_hazards(CLEAN_FUNCTION + HAZARD) == ["_Z3fooPf"]


def test_built_device_code_has_none():
    listings = glob.glob(os.path.join(build.OBJ, "*-hip-amdgcn-amd-amdhsa-*.s"))
    if not listings:
        import pytest
        pytest.skip("the library has not been built here")
    functions = 0
    for path in listings:
        text = open(path).read()
        functions += text.count("@function")
        assert build.return_address_hazards(text) == [], path
    assert functions > 50  # the listings were read
