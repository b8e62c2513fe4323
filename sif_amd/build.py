"""Builds the sif-amd HIP library in-tree: sif_amd/lib/libsif_amd.so.

Every .hip file under sif_amd/csrc is compiled for gfx950 (one object each, in
parallel) and linked into one shared library exporting the C ABI of
include/sif/ and include/sif_amd.h.

    python -m sif_amd.build
"""
import concurrent.futures
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(HERE, "lib", "libsif_amd.so")

ARCH = os.environ.get("SIF_AMD_ARCH", "gfx950")

# -ffp-contract=off: every float expression of the path is evaluated with one
# rounding per operation, as the reference's decisions assume; -packed-fp32-ops
# keeps the compiler from emitting packed FP32 instructions. The host side
# targets x86-64-v2 (SSE4.2, POPCNT): floorf and the bit counts of the host's
# cell walks become single instructions rather than library calls.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xarch_host", "-march=x86-64-v2",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required")
    return path


LABEL = re.compile(r"^([A-Za-z_.$][\w.$]*):")


def return_address_hazards(asm_text):
    """Device functions (not kernels) that jump through s[30:31], the
    register pair holding their return address: the long-branch sequence
    `s_getpc_b64 s[30:31]; ...; s_setpc_b64 s[30:31]` that the compiler emits
    for a branch past the short branch's reach overwrites it, and the
    function's own return then jumps back into itself, forever. Returns the
    names of such functions in an AMDGPU assembly listing."""
    bad, name, is_func, body = [], None, False, []
    functions = set()
    for line in asm_text.splitlines():
        t = line.strip()
        if t.startswith(".type") and t.endswith("@function"):
            functions.add(t.split()[1].split(",")[0])
            continue
        m = LABEL.match(t)
        head = m.group(1) if m else None
        if head and head in functions:
            if is_func and "s_getpc_b64 s[30:31]" in body and "s_setpc_b64 s[30:31]" in body:
                bad.append(name)
            name, body = head, []
            is_func = True
            continue
        if t.startswith(".amdhsa_kernel"):
            is_func = False  # a kernel's descriptor: the function before it was the kernel
        if is_func:
            body.append(" ".join(t.split()))
    if is_func and "s_getpc_b64 s[30:31]" in body and "s_setpc_b64 s[30:31]" in body:
        bad.append(name)
    return bad


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    # -save-temps=obj leaves the device assembly beside the object (under
    # lib/obj, not shipped): it is checked for the hazard above
    cmd = [hipcc(), *FLAGS, "-save-temps=obj", "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    base = os.path.splitext(os.path.basename(src))[0]
    asm = os.path.join(OBJ, f"{base}-hip-amdgcn-amd-amdhsa-{ARCH}.s")
    if os.path.exists(asm):
        bad = return_address_hazards(open(asm).read())
        if bad:
            raise RuntimeError(f"{src}: device functions that branch through their return "
                               f"address (hang on return): {', '.join(bad)}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
