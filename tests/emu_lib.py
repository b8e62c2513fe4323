"""TEST INFRASTRUCTURE ONLY: the CPU builds of the library's device code
under tests/emu. SIF_EMU_VARIANT=asan selects the builds with the address and
undefined-behaviour sanitizers (tests/test_host_sanitize.py runs the
emulated tests that way, in a child process that preloads the runtime)."""
import ctypes as C
import os
import subprocess

EMU_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
VARIANT = os.environ.get("SIF_EMU_VARIANT", "")


def path(name):
    """Builds (make) and returns the path of lib<name>_emu[_asan].so."""
    suffix = f"_{VARIANT}" if VARIANT else ""
    target = f"lib{name}_emu{suffix}.so"
    subprocess.run(["make", "-s", "-C", EMU_DIR, target], check=True)
    return os.path.join(EMU_DIR, target)


def load(name):
    return C.CDLL(path(name))
