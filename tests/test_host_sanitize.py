"""The host side of the commit and the emulated device code under the
sanitizers, on the CPU.

- tests/emu/host_tsan.cpp drives the commit's host code (the catalogue's
  cell list and scan, the host pool, the batch's pair effects) as the
  driver does, built with the thread sanitizer, and with the address and
  undefined-behaviour sanitizers;
- the emulated-device tests (growth kernel, overlap states, painting and the
  per-rung probe) run again in a child process on builds of their libraries
  with the address and undefined-behaviour sanitizers.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")


def _make(target):
    subprocess.run(["make", "-s", "-C", EMU_DIR, target], check=True)


@pytest.mark.parametrize("prog", ["host_tsan", "host_asan"])
def test_commit_host_code_under_sanitizer(prog):
    _make(prog)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([os.path.join(EMU_DIR, prog), "7", "30"], capture_output=True,
                         text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "host harness ok" in res.stdout
    assert "Sanitizer" not in res.stderr, res.stderr[-4000:]


def test_emulated_device_code_under_asan_ubsan():
    _make("asan")
    runtime = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, check=True).stdout.strip()
    env = dict(os.environ, SIF_EMU_VARIANT="asan", LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0")
    tests = [os.path.join(HERE, t) for t in ("test_grow_emulated.py", "test_overlap_split.py",
                                            "test_paint.py", "test_batch_pairs.py")]
    res = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                          "-m", "not gpu", *tests], capture_output=True, text=True, env=env,
                         cwd=os.path.dirname(HERE), timeout=900)
    tail = (res.stdout + res.stderr)[-4000:]
    assert res.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, tail
