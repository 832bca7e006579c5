"""Sanitizer builds of the CPU code (no GPU needed).

* The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
  (make -C oracle asan) runs the oracle's own CPU tests: the closed forms,
  the pinned time step, the preconditioner restatements, particles,
  periodic x, coefficient update and re-initialisation, and the golden
  digests.  Any out-of-bounds access, use after free or undefined operation
  stops the run.
* The C-ABI library with AddressSanitizer on its host code
  (make -C drd_application_amd/csrc asan) is driven on the paths that need
  no GPU: parameter defaults and field sizes for every field on odd, ragged
  and periodic grids, the argument checks of drd_init and the error string.

The instrumented processes run with libasan preloaded (Python itself is not
instrumented) and leak detection off (the interpreter keeps its own
allocations to the end).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "_build")


def _libasan():
    path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not path or not os.path.exists(path):
        pytest.skip("libasan is not installed")
    return path


def _env(extra):
    env = dict(os.environ)
    env.update({"LD_PRELOAD": _libasan(), "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "PYTHONPATH": ROOT})
    env.update(extra)
    return env


def test_oracle_under_address_and_undefined_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    lib = os.path.join(BUILD, "liboracle_asan.so")
    files = ["test_oracle_math.py", "test_oracle_step.py", "test_oracle_prec.py", "test_oracle_particles.py",
             "test_oracle_periodic.py", "test_oracle_coef.py", "test_oracle_reinit.py", "test_golden.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:xdist", "-p",
                        "no:cacheprovider"] + [os.path.join(ROOT, "tests", f) for f in files],
                       env=_env({"DRD_ORACLE_LIB": lib}), capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail


HOST_DRIVER = r'''
import ctypes as C, sys
from drd_application_amd import abi
L = C.CDLL(sys.argv[1])
L.drd_params_default.argtypes = [C.POINTER(abi.Params), C.c_int32, C.c_int32]
L.drd_field_size.argtypes = [C.POINTER(abi.Params), C.c_int32]
L.drd_field_size.restype = C.c_int64
L.drd_init.argtypes = [C.POINTER(abi.Params), C.POINTER(C.c_void_p)]
L.drd_last_error.restype = C.c_char_p
assert L.drd_abi_version() > 0
for nx, ny, lr in ((64, 32, "n"), (37, 11, "d"), (128, 64, "p"), (2, 2, "n")):
    p = abi.Params()
    L.drd_params_default(C.byref(p), nx, ny)
    p.lr_bd_type = ord(lr)
    sizes = [L.drd_field_size(C.byref(p), f) for f in range(abi.F_COUNT)]
    assert all(s > 0 for s in sizes), sizes
bad = abi.Params()
L.drd_params_default(C.byref(bad), 64, 32)
bad.nx = -4
h = C.c_void_p()
rc = L.drd_init(C.byref(bad), C.byref(h))
assert rc != 0 and h.value is None, rc
assert len(L.drd_last_error()) > 0
assert L.drd_init(None, C.byref(h)) != 0
print("host paths ok")
'''


def test_host_runtime_under_address_sanitizer():
    subprocess.run(["make", "-s", "-j", "8", "-C", os.path.join(ROOT, "drd_application_amd", "csrc"), "asan"],
                   check=True)
    lib = os.path.join(BUILD, "libdrd_gpu_asan.so")
    r = subprocess.run([sys.executable, "-c", HOST_DRIVER, lib], env=_env({}), capture_output=True, text=True,
                       timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "host paths ok" in out, out[-3000:]
    assert "ERROR: AddressSanitizer" not in out, out[-3000:]
