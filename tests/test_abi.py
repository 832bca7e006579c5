"""The C-ABI library: exports, struct layouts and loud failure without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from drd_application_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "drd_gpu.h")
LIB = os.path.join(ROOT, "drd_application_amd", "libdrd_gpu.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drd_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libdrd_gpu.so not built")
    return C.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_c_compiler(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "drd_gpu.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(drd_params),"
        " offsetof(drd_params, dx), offsetof(drd_params, bdr_low), sizeof(drd_solve_info),"
        " sizeof(drd_step_info), offsetof(drd_step_info, div_before));"
        " printf(\"%zu %zu\\n\", sizeof(drd_learn_info), offsetof(drd_learn_info, loss_first)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()))
    want = [C.sizeof(abi.Params), abi.Params.dx.offset, abi.Params.bdr_low.offset,
            C.sizeof(abi.SolveInfo), C.sizeof(abi.StepInfo), abi.StepInfo.div_before.offset,
            C.sizeof(abi.LearnInfo), abi.LearnInfo.loss_first.offset]
    assert got == want


def test_params_default_matches_python(lib):
    p = abi.Params()
    lib.drd_params_default.argtypes = [C.POINTER(abi.Params), C.c_int32, C.c_int32]
    lib.drd_params_default(C.byref(p), 64, 32)
    q = abi.default_params(64, 32)
    assert p.as_dict() == q.as_dict()


def test_field_sizes(lib):
    lib.drd_field_size.argtypes = [C.POINTER(abi.Params), C.c_int32]
    lib.drd_field_size.restype = C.c_int64
    p = abi.default_params(10, 6)
    for fid in range(abi.F_COUNT):
        n = 1
        for s in abi.field_shape(p, fid):
            n *= s
        assert lib.drd_field_size(C.byref(p), fid) == n
    assert lib.drd_field_size(C.byref(p), 99) == 0


def test_init_rejects_bad_arguments_and_fails_loudly_without_gpu(lib):
    lib.drd_init.argtypes = [C.POINTER(abi.Params), C.POINTER(C.c_void_p)]
    lib.drd_last_error.restype = C.c_char_p
    h = C.c_void_p()
    p = abi.default_params(2, 2)
    assert lib.drd_init(C.byref(p), C.byref(h)) == abi.DRD_E_ARG
    p = abi.default_params(16, 16)
    p.lr_bd_type = ord("x")
    assert lib.drd_init(C.byref(p), C.byref(h)) == abi.DRD_E_ARG
    p = abi.default_params(16, 16)
    rc = lib.drd_init(C.byref(p), C.byref(h))
    if rc == 0:  # a GPU is present: nothing more to check here
        lib.drd_finalize(h)
        return
    assert rc == abi.DRD_E_HIP
    assert b"device" in lib.drd_last_error()


def test_solver_raises_without_library(monkeypatch, tmp_path):
    from drd_application_amd import solver
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(solver, "_LIB", None)
    with pytest.raises(solver.DrdError):
        solver.lib()
