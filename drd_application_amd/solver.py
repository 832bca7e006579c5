"""Python host interface over the C-ABI (libdrd_gpu.so).

`GpuSolver` mirrors the reference's Fortran interface for the solve path:
the module-`global` fields are set with `set_field`, and the subroutines
keep their names and dummy arguments (TP/VeloSolve_combined.f90:5-18,
TP/VeloSolve_2_combined.f90:4-20, TP/PreSolve_2.f90:1-6,
TP/ph_mu_solve.f90:3-7, LIB/BiCGSTAB.f:2-3).  Outputs are returned as
numpy arrays instead of being written into dummy arguments; INFO/ITER/RESID
come back in a dict, as the reference prints them and continues
(TP/VeloSolve_2_combined.f90:417-421).

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdrd_gpu.so")
_LIB = None

dp = abi.dptr


class DrdError(RuntimeError):
    pass


def lib():
    """Loads libdrd_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DrdError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int32)
        V = C.c_void_p
        L.drd_params_default.argtypes = [C.POINTER(abi.Params), C.c_int32, C.c_int32]
        L.drd_last_error.restype = C.c_char_p
        L.drd_field_size.argtypes = [C.POINTER(abi.Params), C.c_int32]
        L.drd_field_size.restype = C.c_int64
        L.drd_init.argtypes = [C.POINTER(abi.Params), C.POINTER(V)]
        L.drd_finalize.argtypes = [V]
        L.drd_synchronize.argtypes = [V]
        L.drd_set_field.argtypes = [V, C.c_int32, P]
        L.drd_get_field.argtypes = [V, C.c_int32, P]
        L.drd_velosolve_combined.argtypes = [V, P, P, P, P, P, C.c_double, P, P,
                                             C.POINTER(abi.SolveInfo)]
        L.drd_velosolve_2_combined.argtypes = [V, P, P, P, P, P, P, P, P, P, P, C.c_double, P, P,
                                               C.POINTER(abi.SolveInfo)]
        L.drd_presolve_2.argtypes = [V, P, C.c_double, P]
        L.drd_ph_mu_solve.argtypes = [V, P, P, C.c_double, C.POINTER(abi.SolveInfo)]
        L.drd_phsolve.argtypes = [V, P, P, C.c_double, C.POINTER(abi.SolveInfo)]
        L.drd_matvec.argtypes = [V, C.c_int32, C.c_double, P, C.c_double, P]
        L.drd_psolve.argtypes = [V, C.c_int32, C.c_int32, P, P]
        L.drd_bicgstab.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, P, P, I, P, I]
        L.drd_fgmres.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, I, P, I]
        L.drd_step.argtypes = [V, C.c_int32, C.c_double, C.POINTER(abi.StepInfo)]
        L.drd_bench_kernel.argtypes = [V, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.drd_kernel_timer.argtypes = [V, C.c_int32, C.c_int32]
        L.drd_kernel_timer_read.argtypes = [V, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().drd_last_error().decode(errors="replace")
        raise DrdError(f"drd error {rc}: {msg}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class GpuSolver:
    """One device context (drd_ctx) for one grid instance."""

    def __init__(self, params: abi.Params):
        self.p = params.copy()
        self._h = C.c_void_p()
        _check(lib().drd_init(C.byref(self.p), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().drd_finalize(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- module `global` fields ----------------------------------------
    def set_field(self, fid: int, a) -> None:
        a = _f64(a)
        if a.shape != abi.field_shape(self.p, fid):
            raise ValueError(f"field {abi.FIELD_NAMES.get(fid, fid)}: shape {a.shape}, "
                             f"expected {abi.field_shape(self.p, fid)}")
        _check(lib().drd_set_field(self._h, fid, dp(a)))

    def get_field(self, fid: int) -> np.ndarray:
        out = np.zeros(abi.field_shape(self.p, fid))
        _check(lib().drd_get_field(self._h, fid, dp(out)))
        return out

    def synchronize(self) -> None:
        _check(lib().drd_synchronize(self._h))

    # ---- MATVEC / PSOLVE / BICGSTAB --------------------------------------
    def matvec(self, op, alpha, x, beta, y):
        y = _f64(y).copy()
        _check(lib().drd_matvec(self._h, op, alpha, dp(_f64(x)), beta, dp(y)))
        return y

    def psolve(self, prec, op, b):
        b = _f64(b)
        x = np.zeros_like(b)
        _check(lib().drd_psolve(self._h, prec, op, dp(x), dp(b)))
        return x

    def bicgstab(self, op, prec, b, x0, maxit, tol):
        b = _f64(b)
        x = _f64(x0).copy()
        it = C.c_int32(maxit)
        res = C.c_double(tol)
        info = C.c_int32(0)
        _check(lib().drd_bicgstab(self._h, op, prec, len(b), dp(b), dp(x), C.byref(it),
                                  C.byref(res), C.byref(info)))
        return x, it.value, res.value, info.value

    def fgmres(self, op, prec, b, x0, maxit, tol, restart=30):
        b = _f64(b)
        x = _f64(x0).copy()
        it = C.c_int32(maxit)
        res = C.c_double(tol)
        info = C.c_int32(0)
        _check(lib().drd_fgmres(self._h, op, prec, len(b), restart, dp(b), dp(x), C.byref(it),
                                C.byref(res), C.byref(info)))
        return x, it.value, res.value, info.value

    # ---- the Fortran subroutines ----------------------------------------
    def VeloSolve_combined(self, psi, rmu_after, ph_after, t, pf1=None, pf2=None):
        ua = np.zeros(abi.field_shape(self.p, abi.F_U))
        va = np.zeros(abi.field_shape(self.p, abi.F_V))
        info = abi.SolveInfo()
        _check(lib().drd_velosolve_combined(
            self._h, dp(_f64(psi)), dp(ua), dp(va), dp(_f64(rmu_after)), dp(_f64(ph_after)), t,
            dp(None if pf1 is None else _f64(pf1)), dp(None if pf2 is None else _f64(pf2)),
            C.byref(info)))
        return ua, va, info.as_dict()

    def VeloSolve_2_combined(self, psi1, psi2, u_pre, v_pre, ph_pre, ph_after, rmu_pre,
                             rmu_after, t, pf1=None, pf2=None):
        ua = np.zeros(abi.field_shape(self.p, abi.F_U))
        va = np.zeros(abi.field_shape(self.p, abi.F_V))
        info = abi.SolveInfo()
        _check(lib().drd_velosolve_2_combined(
            self._h, dp(_f64(psi1)), dp(_f64(psi2)), dp(_f64(u_pre)), dp(ua), dp(_f64(v_pre)),
            dp(va), dp(_f64(ph_pre)), dp(_f64(ph_after)), dp(_f64(rmu_pre)), dp(_f64(rmu_after)),
            t, dp(None if pf1 is None else _f64(pf1)), dp(None if pf2 is None else _f64(pf2)),
            C.byref(info)))
        return ua, va, info.as_dict()

    def PreSolve_2(self, coef):
        psi = np.zeros(abi.cell_shape(self.p))
        md = C.c_double(0)
        _check(lib().drd_presolve_2(self._h, dp(psi), coef, C.byref(md)))
        return psi, md.value

    def ph_mu_solve(self, t=0.0):
        cs = abi.cell_shape(self.p)
        ph_a = np.zeros(cs)
        mu_a = np.zeros(cs)
        info = abi.SolveInfo()
        _check(lib().drd_ph_mu_solve(self._h, dp(ph_a), dp(mu_a), t, C.byref(info)))
        return ph_a, mu_a, info.as_dict()

    def PhSolve(self, t=0.0):
        cs = abi.cell_shape(self.p)
        ph_a = np.zeros(cs)
        mu_a = np.zeros(cs)
        info = abi.SolveInfo()
        _check(lib().drd_phsolve(self._h, dp(mu_a), dp(ph_a), t, C.byref(info)))
        return mu_a, ph_a, info.as_dict()

    # ---- device-resident stepping -----------------------------------------
    def step(self, bdf, t):
        info = abi.StepInfo()
        _check(lib().drd_step(self._h, bdf, t, C.byref(info)))
        return {"phmu": info.phmu.as_dict(), "velo": info.velo.as_dict(),
                "div_before": info.div_before}

    # ---- measurement -------------------------------------------------------
    def bench_kernel(self, which, reps):
        """(mean ms per launch, algorithmic bytes per launch) of kernel
        `which` (abi.K_*) launched `reps` times on the context's stream."""
        ms, by = C.c_double(0), C.c_double(0)
        _check(lib().drd_bench_kernel(self._h, which, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def kernel_timer(self, which, capacity):
        """Brackets up to `capacity` launches of kernel `which` inside the
        following solves with HIP events (0 disables)."""
        _check(lib().drd_kernel_timer(self._h, which, capacity))

    def kernel_timer_read(self):
        """(timed launches, summed ms, algorithmic bytes per launch)."""
        n, ms, by = C.c_int32(0), C.c_double(0), C.c_double(0)
        _check(lib().drd_kernel_timer_read(self._h, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value
