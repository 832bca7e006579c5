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


# MATVEC(ALPHA, X, BETA, Y) and PSOLVE(X, B) of LIB/BiCGSTAB.f, every argument by reference
MATVEC_CB = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                        C.POINTER(C.c_double))
PSOLVE_CB = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double))


def bicgstab_cb(matvec, psolve, b, x0, maxit, tol, ldw=None):
    """BICGSTAB with the caller's own operator (drd_bicgstab_cb, the
    library's own context): matvec(alpha, x, beta, y) returns alpha A x +
    beta y, psolve(b) returns M^-1 b (None: identity), on numpy vectors.
    Returns (x, iter, resid, info, calls) with calls the (matvec, psolve)
    call counts."""
    b = _f64(b)
    n = len(b)
    x = _f64(x0).copy()
    calls = [0, 0]

    def mv(pa, px, pb, py):
        calls[0] += 1
        xa = np.ctypeslib.as_array(px, shape=(n,))
        ya = np.ctypeslib.as_array(py, shape=(n,))
        ya[...] = matvec(pa[0], xa.copy(), pb[0], ya.copy())

    def ps(px, pb):
        calls[1] += 1
        np.ctypeslib.as_array(px, shape=(n,))[...] = psolve(np.ctypeslib.as_array(pb, shape=(n,)).copy())

    fmv = MATVEC_CB(mv)
    fps = PSOLVE_CB(ps) if psolve is not None else C.cast(None, PSOLVE_CB)
    it = C.c_int32(maxit)
    res = C.c_double(tol)
    info = C.c_int32(0)
    _check(lib().drd_bicgstab_cb(None, n, dp(b), dp(x), max(1, n) if ldw is None else ldw, C.byref(it), C.byref(res), fmv,
                                 fps, C.byref(info)))
    return x, it.value, res.value, info.value, tuple(calls)


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
        L.drd_bicgstab_cb.argtypes = [V, C.c_int32, P, P, C.c_int32, I, P, MATVEC_CB, PSOLVE_CB, I]
        L.drd_phsolve_ori.argtypes = [V, C.c_double, P, P, C.POINTER(abi.SolveInfo)]
        L.drd_reinit_solid.argtypes = [V, C.c_int32, C.c_double, C.c_double, C.POINTER(abi.SolveInfo)]
        L.drd_step.argtypes = [V, C.c_int32, C.c_double, C.POINTER(abi.StepInfo)]
        L.drd_bench_kernel.argtypes = [V, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.drd_kernel_timer.argtypes = [V, C.c_int32, C.c_int32]
        L.drd_kernel_timer_read.argtypes = [V, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.drd_kernel_timer_launches.argtypes = [V, C.c_int32, C.POINTER(C.c_float), I]
        F = C.POINTER(C.c_float)
        L.drd_learn_init.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]
        L.drd_learn_init_crops.argtypes = [V, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]
        L.drd_learn_set_every.argtypes = [V, C.c_int32]
        L.drd_learn_collect.argtypes = [V, P, C.c_double, P, I, I]
        L.drd_learn_record.argtypes = [V, C.c_int32]
        L.drd_learn_train.argtypes = [V, C.c_int32, C.c_int32, C.c_double]
        L.drd_learn_enable.argtypes = [V, C.c_int32]
        L.drd_learn_get_info.argtypes = [V, C.POINTER(abi.LearnInfo)]
        L.drd_learn_get_params.argtypes = [V, F, C.c_int32]
        L.drd_learn_set_params.argtypes = [V, F, C.c_int32]
        L.drd_unet_param_count.argtypes = []
        L.drd_batch_create.argtypes = [C.c_int32, C.POINTER(abi.Params), C.POINTER(V)]
        L.drd_batch_destroy.argtypes = [V]
        L.drd_batch_set_field.argtypes = [V, C.c_int32, C.c_int32, P]
        L.drd_batch_get_field.argtypes = [V, C.c_int32, C.c_int32, P]
        L.drd_batch_step.argtypes = [V, C.c_int32, C.c_double, C.POINTER(abi.StepInfo)]
        L.drd_batch_groups.argtypes = [V, I, I, C.c_int32]
        L.drd_use_graphs.argtypes = [V, C.c_int32]
        L.drd_mg_fused.argtypes = [V, C.c_int32]
        L.drd_graph_count.argtypes = [V, I, C.POINTER(C.c_int64)]
        L.drd_write_snapshot.argtypes = [V, C.c_char_p, C.c_int32, C.c_int32, C.c_double]
        L.drd_interpolate_coef.argtypes = [V, C.c_double, P, P, C.c_int32, C.c_double, P, P, C.c_int32,
                                           C.c_double]
        L.drd_particle_force.argtypes = [V, P, P, P, P, P, C.c_int32, C.c_double, C.c_double]
        L.drd_particle_velocity.argtypes = [V, P, P, P, C.c_int32, C.c_double, C.c_double, P, P]
        L.drd_read_snapshot.argtypes = [V, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
        L.drd_set_initial_guess.argtypes = [V, C.c_int32]
        L.drd_batch_set_initial_guess.argtypes = [V, C.c_int32]
        L.drd_tune.argtypes = [V, I, C.POINTER(C.c_double)]
        L.drd_batch_tune.argtypes = [V, I, C.POINTER(C.c_double), C.c_int32]
        L.drd_batch_use_graphs.argtypes = [V, C.c_int32]
        L.drd_unet_run.argtypes = [C.c_int32, C.c_int32, F, C.POINTER(C.c_uint16), F, F, F,
                                   C.POINTER(C.c_double), C.c_float, F]
        _LIB = L
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().drd_last_error().decode(errors="replace")
        raise DrdError(f"drd error {rc}: {msg}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _fptr(a):
    if a is None:
        return None
    if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(C.POINTER(C.c_float))


def unet_param_count() -> int:
    return int(lib().drd_unet_param_count())


def unet_run(params, img_bits, target=None, lr=None):
    """One forward (and, with a target, backward) pass of the learned
    correction's network on the GPU (test hook drd_unet_run).

    img_bits: (H, W, 8) uint16 BF16 bit patterns; target: (H, W, 2) float32.
    Returns (out, grads, loss, params_after) -- grads/loss None without a
    target, params_after None without lr (one Adam step otherwise)."""
    H, W, c = img_bits.shape
    if c != 8:
        raise ValueError("the network takes 8 input channels")
    n = unet_param_count()
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != n:
        raise ValueError(f"expected {n} parameters")
    img = np.ascontiguousarray(img_bits, dtype=np.uint16)
    out = np.zeros((H, W, 2), np.float32)
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
    grads = None if t is None else np.zeros(n, np.float32)
    after = None if (t is None or lr is None) else np.zeros(n, np.float32)
    loss = C.c_double(0.0)
    _check(lib().drd_unet_run(H, W, _fptr(p), img.ctypes.data_as(C.POINTER(C.c_uint16)), _fptr(t),
                              _fptr(out), _fptr(grads), C.byref(loss),
                              0.0 if lr is None else float(lr), _fptr(after)))
    return out, grads, (loss.value if t is not None else None), after


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

    def interpolate_coef(self, lambda_eta, particles, radius, swimmers=None, swm_radius=0.0):
        """interpolate_coef TP/interpolate_coef.f90:4-115 on the device: fpd_ph,
        rho, eta from the particle and swimmer centres ((n, 2) arrays)."""
        pa = np.ascontiguousarray(np.asarray(particles, dtype=np.float64).reshape(-1, 2))
        sa = np.ascontiguousarray(np.asarray(swimmers if swimmers is not None else np.zeros((0, 2)),
                                             dtype=np.float64).reshape(-1, 2))
        px, py = np.ascontiguousarray(pa[:, 0]), np.ascontiguousarray(pa[:, 1])
        sx, sy = np.ascontiguousarray(sa[:, 0]), np.ascontiguousarray(sa[:, 1])
        P = C.POINTER(C.c_double)
        cast = lambda a: a.ctypes.data_as(P)
        _check(lib().drd_interpolate_coef(self._h, lambda_eta, cast(px), cast(py), len(px), radius, cast(sx),
                                          cast(sy), len(sx), swm_radius))

    def particle_force(self, bodies, lx, ly):
        """force_caculation_vel TP/particles_move.f90:758-839 on the device:
        particle_force1/2 (F_PFORCE1/2, used by step) from bodies, an (n, 5)
        array of (x, y, radius, fx, fy) in the reference's loop order."""
        b = np.asarray(bodies, dtype=np.float64).reshape(-1, 5)
        col = [np.ascontiguousarray(b[:, j]) for j in range(5)]
        P = C.POINTER(C.c_double)
        _check(lib().drd_particle_force(self._h, *[a.ctypes.data_as(P) for a in col], len(b), lx, ly))

    def particle_velocity(self, bodies, lx, ly):
        """update_position_vel TP/particles_move.f90:842-920 on the device: the
        weighted velocity averages (u_aver, v_aver) of bodies (n, >= 3)."""
        b = np.asarray(bodies, dtype=np.float64).reshape(len(bodies), -1)
        col = [np.ascontiguousarray(b[:, j]) for j in range(3)]
        ua, va = np.zeros(len(b)), np.zeros(len(b))
        P = C.POINTER(C.c_double)
        _check(lib().drd_particle_velocity(self._h, *[a.ctypes.data_as(P) for a in col], len(b), lx, ly,
                                           ua.ctypes.data_as(P), va.ctypes.data_as(P)))
        return ua, va

    def write_snapshot(self, path, f_field=abi.F_BDR_PH, f2_field=abi.F_PH, t=0.0):
        """The reference's data file (TP/inout.f90 output): x, y, u, v, p, f, f2, t."""
        _check(lib().drd_write_snapshot(self._h, os.fsencode(path), f_field, f2_field, t))

    def read_snapshot(self, path, f_field=abi.F_BDR_PH, f2_field=abi.F_PH):
        """Sets u, v, p, f, f2 (0:nx, 0:ny) from a data file; returns its t."""
        t = C.c_double(0.0)
        _check(lib().drd_read_snapshot(self._h, os.fsencode(path), f_field, f2_field, C.byref(t)))
        return t.value

    def PhSolve_ori(self, alpha_s=0.0):
        """PhSolve_ori TP/PhSolve_ori.f90:1-234 -> (rmu_after, ph_after, info)."""
        cs = abi.cell_shape(self.p)
        ph_a, mu_a = np.zeros(cs), np.zeros(cs)
        info = abi.SolveInfo()
        _check(lib().drd_phsolve_ori(self._h, alpha_s, dp(mu_a), dp(ph_a), C.byref(info)))
        return mu_a, ph_a, info.as_dict()

    def reinit_solid(self, steps=51, dt=1e-4, alpha_s=0.0):
        """The CH re-initialisation of the solid domain MF/finit.f90:132-188."""
        info = abi.SolveInfo()
        _check(lib().drd_reinit_solid(self._h, steps, dt, alpha_s, C.byref(info)))
        return info.as_dict()

    def set_initial_guess(self, mode):
        """0: BDF2 solves start from the previous step (the reference); 1: from 2x^n - x^{n-1}."""
        _check(lib().drd_set_initial_guess(self._h, mode))

    def mg_fused(self, on=True):
        """V(2,2) multigrid cycles by one LDS-tiled launch per level and leg
        (on, the default) or one launch per sweep and transfer (off); the
        same values either way."""
        _check(lib().drd_mg_fused(self._h, 1 if on else 0))

    def use_graphs(self, on=True):
        """Capture BiCGSTAB iteration chunks as hipGraphs and replay them."""
        _check(lib().drd_use_graphs(self._h, 1 if on else 0))

    def graph_count(self):
        """(captured chunks held, replays refused for a changed fingerprint)."""
        live, rec = C.c_int32(), C.c_int64()
        _check(lib().drd_graph_count(self._h, C.byref(live), C.byref(rec)))
        return live.value, rec.value

    def tune(self):
        """(chosen transpose tile shape, [ms per preconditioner application per shape])."""
        sh = C.c_int32(0)
        ms = (C.c_double * 3)()
        _check(lib().drd_tune(self._h, C.byref(sh), ms))
        return sh.value, list(ms)

    # ---- learned correction of the velocity preconditioner ------------------
    def learn_init(self, capacity, crop_h, crop_w, seed=1):
        _check(lib().drd_learn_init(self._h, capacity, crop_h, crop_w, seed))

    def learn_init_crops(self, capacity, crop_h, crop_w, crops_per_iterate, seed=1):
        _check(lib().drd_learn_init_crops(self._h, capacity, crop_h, crop_w, crops_per_iterate, seed))

    def learn_set_every(self, every):
        _check(lib().drd_learn_set_every(self._h, every))

    def learn_collect(self, b, tol=1e-12):
        """Recorded BiCGSTAB(FP32 DCT) solve of A x = b (BDF2 velocity
        operator) from 0; returns (x, pairs added, iterations)."""
        b = _f64(b)
        x = np.zeros_like(b)
        added, its = C.c_int32(0), C.c_int32(0)
        _check(lib().drd_learn_collect(self._h, dp(b), tol, dp(x), C.byref(added), C.byref(its)))
        return x, added.value, its.value

    def learn_record(self, on):
        _check(lib().drd_learn_record(self._h, 1 if on else 0))

    def learn_train(self, iters, batch, lr):
        _check(lib().drd_learn_train(self._h, iters, batch, lr))
        return self.learn_info()

    def learn_enable(self, on):
        _check(lib().drd_learn_enable(self._h, 1 if on else 0))

    def learn_info(self):
        info = abi.LearnInfo()
        _check(lib().drd_learn_get_info(self._h, C.byref(info)))
        return info.as_dict()

    def learn_params(self):
        p = np.zeros(unet_param_count(), np.float32)
        _check(lib().drd_learn_get_params(self._h, _fptr(p), p.size))
        return p

    def learn_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().drd_learn_set_params(self._h, _fptr(p), p.size))

    def kernel_timer_read(self):
        """(timed launches, summed ms, algorithmic bytes per launch)."""
        n, ms, by = C.c_int32(0), C.c_double(0), C.c_double(0)
        _check(lib().drd_kernel_timer_read(self._h, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value

    def kernel_timer_launches(self):
        """Durations (ms, float32 array) of the timed launches in launch order."""
        n = C.c_int32(0)
        _check(lib().drd_kernel_timer_launches(self._h, 0, None, C.byref(n)))
        out = np.zeros(n.value, dtype=np.float32)
        _check(lib().drd_kernel_timer_launches(self._h, n.value, out.ctypes.data_as(C.POINTER(C.c_float)),
                                               C.byref(n)))
        return out


class GpuBatch:
    """A batch of independent instances (drd_batch): equal-size instances
    step together through batched launches."""

    def __init__(self, params_list):
        self.params = [q.copy() for q in params_list]
        arr = (abi.Params * len(self.params))(*self.params)
        self._h = C.c_void_p()
        _check(lib().drd_batch_create(len(self.params), arr, C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().drd_batch_destroy(self._h)
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

    def __len__(self):
        return len(self.params)

    def set_field(self, inst: int, fid: int, a) -> None:
        a = _f64(a)
        if a.shape != abi.field_shape(self.params[inst], fid):
            raise ValueError(f"field {abi.FIELD_NAMES.get(fid, fid)}: shape {a.shape}")
        _check(lib().drd_batch_set_field(self._h, inst, fid, dp(a)))

    def get_field(self, inst: int, fid: int) -> np.ndarray:
        out = np.zeros(abi.field_shape(self.params[inst], fid))
        _check(lib().drd_batch_get_field(self._h, inst, fid, dp(out)))
        return out

    def step(self, bdf, t):
        """One step of every instance; a list of per-instance reports."""
        infos = (abi.StepInfo * len(self.params))()
        _check(lib().drd_batch_step(self._h, bdf, t, infos))
        return [{"phmu": i.phmu.as_dict(), "velo": i.velo.as_dict(), "div_before": i.div_before} for i in infos]

    def use_graphs(self, on=True):
        _check(lib().drd_batch_use_graphs(self._h, 1 if on else 0))

    def set_initial_guess(self, mode):
        _check(lib().drd_batch_set_initial_guess(self._h, mode))

    def tune(self):
        """[(nx, ny, chosen tile shape, [ms per shape])] for every size group."""
        grp = self.groups()
        sh = (C.c_int32 * len(grp))()
        ms = (C.c_double * (3 * len(grp)))()
        _check(lib().drd_batch_tune(self._h, sh, ms, len(grp)))
        return [(g[0], g[1], sh[i], list(ms[3 * i:3 * i + 3])) for i, g in enumerate(grp)]

    def groups(self):
        """[(nx, ny, instances)] of the batch's size groups."""
        n = C.c_int32(0)
        _check(lib().drd_batch_groups(self._h, C.byref(n), None, 0))
        desc = (C.c_int32 * (3 * n.value))()
        _check(lib().drd_batch_groups(self._h, C.byref(n), desc, n.value))
        return [(desc[3 * g], desc[3 * g + 1], desc[3 * g + 2]) for g in range(n.value)]
