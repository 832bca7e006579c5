"""ctypes mirror of include/drd_gpu.h (structs, constants, field layout).

Array layout (Fortran column-major, x fastest, FP64, reference ghost layers):

* cell field ``f(-1:nx,-1:ny)``  -> numpy shape ``(ny+2, nx+2)``
* u field    ``u(-1:nx+1,-1:ny)`` -> numpy shape ``(ny+2, nx+3)``
* v field    ``v(-1:nx,0:ny)``    -> numpy shape ``(ny+1, nx+2)``

so ``a[k+1, i+1]`` is the Fortran element ``(i, k)`` of a cell field
(``TP/global.f90:54-59``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

DRD_OK = 0
DRD_E_ARG = -1
DRD_E_HIP = -2
DRD_E_STATE = -3
DRD_E_UNSUPPORTED = -4

VARIANT_TP = 0
VARIANT_MF = 1

PREC_REFERENCE = 0
PREC_IDENTITY = 1
PREC_JACOBI = 2
PREC_DCT32 = 3
PREC_DCT64 = 4
PREC_LEARNED = 5
PREC_MG = 6
PREC_MG32 = 7

CH_PREC_REFERENCE = 0
CH_PREC_BLOCK = 1
CH_PREC_RBGS = 2
CH_PREC_RBGS32 = 3

KRYLOV_BICGSTAB = 0
KRYLOV_FGMRES = 1

OP_VELO1 = 0
OP_VELO2 = 1
OP_PHMU = 2
OP_PH = 3
OP_PHORI = 4  # MATRIX_VECTOR_CG_ori TP/PhSolve_ori.f90:348-397

PS_VELO1_JACOBI = 0
PS_IDENTITY = 1
PS_PHMU_JACOBI = 2
PS_PH_DCT = 3
PS_VELO_DCT32 = 4
PS_PHMU_BLOCK = 5
PS_VELO_DCT64 = 6
PS_VELO_ESOLP = 7
PS_VELO_LEARNED = 8
PS_VELO_MG = 9
PS_VELO_MG32 = 11
PS_PHMU_RBGS = 10
PS_PHMU_RBGS32 = 12
PS_VELO2_JACOBI = 17
PS_VELO1_DCT32 = 13
PS_VELO1_DCT64 = 14
PS_VELO1_MG = 15
PS_VELO1_MG32 = 16

LEARN_OFF, LEARN_JUDGE, LEARN_ON, LEARN_FALLBACK = 0, 1, 2, 3

K_VELO_MATVEC, K_PHMU_MATVEC, K_BICG_XR, K_PRESSURE, K_COPY, K_UNET_FWD, K_VELO_PREC, K_VELO_MG = 0, 1, 2, 3, 4, 5, 6, 7
K_PHMU_SWEEP32 = 8

F_U, F_V, F_P, F_PH, F_RMU, F_RHO, F_ETA = 0, 1, 2, 3, 4, 5, 6
F_BDR_PH, F_BDR_PH_USED, F_BDR_PH_PENALTY, F_BDR_PH_GRAD = 7, 8, 9, 10
F_PH_M_INNER, F_PH_M_0, F_ANGLEB, F_ANGLET = 11, 12, 13, 14
F_U_PRE, F_V_PRE, F_PH_PRE, F_RMU_PRE, F_PSI1, F_PSI2 = 15, 16, 17, 18, 19, 20
F_FPD_PH = 21
F_PFORCE1, F_PFORCE2 = 22, 23
F_COUNT = 24

FIELD_NAMES = {
    F_U: "u", F_V: "v", F_P: "p", F_PH: "ph", F_RMU: "rmu", F_RHO: "rho",
    F_ETA: "eta", F_BDR_PH: "bdr_ph", F_BDR_PH_USED: "bdr_ph_used",
    F_BDR_PH_PENALTY: "bdr_ph_penalty", F_BDR_PH_GRAD: "bdr_ph_grad",
    F_PH_M_INNER: "ph_M_inner", F_PH_M_0: "ph_M_0", F_ANGLEB: "angleB",
    F_ANGLET: "angleT", F_U_PRE: "u_pre", F_V_PRE: "v_pre",
    F_PH_PRE: "ph_pre", F_RMU_PRE: "rmu_pre", F_PSI1: "psi1", F_PSI2: "psi2",
    F_FPD_PH: "fpd_ph", F_PFORCE1: "particle_force1", F_PFORCE2: "particle_force2",
}
FIELD_IDS = {v: k for k, v in FIELD_NAMES.items()}


class Params(C.Structure):
    """``drd_params`` (include/drd_gpu.h)."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("variant", C.c_int32),
        ("phi_mode", C.c_int32), ("lr_bd_type", C.c_int32),
        ("velo_prec", C.c_int32), ("maxit_velo1", C.c_int32),
        ("maxit_velo2", C.c_int32), ("maxit_phmu", C.c_int32),
        ("maxit_ph", C.c_int32), ("ch_prec", C.c_int32), ("krylov", C.c_int32),
        ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("re", C.c_double), ("ca", C.c_double), ("xld", C.c_double),
        ("s_implicit", C.c_double), ("I_thickness", C.c_double),
        ("ggx", C.c_double), ("ggz", C.c_double), ("uwT", C.c_double),
        ("uwB", C.c_double), ("xlsB", C.c_double), ("tol", C.c_double),
        ("penalty_k", C.c_double), ("angle", C.c_double),
        ("lambda_rho", C.c_double), ("bdr_ph_used_adjust", C.c_double),
        ("twall", C.c_double), ("UU", C.c_double), ("bdr_upp", C.c_double),
        ("bdr_low", C.c_double),
    ]

    def copy(self) -> "Params":
        q = Params()
        C.pointer(q)[0] = self
        return q

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class SolveInfo(C.Structure):
    _fields_ = [("iter", C.c_int32), ("info", C.c_int32),
                ("resid", C.c_double), ("true_resid", C.c_double)]

    def as_dict(self) -> dict:
        return {"iter": self.iter, "info": self.info, "resid": self.resid,
                "true_resid": self.true_resid}


class LearnInfo(C.Structure):
    """``drd_learn_info`` (include/drd_gpu.h)."""

    _fields_ = [("state", C.c_int32), ("pairs", C.c_int32), ("capacity", C.c_int32),
                ("train_iters", C.c_int32), ("ref_iters", C.c_int32),
                ("judged_iters", C.c_int32), ("last_iters", C.c_int32),
                ("fallbacks", C.c_int32), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("train_ms", C.c_double),
                ("judged_ms", C.c_double), ("plain_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class StepInfo(C.Structure):
    _fields_ = [("phmu", SolveInfo), ("velo", SolveInfo),
                ("div_before", C.c_double), ("div_after_max", C.c_double)]


LR_NEUMANN, LR_DIRICHLET, LR_PERIODIC = ord("n"), ord("d"), ord("p")


def default_params(nx: int, ny: int) -> Params:
    """TP defaults (TP/input, TP/mcl6.f90, TP/finit.f90), grid nx x ny on
    the unit-height domain used by BASELINE-style configs."""
    p = Params()
    p.nx, p.ny = nx, ny
    p.variant = VARIANT_TP
    p.phi_mode = 2
    p.lr_bd_type = ord("n")
    p.velo_prec = PREC_REFERENCE
    p.maxit_velo1, p.maxit_velo2, p.maxit_phmu, p.maxit_ph = 800, 2000, 5000, 500
    p.dx, p.dz = 1.0 / nx, 1.0 / ny
    p.dt = 1.0e-4 / 2.0
    p.re, p.ca, p.xld = 3.0, 12.0, 5.0e-2
    p.s_implicit, p.I_thickness = 1.5, 0.01
    p.ggx = p.ggz = p.uwT = p.uwB = p.xlsB = 0.0
    p.tol = 1.0e-8
    p.penalty_k = 1.0e8
    p.angle = 60.0
    p.lambda_rho = 1.0
    p.bdr_ph_used_adjust = 1.0e-6
    p.twall, p.UU, p.bdr_upp, p.bdr_low = 0.1, 0.0, 1.0, 0.0
    return p


def field_shape(p: Params, field: int) -> tuple:
    nx, ny = p.nx, p.ny
    if field in (F_U, F_U_PRE):
        return (ny + 2, nx + 3)
    if field in (F_V, F_V_PRE):
        return (ny + 1, nx + 2)
    if field in (F_ANGLEB, F_ANGLET):
        return (nx + 1,)
    if 0 <= field < F_COUNT:
        return (ny + 2, nx + 2)
    raise ValueError(f"unknown field id {field}")


def cell_shape(p: Params) -> tuple:
    return (p.ny + 2, p.nx + 2)


def periodic_x(p: Params) -> bool:
    """lr_bd_type 'p': periodic in x (u has nx unknowns per row, u(nx) = u(0))."""
    return p.lr_bd_type == LR_PERIODIC


def nx_u(p: Params) -> int:
    """u unknowns per row of the packed velocity vector."""
    return p.nx + (0 if periodic_x(p) else 1)


def n_uv(p: Params) -> int:
    return nx_u(p) * p.ny + p.nx * (p.ny - 1)


def op_size(p: Params, op: int) -> int:
    if op in (OP_VELO1, OP_VELO2):
        return n_uv(p)
    if op == OP_PHMU:
        return 2 * p.nx * p.ny
    if op in (OP_PH, OP_PHORI):
        return p.nx * p.ny
    raise ValueError(op)


def dptr(a: np.ndarray):
    """ctypes double* of a C-contiguous float64 array (None passes NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))
