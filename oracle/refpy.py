"""ctypes harness over the reference Fortran built from its own sources.

TEST INFRASTRUCTURE ONLY.  `make -C oracle ref` compiles the reference's
routines in place (oracle/Makefile) into oracle/_ref/libdrdref_{tp,mf}.so.
Their grid is the compile-time one of each module `global`
(TP/global.f90:15-16: 128 x 128; MF/global.f90:18-19: 420 x 180) and
penalty_k is a compile-time parameter (1e8 TP, 1e5 MF).

Module variables are reached through flang's exported symbols
(`_QMglobalE<name>`), routines through their `name_` entry points; every
Fortran dummy argument is passed by reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")

_P = C.POINTER(C.c_double)


def available(which: str = "tp") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libdrdref_{which}.so"))


class Ref:
    """One loaded reference library (TP or MF) and its module `global`."""

    def __init__(self, which: str = "tp"):
        path = os.path.join(REF_DIR, f"libdrdref_{which}.so")
        # RTLD_LAZY: the unresolved BICGSTAB is never called
        self.lib = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        self.nr = C.c_int32.in_dll(self.lib, "_QMglobalECnr").value
        self.nz = C.c_int32.in_dll(self.lib, "_QMglobalECnz").value
        self.penalty_k = C.c_double.in_dll(self.lib, "_QMglobalECpenalty_k").value
        self.pi = C.c_double.in_dll(self.lib, "_QMglobalECpi").value
        nr, nz = self.nr, self.nz
        self.cell_shape = (nz + 2, nr + 2)
        self.u_shape = (nz + 2, nr + 3)
        self.v_shape = (nz + 1, nr + 2)
        self.nuv = (nr + 1) * nz + nr * (nz - 1)

    # ---- module variables -------------------------------------------------
    def scalar(self, name: str) -> C.c_double:
        return C.c_double.in_dll(self.lib, f"_QMglobalE{name.lower()}")

    def set_scalar(self, name: str, value: float) -> None:
        self.scalar(name).value = float(value)

    def set_int(self, name: str, value: int) -> None:
        C.c_int32.in_dll(self.lib, f"_QMglobalE{name.lower()}").value = int(value)

    def set_char(self, name: str, ch: str) -> None:
        C.c_char.in_dll(self.lib, f"_QMglobalE{name.lower()}").value = ch.encode()

    def array(self, name: str, shape) -> np.ndarray:
        n = int(np.prod(shape))
        buf = (C.c_double * n).in_dll(self.lib, f"_QMglobalE{name.lower()}")
        return np.ctypeslib.as_array(buf).reshape(shape)

    def set_params(self, p) -> None:
        """Copy the scalar fields of an abi.Params into module `global`."""
        assert p.nx == self.nr and p.ny == self.nz, "grid is fixed by the reference build"
        self.set_scalar("dt", p.dt)
        self.set_scalar("dr", p.dx)
        self.set_scalar("dz", p.dz)
        self.set_scalar("dr2", p.dx ** 2)
        self.set_scalar("dz2", p.dz ** 2)
        # the MF module lacks some TP names (e.g. bdr_ph_used_adjust, phi_mode)
        for name in ("re", "ca", "xld", "s_implicit", "I_thickness", "ggx", "ggz", "uwT",
                     "uwB", "tol", "angle", "bdr_ph_used_adjust", "twall", "UU", "bdr_upp",
                     "bdr_low", "xlsB", "lambda_rho"):
            if self.has(name):
                self.set_scalar(name, getattr(p, name))
        if self.has("phi_mode"):
            self.set_int("phi_mode", p.phi_mode)
        if self.has("lr_bd_type"):
            self.set_char("lr_bd_type", chr(p.lr_bd_type))

    def has(self, name: str) -> bool:
        try:
            C.c_char.in_dll(self.lib, f"_QMglobalE{name.lower()}")
            return True
        except ValueError:
            return False

    # ---- routines ------------------------------------------------------------
    def _fn(self, name):
        return getattr(self.lib, name)

    def matvec(self, name: str, alpha: float, x: np.ndarray, beta: float, y: np.ndarray):
        """y := alpha*A*x + beta*y through MATVEC(alpha,xx,beta,yy)."""
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        y = np.ascontiguousarray(y, dtype=np.float64).copy()
        a = C.c_double(alpha)
        b = C.c_double(beta)
        self._fn(name)(C.byref(a), x.ctypes.data_as(_P), C.byref(b), y.ctypes.data_as(_P))
        return y

    def psolve(self, name: str, b: np.ndarray):
        """x from PSOLVE(x_out, b_in)."""
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        x = np.zeros_like(b)
        self._fn(name)(x.ctypes.data_as(_P), b.ctypes.data_as(_P))
        return x

    def setup(self) -> None:
        self._fn("setup_")()

    def op_double_gradient_neum(self, rmu, mo):
        nr, nz = C.c_int32(self.nr), C.c_int32(self.nz)
        dx2 = C.c_double(self.scalar("dr2").value)
        dy2 = C.c_double(self.scalar("dz2").value)
        rmu = np.ascontiguousarray(rmu, dtype=np.float64)
        mo = np.ascontiguousarray(mo, dtype=np.float64)
        out = np.zeros(self.cell_shape)
        self._fn("op_double_gradient_neum_")(C.byref(nr), C.byref(nz), C.byref(dx2), C.byref(dy2),
                                             rmu.ctypes.data_as(_P), mo.ctypes.data_as(_P),
                                             out.ctypes.data_as(_P))
        return out

    def ch_bulk_energy(self, phi: float, psi: float, mode: int | None = None) -> float:
        f = self._fn("ch_bulk_energy_")
        f.restype = C.c_double
        a, b = C.c_double(phi), C.c_double(psi)
        if mode is None:
            return f(C.byref(a), C.byref(b))
        m = C.c_int32(mode)
        return f(C.byref(a), C.byref(b), C.byref(m))

    def ch_bdr_condition(self, angle: float, phi: float, mode: int) -> float:
        f = self._fn("ch_bdr_condition_")
        f.restype = C.c_double
        a, b, pi, m = C.c_double(angle), C.c_double(phi), C.c_double(self.pi), C.c_int32(mode)
        return f(C.byref(a), C.byref(b), C.byref(pi), C.byref(m))

    def vcost(self, x: np.ndarray) -> np.ndarray:
        """VCOST over the rows of x (m sequences of length n)."""
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        m, n = x.shape
        # Fortran X(MDIMX,N): sequence j is X(j,:) -> store column-major
        xf = np.asfortranarray(x.copy())
        xt = np.zeros((m, n), order="F")
        ws = np.zeros(3 * n + 115)
        nn = C.c_int32(n)
        self._fn("vcosti_")(C.byref(nn), ws.ctypes.data_as(_P))
        mm, md = C.c_int32(m), C.c_int32(m)
        self._fn("vcost_")(C.byref(mm), C.byref(nn), xf.ctypes.data_as(_P),
                           xt.ctypes.data_as(_P), C.byref(md), ws.ctypes.data_as(_P))
        return np.ascontiguousarray(xf)

    def dct_solve(self, nr1: int, nz1: int, f0: np.ndarray, cof: np.ndarray) -> np.ndarray:
        """DCT(nr1,nz1,fO,fN,cof,wsavex,wsavey): fO,fN (-1:nr1+1,-1:nz1+1)."""
        fo = np.ascontiguousarray(f0, dtype=np.float64).copy()
        fn = np.zeros_like(fo)
        cof = np.ascontiguousarray(cof, dtype=np.float64)
        wx = np.zeros(3 * (nr1 + 1) + 100)
        wy = np.zeros(3 * (nz1 + 1) + 100)
        nx1, ny1 = C.c_int32(nr1 + 1), C.c_int32(nz1 + 1)
        self._fn("vcosti_")(C.byref(nx1), wx.ctypes.data_as(_P))
        self._fn("vcosti_")(C.byref(ny1), wy.ctypes.data_as(_P))
        a, b = C.c_int32(nr1), C.c_int32(nz1)
        self._fn("dct_")(C.byref(a), C.byref(b), fo.ctypes.data_as(_P), fn.ctypes.data_as(_P),
                         cof.ctypes.data_as(_P), wx.ctypes.data_as(_P), wy.ctypes.data_as(_P))
        return fn

    def esolp_x(self, nr1: int, nz1: int, pa, pb, pc, b) -> np.ndarray:
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        p = np.zeros_like(b)
        wx = np.zeros(3 * (nr1 + 1) + 100)
        nx1 = C.c_int32(nr1 + 1)
        self._fn("vcosti_")(C.byref(nx1), wx.ctypes.data_as(_P))
        arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in (pa, pb, pc)]
        a, c = C.c_int32(nr1), C.c_int32(nz1)
        self._fn("esolp_x_")(C.byref(a), C.byref(c), *[t.ctypes.data_as(_P) for t in arrs],
                             b.ctypes.data_as(_P), p.ctypes.data_as(_P), wx.ctypes.data_as(_P))
        return p

    def rfords(self, ff, aa, bb, cc) -> np.ndarray:
        ff = np.ascontiguousarray(ff, dtype=np.float64).copy()
        arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in (aa, bb, cc)]
        n = C.c_int32(len(ff))
        self._fn("rfords_")(C.byref(n), ff.ctypes.data_as(_P), *[t.ctypes.data_as(_P) for t in arrs])
        return ff
