"""drd_particle_velocity_noalloc (no context, no device allocation) against
drd_particle_velocity on the same fields and bodies: bit for bit, on two
calls in a row (the second on a smaller grid with fewer bodies, so it reuses
the per-device work array the first call left behind), and its size limit.

The test is the caller, so it allocates the device arrays itself through the
HIP runtime the library is linked against."""
import ctypes as C

import numpy as np
import pytest

from drd_application_amd import abi, geometry

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs (the library's own runtime)."""

    def __init__(self):
        from drd_application_amd import solver
        solver.lib()  # loads libamdhip64 as the library's dependency
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.live = []

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)))
        self.live.append(p.value)
        self.ok(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))  # host to device
        return p.value

    def get(self, p, n):
        a = np.empty(n)
        self.ok(self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(p), a.nbytes, 2))  # device to host
        return a

    def close(self):
        for p in self.live:
            self.h.hipFree(C.c_void_p(p))
        self.live = []


def _lib():
    from drd_application_amd import solver
    L = solver.lib()
    P = C.c_void_p
    L.drd_particle_velocity_noalloc.argtypes = [C.POINTER(abi.Params), P, P, P, P, P, C.c_int32, C.c_double,
                                                 C.c_double, P, P, P]
    L.drd_noalloc_scratch_bytes.restype = C.c_int64
    return L


def _bodies(rng, n, lx, ly):
    b = np.stack([rng.uniform(0, lx, n), rng.uniform(0, ly, n), rng.choice([0.03, 0.05], n)], axis=1)
    # the periodic corners and edges
    b[:4, :2] = [[0.001, 0.001], [lx - 0.001, ly - 0.001], [0.0, ly / 2], [lx / 2, ly - 1e-9]]
    return b


def _case(L, hip, stream, nx, ny, nbody, seed):
    from drd_application_amd.solver import GpuSolver
    print(f"noalloc particle velocity {nx}x{ny} bodies {nbody}", flush=True)
    cfg = geometry.tp_case(nx, ny, lambda_eta=10.0, tol=1e-8)
    p = cfg.params
    lx, ly = nx * p.dx, ny * p.dz
    rng = np.random.default_rng(seed)
    g = GpuSolver(p)
    geometry.load(g, cfg)
    u = rng.normal(size=g.get_field(abi.F_U).shape)
    v = rng.normal(size=g.get_field(abi.F_V).shape)
    g.set_field(abi.F_U, u)
    g.set_field(abi.F_V, v)
    b = _bodies(rng, nbody, lx, ly)
    ref_u, ref_v = g.particle_velocity(np.concatenate([b, np.zeros((nbody, 2))], axis=1), lx, ly)
    g.close()
    du, dv = hip.put(u), hip.put(v)
    dx_, dy_, dr_ = hip.put(b[:, 0]), hip.put(b[:, 1]), hip.put(b[:, 2])
    ou, ov = hip.put(np.full(nbody, np.nan)), hip.put(np.full(nbody, np.nan))
    rc = L.drd_particle_velocity_noalloc(C.byref(p), du, dv, dx_, dy_, dr_, nbody, lx, ly, ou, ov, stream)
    assert rc == 0, L.drd_last_error()
    hip.ok(hip.h.hipStreamSynchronize(C.c_void_p(stream)))
    got_u, got_v = hip.get(ou, nbody), hip.get(ov, nbody)
    np.testing.assert_array_equal(got_u, ref_u)
    np.testing.assert_array_equal(got_v, ref_v)


def test_noalloc_particle_velocity_matches_the_context_form():
    L = _lib()
    hip = Hip()
    s = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(s)))
    try:
        _case(L, hip, s.value, 257, 130, 600, 3)  # several batches of 256 bodies per tile
        _case(L, hip, s.value, 64, 32, 7, 4)      # smaller: the work array keeps the first call's data
    finally:
        hip.ok(hip.h.hipStreamSynchronize(s))
        hip.h.hipStreamDestroy(s)
        hip.close()


def test_noalloc_rejects_beyond_its_work_array():
    L = _lib()
    p = abi.default_params(4096, 2048)
    cap = L.drd_noalloc_scratch_bytes() // 8
    nbody = 4096  # 8 n + 3 * 256 n doubles > 2 Mi doubles
    assert 8 * nbody + 3 * 256 * nbody > cap
    one = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = L.drd_particle_velocity_noalloc(C.byref(p), one, one, one, one, one, nbody, 2.0, 1.0, one, one, None)
    assert rc == abi.DRD_E_ARG
    assert b"work array" in L.drd_last_error()
