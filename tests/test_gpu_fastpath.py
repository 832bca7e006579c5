"""The fast path (additions without a reference counterpart) on the GPU.

* CH 2x2 block Jacobi: bit for bit against the oracle's (itself pinned to
  blocks probed from the oracle's operator), and the block-preconditioned
  ph_mu_solve bit for bit against the oracle's (the dot products of both
  sides use the same canonical summation order);
* DCT/DST velocity preconditioner: against a scipy implementation of
  S^-1 A0^-1 S^-1 with S probed from the oracle's operator; 1e-11 (FP64),
  2e-5 relative (FP32);
* solves and time steps with the fast preconditioners converge to the
  oracle's solution (unique solution; differences are bounded by the
  conditioning times the tolerance).
"""
import numpy as np
import pytest

from drd_application_amd import abi, geometry
from prec_ref import probe_diag, scipy_dct_prec

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _pair(cfg):
    from drd_application_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    g = GpuSolver(cfg.params)
    geometry.load(g, cfg)
    o = Oracle(cfg.params)
    geometry.load(o, cfg)
    return g, o


def test_ch_block_jacobi_bitexact():
    cfg = geometry.tp_case(64, 32, lambda_eta=10.0)
    g, o = _pair(cfg)
    p = cfg.params
    rng = np.random.default_rng(2)
    m0 = rng.uniform(0.0, 1.0, abi.cell_shape(p)) ** 3  # includes near-zero mobility (floor)
    g.set_field(abi.F_PH_M_0, m0)
    o.set_field(abi.F_PH_M_0, m0)
    b = rng.standard_normal(2 * p.nx * p.ny)
    np.testing.assert_array_equal(g.psolve(abi.PS_PHMU_BLOCK, abi.OP_PHMU, b),
                                  o.psolve(abi.PS_PHMU_BLOCK, abi.OP_PHMU, b))
    g.close()


@pytest.mark.parametrize("prec,tol", [(abi.PS_VELO_DCT64, 1e-11), (abi.PS_VELO_DCT32, 2e-5)])
@pytest.mark.parametrize("nx,ny", [(64, 32), (128, 128)])
def test_velocity_dct_preconditioner(prec, tol, nx, ny):
    print(f"dct prec {prec} {nx}x{ny}", flush=True)
    cfg = geometry.two_phase_config(nx, ny, 5, n_drops=3, r_drop=(3, 8), n_discs=2, r_disc=(3, 6),
                                    lambda_eta=50.0, penalty=1e5, dt=5e-5, name="t")
    g, o = _pair(cfg)
    r = np.random.default_rng(4).standard_normal(abi.n_uv(cfg.params))
    got = g.psolve(prec, abi.OP_VELO2, r)
    want = scipy_dct_prec(o, cfg.params, r)
    e = rel_l2(got, want)
    print("rel err", e)
    assert e < tol
    g.close()


@pytest.mark.parametrize("vprec", [abi.PREC_DCT32, abi.PREC_DCT64])
def test_fast_velocity_solve_converges_to_reference_solution(vprec):
    cfg = geometry.two_phase_config(128, 64, 9, n_drops=4, r_drop=(4, 10), n_discs=2, r_disc=(4, 8),
                                    lambda_eta=50.0, penalty=1e5, dt=5e-5, tol=1e-10, name="t")
    p = cfg.params
    g, o = _pair(cfg)
    cs = abi.cell_shape(p)
    args = (np.zeros(cs), np.zeros(cs), np.zeros(abi.field_shape(p, abi.F_U)),
            np.zeros(abi.field_shape(p, abi.F_V)), cfg.fields[abi.F_PH], cfg.fields[abi.F_PH],
            cfg.fields[abi.F_RMU], cfg.fields[abi.F_RMU], p.dt)
    uo, vo, io = o.VeloSolve_2_combined(*args)
    ur, vr, ir = g.VeloSolve_2_combined(*args)
    q = p.copy()
    q.velo_prec = vprec
    from drd_application_amd.solver import GpuSolver
    gf = GpuSolver(q)
    geometry.load(gf, cfg)
    uf, vf, inf = gf.VeloSolve_2_combined(*args)
    print("oracle", io, "gpu ref", ir, "gpu fast", inf)
    assert inf["info"] == 0 and inf["true_resid"] < 1e-9
    assert inf["iter"] < ir["iter"]
    assert rel_l2(uf, uo) < 1e-7 and rel_l2(vf, vo) < 1e-7
    g.close()
    gf.close()


@pytest.mark.parametrize("prec", [abi.PS_IDENTITY, abi.PS_VELO_DCT32])
def test_fgmres_matches_oracle(prec):
    """drd_fgmres (CGS twice-is-enough) against the oracle's FGMRES (MGS) on
    the velocity operator: same solution, close iteration counts."""
    cfg = geometry.two_phase_config(64, 32, 13, n_drops=2, r_drop=(3, 6), n_discs=1, r_disc=(3, 5),
                                    lambda_eta=20.0, penalty=1e5, dt=5e-5, name="t")
    p = cfg.params
    g, o = _pair(cfg)
    n = abi.n_uv(p)
    b = np.random.default_rng(14).standard_normal(n)
    xg, itg, rg, ig = g.fgmres(abi.OP_VELO2, prec, b, np.zeros(n), 3000, 1e-10, restart=20)
    if prec == abi.PS_IDENTITY:
        xo, ito, ro, io = o.fgmres(abi.OP_VELO2, prec, b, np.zeros(n), 3000, 1e-10, restart=20)
        assert abs(itg - ito) <= 2
    else:
        xo, ito, ro, io = o.bicgstab(abi.OP_VELO2, abi.PS_IDENTITY, b, np.zeros(n), 3000, 1e-12)
    print("gpu", itg, rg, ig, "oracle", ito, ro, io)
    assert ig == 0 and rg <= 1e-10
    assert rel_l2(xg, xo) < 1e-7
    g.close()


def test_block_ch_solve_bitexact():
    """ph_mu_solve with the block preconditioner: GPU and oracle take the same
    BiCGSTAB trajectory (85 iterations instead of the reference Jacobi's
    5000-iteration stall on this case)."""
    cfg = geometry.tp_case(64, 64, tol=1e-10)
    p = cfg.params
    p.ch_prec = abi.CH_PREC_BLOCK
    g, o = _pair(cfg)
    pg, mg, ig = g.ph_mu_solve(0.0)
    po, mo, io = o.ph_mu_solve(0.0)
    print("gpu", ig, "oracle", io)
    assert ig["info"] == io["info"] == 0
    assert ig["iter"] == io["iter"] and ig["resid"] == io["resid"]
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(mg, mo)
    g.close()


def test_fast_steps_track_oracle_steps():
    """Five steps of the full fast path (block CH, FP32 fast-Poisson velocity
    preconditioner) against five oracle steps (block CH, reference velocity
    preconditioners), all solves converged to 1e-10."""
    cfg = geometry.tp_case(64, 64, lambda_eta=10.0, tol=1e-10)
    p = cfg.params
    p.ch_prec = abi.CH_PREC_BLOCK
    q = p.copy()
    q.velo_prec = abi.PREC_DCT32
    from drd_application_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    gf = GpuSolver(q)
    geometry.load(gf, cfg)
    o = Oracle(p)
    geometry.load(o, cfg)
    t = 0.0
    for s in range(1, 6):
        bdf = 1 if s == 1 else 2
        i_f = gf.step(bdf, t)
        i_o = o.step(bdf, t)
        print(s, "fast", i_f["phmu"]["iter"], i_f["velo"]["iter"], "oracle", i_o["phmu"]["iter"],
              i_o["velo"]["iter"], flush=True)
        assert i_f["phmu"]["info"] == i_f["velo"]["info"] == 0
        assert i_o["phmu"]["info"] == i_o["velo"]["info"] == 0
        t += p.dt
    for fid in (abi.F_U, abi.F_V, abi.F_P, abi.F_PH):
        e = rel_l2(gf.get_field(fid), o.get_field(fid))
        print(abi.FIELD_NAMES[fid], e)
        assert e < 1e-6
    gf.close()


def test_extrapolated_initial_guess_is_bitexact():
    """drd_set_initial_guess(1) (x0 = 2x^n - x^{n-1}) against the oracle with
    the same option, bit for bit (reference velocity preconditioner, block
    CH); the iterations against the default start are printed (they depend
    on how smoothly the solution varies: startup steps can need more)."""
    from drd_application_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    cfg = geometry.tp_case(64, 64, lambda_eta=10.0, tol=1e-10)
    p = cfg.params
    p.ch_prec = abi.CH_PREC_BLOCK
    g, g0 = GpuSolver(p), GpuSolver(p)
    o = Oracle(p)
    for s in (g, g0, o):
        geometry.load(s, cfg)
    g.set_initial_guess(1)
    o.set_initial_guess(1)
    its, its0 = 0, 0
    t = 0.0
    for k in range(6):
        bdf = 1 if k == 0 else 2
        a, b, c = g.step(bdf, t), o.step(bdf, t), g0.step(bdf, t)
        for key in ("phmu", "velo"):
            assert a[key]["iter"] == b[key]["iter"] and a[key]["resid"] == b[key]["resid"], (k, key)
            assert a[key]["info"] == 0 and c[key]["info"] == 0
        if k > 0:
            its += a["phmu"]["iter"] + a["velo"]["iter"]
            its0 += c["phmu"]["iter"] + c["velo"]["iter"]
        t += p.dt
    for fid in (abi.F_U, abi.F_V, abi.F_P, abi.F_PH, abi.F_RMU):
        np.testing.assert_array_equal(g.get_field(fid), o.get_field(fid))
    print(f"BDF2 iterations with the extrapolated start {its}, from the previous step {its0}")
    g.close()
    g0.close()
