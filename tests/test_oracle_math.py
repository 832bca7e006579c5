"""Pins for the oracle parts the reference build cannot run.

The reference's BICGSTAB needs MKL BLAS-1 and its pressure solve FFTW, which
this image lacks (both are prebuilt binaries in the checkout, never loaded),
and the right-hand-side assembly lives inside the subroutines that call
BICGSTAB.  These are pinned by what the mathematics fixes:

* REDFT10 / REDFT01 against scipy's DCT-II / DCT-III (the same definitions,
  FFTW manual fftw3.info-1:2755-2800) and their inverse pair identity;
* the pressure solve against the reference's own commented manufactured
  solution (TP/PreSolve_2.f90:32-64): a cosine mode is reproduced exactly;
* BiCGSTAB against dense LU on small nonsymmetric systems, and each
  documented exit path of LIB/BiCGSTABREVCOM.f;
* operator invariants: symmetry of the BDF2 velocity operator, conservation
  of op_double_gradient_Neum, constant-coefficient reduction to the textbook
  5-point Neumann Laplacian;
* right-hand sides: each term isolated and compared with its closed form.
"""
import math

import numpy as np
import pytest
import scipy.fft

from drd_application_amd import abi, geometry
from oracle import oracle as oc


# ---- transforms ---------------------------------------------------------------
@pytest.mark.parametrize("nx,ny", [(8, 4), (16, 12), (7, 5), (32, 32)])
def test_dct_definitions_match_scipy(nx, ny):
    rng = np.random.default_rng(nx * 100 + ny)
    a = rng.standard_normal((ny, nx))
    np.testing.assert_allclose(oc.dct2_2d(a), scipy.fft.dctn(a, type=2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oc.dct3_2d(a), scipy.fft.dctn(a, type=3), rtol=1e-12, atol=1e-12)
    # REDFT01(REDFT10(x)) = 4 nx ny x
    np.testing.assert_allclose(oc.dct3_2d(oc.dct2_2d(a)), 4 * nx * ny * a, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("nx,ny", [(64, 32), (16, 128), (24, 20), (1, 8), (4096, 2)])
def test_fft_factored_dct_matches_definition(nx, ny):
    """The pressure solve's FFT-factored REDFT10/REDFT01 (Makhoul + radix-2
    Stockham for powers of two, table sums otherwise) against the O(n^2)
    definitions and scipy."""
    a = np.random.default_rng(nx + ny).standard_normal((ny, nx))
    f = oc.dct_2d_fast(a, 0)
    d = oc.dct2_2d(a)
    assert np.abs(f - d).max() <= 1e-13 * np.abs(d).max()
    ref = scipy.fft.dctn(a, type=2)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max()
    g = oc.dct_2d_fast(a, 1)
    d3 = oc.dct3_2d(a)
    assert np.abs(g - d3).max() <= 1e-13 * np.abs(d3).max()
    # REDFT01 inverts REDFT10 up to 4 nx ny
    back = oc.dct_2d_fast(f, 1) / (4.0 * nx * ny)
    assert np.abs(back - a).max() <= 1e-13 * np.abs(a).max()

@pytest.mark.parametrize("n", [2, 5, 16, 33])
def test_vcost_self_inverse_and_definition(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n))
    np.testing.assert_allclose(oc.vcost(oc.vcost(x)), x, rtol=1e-12, atol=1e-12)
    # definition LIB/vfftpack.f:689-695 evaluated independently
    k = np.arange(n)
    C = 2 * np.cos(np.outer(k, k) * np.pi / (n - 1))
    C[:, 0] = 1.0
    C[:, -1] = (-1.0) ** k
    ref = (x @ C.T) / math.sqrt(2 * (n - 1))
    np.testing.assert_allclose(oc.vcost(x), ref, rtol=1e-12, atol=1e-12)


def test_pressure_manufactured_solution():
    """TP/PreSolve_2.f90:32-64: p = cos(20 pi x/L) cos(10 pi z/H) sampled at
    cell centres; with div u = Lap_h p the DCT solve returns p exactly (the
    mode is an eigenvector of the discrete Neumann Laplacian)."""
    nx, ny = 64, 32
    p = abi.default_params(nx, ny)
    p.dx, p.dz = 2.0 / nx, 1.0 / ny
    o = oc.Oracle(p)
    L, H = 2.0, 1.0
    X, Z = geometry._cell_centres(p)
    pp = np.cos(20 * np.pi * X / L) * np.cos(10 * np.pi * Z / H)
    u = np.zeros(abi.field_shape(p, abi.F_U))
    v = np.zeros(abi.field_shape(p, abi.F_V))
    # u(i,k) = (p(i,k) - p(i-1,k))/dx  for i=0..nx (ghost-reflected p), v likewise
    u[1:-1, 1:-1] = (pp[1:-1, 1:] - pp[1:-1, :-1]) / p.dx
    u[1:-1, 1] = 0.0
    u[1:-1, nx + 1] = 0.0
    v[1:-1, 1:-1] = (pp[2:-1, 1:-1] - pp[1:-2, 1:-1]) / p.dz
    v[0, :] = v[-1, :] = 0.0
    o.set_field(abi.F_U, u)
    o.set_field(abi.F_V, v)
    psi, mean = o.PreSolve_2(1.0)
    err = np.max(np.abs(psi[1:-1, 1:-1] - pp[1:-1, 1:-1]))
    assert err < 1e-11, err
    assert abs(mean) < 1e-12
    # Neumann ghosts psi(-1,k) = psi(0,k) ... (TP/dct_fftw.f90:43-51)
    np.testing.assert_array_equal(psi[:, 0], psi[:, 1])
    np.testing.assert_array_equal(psi[0, :], psi[1, :])


def test_pressure_mean_mode_passes_with_sign_flip():
    """Pre_Lcc(0,0) = -1 (TP/setup.f90:131): a constant divergence d comes back
    as psi = -d (not zeroed)."""
    nx, ny = 16, 8
    p = abi.default_params(nx, ny)
    o = oc.Oracle(p)
    u = np.zeros(abi.field_shape(p, abi.F_U))
    # div = (u(i+1)-u(i))/dx = d for a linear u
    d = 0.37
    i = np.arange(-1, nx + 2)
    u[:, :] = (d * p.dx) * i[None, :]
    o.set_field(abi.F_U, u)
    psi, mean = o.PreSolve_2(1.0)
    assert abs(mean - d) < 1e-12
    np.testing.assert_allclose(psi, -d, rtol=1e-12)


def _pairwise_spec(t):
    """The canonical order written out: zero-pad to a power of two, then add
    neighbours level by level."""
    t = [float(v) for v in t]
    if not t:
        return 0.0
    n = 1
    while n < len(t):
        n *= 2
    t += [0.0] * (n - len(t))
    while len(t) > 1:
        t = [t[2 * i] + t[2 * i + 1] for i in range(len(t) // 2)]
    return t[0]


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 1000, 1024, 4097])
def test_canonical_sum_definition(n):
    """oc_pairwise (the order of every dot product and norm on the path)
    against the level-by-level definition, bit for bit, for sums and dots."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    y = rng.standard_normal(n)
    assert oc.pairwise(x) == _pairwise_spec(x)
    assert oc.pairwise(x, y) == _pairwise_spec(x * y)


def test_canonical_sum_exact_and_not_sequential():
    rng = np.random.default_rng(5)
    ints = rng.integers(-2 ** 20, 2 ** 20, 3001).astype(np.float64)
    assert oc.pairwise(ints) == float(ints.astype(np.int64).sum())
    # (2^53 + 1) + (1 - 2^53) = 1 pairwise; left to right the ones round away
    t = np.array([2.0 ** 53, 1.0, 1.0, -(2.0 ** 53)])
    assert oc.pairwise(t) == 1.0
    assert ((t[0] + t[1]) + t[2]) + t[3] == 0.0
    assert oc.pairwise(np.zeros(0)) == 0.0

# ---- BiCGSTAB -------------------------------------------------------------------
@pytest.mark.parametrize("n,seed", [(5, 1), (20, 2), (60, 3)])
def test_bicgstab_dense_matches_lu(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    x, it, res, info = oc.bicgstab_dense(a, b, np.zeros(n), 500, 1e-12)
    xs = np.linalg.solve(a, b)
    assert info == 0 and res <= 1e-12
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    assert np.linalg.norm(b - a @ x) / np.linalg.norm(b) <= 2e-12
    # Jacobi-preconditioned variant converges to the same solution
    x2, *_ = oc.bicgstab_dense(a, b, np.zeros(n), 500, 1e-12, jacobi=True)
    assert np.linalg.norm(x2 - xs) / np.linalg.norm(xs) < 1e-10


def test_bicgstab_exit_paths():
    n = 12
    rng = np.random.default_rng(4)
    a = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    # ITER = 1 -> INFO = 1, RESID = ||r||/||b|| of that iteration
    x, it, res, info = oc.bicgstab_dense(a, b, np.zeros(n), 1, 1e-14)
    assert (it, info) == (1, 1)
    assert abs(res - np.linalg.norm(b - a @ x) / np.linalg.norm(b)) < 1e-10
    # b = 0, x0 = 0: rho = 0 -> breakdown -10 at ITER 1
    assert oc.bicgstab_dense(a, np.zeros(n), np.zeros(n), 10, 1e-8)[1:] == (1, 1e-8, -10)
    # x0 the solution: absolute ||r0|| <= tol exits with ITER and RESID unchanged
    xs = np.linalg.solve(a, b)
    _, it, res, info = oc.bicgstab_dense(a, b, xs, 123, 1e-8)
    assert (it, res, info) == (123, 1e-8, 0)
    # A = I: s = r - alpha v = 0 at ITER 1 -> early exit, x = alpha p = b
    x, it, res, info = oc.bicgstab_dense(np.eye(n), b, np.zeros(n), 10, 1e-10)
    assert (it, info) == (1, 0) and res <= 1e-15
    np.testing.assert_allclose(x, b, rtol=1e-15)
    # ITER <= 0 -> INFO = -3
    assert oc.bicgstab_dense(a, b, np.zeros(n), 0, 1e-8)[3] == -3


@pytest.mark.parametrize("m", [1, 5, 80])
def test_fgmres_dense_matches_lu(m):
    """FGMRES(m) (Saad 1993) converges to the LU solution; with m >= n it is
    full GMRES and terminates in at most n inner iterations."""
    n = 40
    rng = np.random.default_rng(11)
    # field of values away from 0 so that even GMRES(1) converges
    a = rng.standard_normal((n, n)) + 15 * np.eye(n)
    b = rng.standard_normal(n)
    xs = np.linalg.solve(a, b)
    x, it, res, info = oc.fgmres_dense(a, b, np.zeros(n), 2000, 1e-12, restart=m)
    assert info == 0 and res <= 1e-12
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    if m >= n:
        assert it <= n
    # the true residual is what RESID reports
    assert abs(res - np.linalg.norm(b - a @ x) / np.linalg.norm(b)) < 1e-14


def test_fgmres_exit_paths():
    n = 10
    rng = np.random.default_rng(12)
    a = rng.standard_normal((n, n)) + 5 * np.eye(n)
    b = rng.standard_normal(n)
    x, it, res, info = oc.fgmres_dense(a, b, np.zeros(n), 3, 1e-14, restart=10)
    assert (it, info) == (3, 1)
    x, it, res, info = oc.fgmres_dense(a, np.zeros(n), np.ones(n), 5, 1e-10)
    assert (it, info, res) == (0, 0, 0.0) and not x.any()
    xs = np.linalg.solve(a, b)
    x, it, res, info = oc.fgmres_dense(a, b, xs, 5, 1e-10)
    assert (it, info) == (0, 0)


# ---- operator invariants ----------------------------------------------------------
def _coef_oracle(nx, ny, seed, lr="n"):
    p = abi.default_params(nx, ny)
    p.lr_bd_type = ord(lr)
    p.penalty_k = 1e4
    o = oc.Oracle(p)
    rng = np.random.default_rng(seed)
    cs = abi.cell_shape(p)
    o.set_field(abi.F_RHO, rng.uniform(0.5, 2, cs))
    o.set_field(abi.F_ETA, rng.uniform(1, 10, cs))
    o.set_field(abi.F_BDR_PH_PENALTY, rng.uniform(0, 1, cs))
    return p, o


@pytest.mark.parametrize("op", [abi.OP_VELO1, abi.OP_VELO2, abi.OP_PHMU])
def test_matvec_with_beta_zero_does_not_read_y(op):
    """y = alpha A x + beta y with beta = 0 leaves y unread (BLAS), as the
    GPU sweeps do: Krylov bases hand the matvec fresh, uninitialised memory,
    and 0 * NaN would otherwise poison the result."""
    cfg = geometry.config("tiny")
    o = oc.Oracle(cfg.params)
    geometry.load(o, cfg)
    n = abi.op_size(cfg.params, op)
    x = np.random.default_rng(4).standard_normal(n)
    want = o.matvec(op, 1.5, x, 0.0, np.zeros(n))
    got = o.matvec(op, 1.5, x, 0.0, np.full(n, np.nan))
    assert np.all(np.isfinite(want))
    np.testing.assert_array_equal(got, want)
    y0 = np.arange(n, dtype=np.float64)
    np.testing.assert_array_equal(o.matvec(op, 1.5, x, 2.0, y0.copy()), want + 2.0 * y0)


@pytest.mark.parametrize("op", [abi.OP_VELO1, abi.OP_VELO2])
@pytest.mark.parametrize("lr", ["n", "d"])
def test_velocity_operator_symmetric_positive(op, lr):
    """Away from the u unknowns on the side walls (i = 0, nx, whose mirrored
    ghosts make those rows one-sided) the staggered full-stress operator is
    symmetric: the u-row coefficient of v(i,k+1) equals the v-row coefficient
    of u(i,k), and so on, and so is its sparsity pattern.  A transposed or
    shifted index in any cross term breaks this.  The whole operator has
    eigenvalues with positive real part."""
    p, o = _coef_oracle(12, 9, 5, lr)
    n = abi.n_uv(p)
    A = np.column_stack([o.matvec(op, 1.0, e, 0.0, np.zeros(n)) for e in np.eye(n)])
    mn = (p.nx + 1) * p.ny
    side = np.zeros(n, dtype=bool)
    iu = np.arange(mn) % (p.nx + 1)
    side[:mn] = (iu == 0) | (iu == p.nx)
    keep = ~side
    S = A[np.ix_(keep, keep)]
    assert np.max(np.abs(S - S.T)) <= 1e-12 * np.max(np.abs(A))
    assert np.array_equal(S != 0, S.T != 0)
    assert np.linalg.eigvals(A).real.min() > 0


def test_op_double_gradient_conservative_and_laplacian():
    nx, ny = 10, 7
    rng = np.random.default_rng(6)
    f = rng.standard_normal((ny + 2, nx + 2))
    m = rng.uniform(0.1, 1.0, (ny + 2, nx + 2))
    dx2, dy2 = 0.3, 0.7
    out = oc.op_double_gradient_neum(nx, ny, dx2, dy2, f, m)
    # divergence form with zero-flux walls: the sum over cells vanishes
    assert abs(out[1:-1, 1:-1].sum()) < 1e-12 * np.abs(out).sum()
    # unit mobility: textbook 5-point Laplacian with reflected ghosts
    g = f.copy()
    g[:, 0], g[:, -1], g[0, :], g[-1, :] = g[:, 1], g[:, -2], g[1, :], g[-2, :]
    c = g[1:-1, 1:-1]
    lap = (g[1:-1, 2:] + g[1:-1, :-2] - 2 * c) / dx2 + (g[2:, 1:-1] + g[:-2, 1:-1] - 2 * c) / dy2
    out1 = oc.op_double_gradient_neum(nx, ny, dx2, dy2, f, np.ones_like(m))
    np.testing.assert_allclose(out1[1:-1, 1:-1], lap, rtol=1e-12, atol=1e-12)


# ---- right-hand sides ---------------------------------------------------------------
def _rhs_oracle(nx=10, ny=6):
    p = abi.default_params(nx, ny)
    p.dx, p.dz, p.dt, p.re, p.ca = 0.1, 0.2, 0.01, 2.0, 3.0
    o = oc.Oracle(p)
    cs = abi.cell_shape(p)
    o.set_field(abi.F_RHO, np.ones(cs))
    o.set_field(abi.F_ETA, np.ones(cs))
    o.set_field(abi.F_BDR_PH_USED, np.ones(cs))
    return p, o


def _split(p, rhs):
    mn = (p.nx + 1) * p.ny
    return rhs[:mn].reshape(p.ny, p.nx + 1), rhs[mn:].reshape(p.ny - 1, p.nx)


def test_velocity_rhs_isolated_terms():
    p, o = _rhs_oracle()
    cs = abi.cell_shape(p)
    X, Z = geometry._cell_centres(p)
    zero = np.zeros(cs)
    # pressure gradient of p = a x + c z: fu = -a dt, fv = -c dt
    o.set_field(abi.F_P, 0.7 * X - 0.4 * Z)
    fu, fv = _split(p, o.velo_rhs(1, ph_after=zero, rmu_after=zero))
    np.testing.assert_allclose(fu, -0.7 * p.dt, rtol=1e-12)
    np.testing.assert_allclose(fv, 0.4 * p.dt, rtol=1e-12)
    o.set_field(abi.F_P, zero)
    # inertia: BDF1 rho u re/dt * dt = re u ; BDF2 (2u - u_pre/2) re
    u = np.full(abi.field_shape(p, abi.F_U), 1.5)
    o.set_field(abi.F_U, u)
    fu, _ = _split(p, o.velo_rhs(1, ph_after=zero, rmu_after=zero))
    np.testing.assert_allclose(fu, p.re * 1.5, rtol=1e-12)
    up = np.full_like(u, 0.5)
    fu, _ = _split(p, o.velo_rhs(2, u_pre=up, ph_after=zero, rmu_after=zero))
    np.testing.assert_allclose(fu, (2 * 1.5 - 0.5 / 2) * p.re, rtol=1e-12)
    o.set_field(abi.F_U, np.zeros_like(u))
    # body force: BDF1 rho ggx dt ; BDF2 rho phi_avg ggx dt (TP/VeloSolve_2_combined.f90:212)
    p2 = p.copy()
    p2.ggx = 2.0
    o2 = oc.Oracle(p2)
    for fid in (abi.F_RHO, abi.F_ETA, abi.F_BDR_PH_USED):
        o2.set_field(fid, np.ones(cs))
    fu, _ = _split(p2, o2.velo_rhs(1, ph_after=zero, rmu_after=zero))
    np.testing.assert_allclose(fu, 2.0 * p2.dt, rtol=1e-12)
    fu, _ = _split(p2, o2.velo_rhs(2, ph_after=np.full(cs, 0.5), rmu_after=zero))
    np.testing.assert_allclose(fu, 0.5 * 2.0 * p2.dt, rtol=1e-12)
    # surface tension: ca mu dphi/dx for mu const, phi = b x
    fu, fv = _split(p, o.velo_rhs(1, ph_after=0.3 * X, rmu_after=np.full(cs, 2.0)))
    np.testing.assert_allclose(fu, p.ca * 2.0 * 0.3 * p.dt, rtol=1e-12)
    np.testing.assert_allclose(fv, 0.0, atol=1e-15)
    # convection of u = a x: u du/dx = a^2 x_face, re rho (...) dt subtracted
    xf = (np.arange(-1, p.nx + 2)) * p.dx
    uu = np.tile(0.5 * xf, (p.ny + 2, 1))
    o.set_field(abi.F_U, uu)
    fu, _ = _split(p, o.velo_rhs(1, ph_after=zero, rmu_after=zero))
    uf = 0.5 * xf[1:-1]
    expect = np.tile(p.re * uf - p.re * (uf * 0.5) * p.dt, (p.ny, 1))
    np.testing.assert_allclose(fu, expect, rtol=1e-12, atol=1e-14)
    o.set_field(abi.F_U, np.zeros_like(uu))
    # moving bottom wall: fu(i,0) += dt 2 uwB/dz^2 * eta
    p3 = p.copy()
    p3.uwB = 0.25
    o3 = oc.Oracle(p3)
    for fid in (abi.F_RHO, abi.F_ETA, abi.F_BDR_PH_USED):
        o3.set_field(fid, np.ones(cs))
    fu, _ = _split(p3, o3.velo_rhs(1, ph_after=zero, rmu_after=zero))
    np.testing.assert_allclose(fu[0], p3.dt * 2 * 0.25 / p3.dz ** 2, rtol=1e-12)
    np.testing.assert_allclose(fu[1:], 0.0, atol=1e-15)


def test_ch_rhs_isolated_terms():
    nx, ny = 8, 6
    p = abi.default_params(nx, ny)
    p.I_thickness, p.s_implicit, p.angle, p.phi_mode = 0.05, 1.5, 90.0, 2
    o = oc.Oracle(p)
    cs = abi.cell_shape(p)
    rng = np.random.default_rng(9)
    ph = rng.uniform(-1, 1, cs)
    bu = rng.uniform(0, 1, cs)
    bdr = rng.uniform(0, 1, cs)
    o.set_field(abi.F_PH, ph)
    o.set_field(abi.F_BDR_PH_USED, bu)
    o.set_field(abi.F_BDR_PH, bdr)
    o.set_field(abi.F_ANGLEB, np.full(nx + 1, 90.0))
    o.set_field(abi.F_ANGLET, np.full(nx + 1, 90.0))
    rhs = o.ph_mu_rhs()
    C = nx * ny
    c = (slice(1, -1), slice(1, -1))
    # u = v = 0: f2 = phi (bu + adj)                               :40
    np.testing.assert_allclose(rhs[C:].reshape(ny, nx), ph[c] * (bu[c] + p.bdr_ph_used_adjust), rtol=1e-15)
    # angle 90 and no wall gradient: f1 = s phi/eps - (phi^3 - phi)(1 - psi)/eps
    # plus the wall term cos(90)/sqrt2 (1 - phi_w^2) ~ 0
    f1 = p.s_implicit * ph[c] / p.I_thickness - (ph[c] ** 3 - ph[c]) * (1 - bdr[c]) / p.I_thickness
    np.testing.assert_allclose(rhs[:C].reshape(ny, nx), f1, rtol=1e-12, atol=1e-12)
    # M_0 = (1 - bdr)(phi^2 - 1)^2 over the whole field                :31-35
    np.testing.assert_allclose(o.get_field(abi.F_PH_M_0), (1 - bdr) * (ph ** 2 - 1) ** 2, rtol=1e-14)


def test_oracle_threads_do_not_change_results():
    """The OpenMP loops and the task-parallel pairwise tree give the serial
    results bit for bit (1024x512: above the parallel threshold)."""
    from drd_application_amd import abi, geometry
    from oracle import oracle as O
    cfg = geometry.two_phase_config(1024, 512, 9, n_drops=4, r_drop=(6, 20), n_discs=2, r_disc=(8, 12),
                                    lambda_eta=20.0, penalty=1e5, dt=5e-5, name="omp")
    out = {}
    for nt in (1, 8):
        O.set_threads(nt)
        o = O.Oracle(cfg.params)
        geometry.load(o, cfg)
        nv = abi.n_uv(cfg.params)
        b = np.random.default_rng(3).standard_normal(nv)
        x, it, res, info = o.bicgstab(abi.OP_VELO2, abi.PS_IDENTITY, b, np.zeros(nv), 6, 1e-300)
        nc = 2 * cfg.params.nx * cfg.params.ny
        bc = o.ph_mu_rhs()
        y, it2, res2, info2 = o.bicgstab(abi.OP_PHMU, abi.PS_PHMU_BLOCK, bc, np.zeros(nc), 6, 1e-300)
        out[nt] = (x, res, y, res2, O.pairwise(b, b) if hasattr(O, "pairwise") else 0.0)
        o.close()
    O.set_threads(1)
    a, b8 = out[1], out[8]
    np.testing.assert_array_equal(a[0], b8[0])
    np.testing.assert_array_equal(a[2], b8[2])
    assert a[1] == b8[1] and a[3] == b8[3] and a[4] == b8[4]


def test_velocity_rhs_mf_terms():
    """The MF variant's two different RHS terms (MF/VeloSolve_2_combined.f90):
    surface tension ca mu (phi(i) - phi(i-1))/dr (1 - mean bdr_ph) (:64-65)
    and the BDF2 body force plain ggx (:201), and its always-on inflow."""
    p, _ = _rhs_oracle()
    p.variant = abi.VARIANT_MF
    p.ggx = 2.0
    o = oc.Oracle(p)
    cs = abi.cell_shape(p)
    for fid in (abi.F_RHO, abi.F_ETA):
        o.set_field(fid, np.ones(cs))
    X, Z = geometry._cell_centres(p)
    zero = np.zeros(cs)
    # plain ggx in BDF2 whatever phi is (TP would give phi_avg ggx)
    fu, _ = _split(p, o.velo_rhs(2, ph_after=np.full(cs, 0.5), rmu_after=zero))
    np.testing.assert_allclose(fu, 2.0 * p.dt, rtol=1e-12)
    o.p.ggx = 0.0
    o2 = oc.Oracle(o.p)
    for fid in (abi.F_RHO, abi.F_ETA):
        o2.set_field(fid, np.ones(cs))
    # surface tension with bdr_ph = 0.25 everywhere: (1 - 0.25) ca mu b dt; bdr_ph_used plays no part
    o2.set_field(abi.F_BDR_PH, np.full(cs, 0.25))
    o2.set_field(abi.F_BDR_PH_USED, np.full(cs, 0.1))
    fu, fv = _split(p, o2.velo_rhs(1, ph_after=0.3 * X, rmu_after=np.full(cs, 2.0)))
    np.testing.assert_allclose(fu, 0.75 * p.ca * 2.0 * 0.3 * p.dt, rtol=1e-12)
    np.testing.assert_allclose(fv, 0.0, atol=1e-15)
    fu, fv = _split(p, o2.velo_rhs(1, ph_after=-0.2 * Z, rmu_after=np.full(cs, 1.5)))
    np.testing.assert_allclose(fu, 0.0, atol=1e-15)
    np.testing.assert_allclose(fv, 0.75 * p.ca * 1.5 * -0.2 * p.dt, rtol=1e-12)
    # inflow on although lr_bd_type is 'n': u(0,k) rows get 2 uL eta 2 dt / dr^2
    q = o.p.copy()
    q.UU, q.bdr_low, q.bdr_upp, q.twall = 1.0, 0.0, q.ny * q.dz, 1e-9
    o3 = oc.Oracle(q)
    for fid in (abi.F_RHO, abi.F_ETA):
        o3.set_field(fid, np.ones(cs))
    fu, _ = _split(q, o3.velo_rhs(1, ph_after=zero, rmu_after=zero, t=1.0))
    zc = (np.arange(q.ny) + 0.5) * q.dz
    h = q.bdr_upp - q.bdr_low
    uL = q.UU * (-((zc - q.bdr_low - h / 2) ** 2) / h / h * 4 + 1)
    np.testing.assert_allclose(fu[:, 0], 2.0 * uL * 2.0 * q.dt / q.dx ** 2, rtol=1e-12)
    np.testing.assert_allclose(fu[:, -1], 2.0 * uL * 2.0 * q.dt / q.dx ** 2, rtol=1e-12)
    np.testing.assert_allclose(fu[:, 1:-1], 0.0, atol=1e-15)
