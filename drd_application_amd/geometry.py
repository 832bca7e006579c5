"""Seeded synthetic inputs for the solve path (droplets, solid masks, coefficients).

The reference builds its geometry in `finit` (TP/finit.f90:4-628) by
commenting blocks in and out and fixes the grid at compile time.  These
generators produce the same kinds of fields at run time, deterministically
from a seed (SplitMix64 -> uniform doubles), and derive the secondary fields
the way TP/finit.f90:540-588 does:

  bdr_ph_grad = |grad bdr_ph| (central differences)        TP/finit.f90:544-549
  ph_M_inner  = 1 - bdr_ph, bdr_ph_used = 1 - bdr_ph      TP/finit.f90:569-571
  ph          = (1 - bdr_ph) * phi                         TP/finit.f90:451
  rmu         = -eps*lap(ph) - ph/eps + ph^3/eps           TP/finit.f90:474-480
  angleB = angleT = 90 in a complex domain                 TP/finit.f90:586-588

Viscosity follows the phase field, eta = 1 + (lambda_eta - 1)(1 + phi)/2
(the two-phase form of TP/interpolate_coef.f90:92-112), rho = 1.
Arrays use the abi layout: a[k+1, i+1] is the Fortran element (i, k).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform() in [0, 1) with 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def splitmix_uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """n SplitMix64 uniforms in [lo, hi), vectorised: element j is the
    (j+1)-th output of SplitMix64(seed), identical to SplitMix64.uniform()."""
    with np.errstate(over="ignore"):
        j = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + j * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


@dataclass
class Config:
    name: str
    params: abi.Params
    fields: dict = field(default_factory=dict)  # field id -> ndarray
    steps: int = 1
    notes: str = ""


def _cell_centres(p: abi.Params):
    i = np.arange(-1, p.nx + 1)
    k = np.arange(-1, p.ny + 1)
    x = (i + 0.5) * p.dx
    z = (k + 0.5) * p.dz
    return np.meshgrid(x, z)  # shapes (ny+2, nx+2)


def _neumann_ghosts(a: np.ndarray) -> np.ndarray:
    a[:, 0] = a[:, 1]
    a[:, -1] = a[:, -2]
    a[0, :] = a[1, :]
    a[-1, :] = a[-2, :]
    return a


def _ghosts(a: np.ndarray, periodic: bool) -> np.ndarray:
    """Ghost rows copy their neighbour row (Neumann in z); ghost columns copy
    their neighbour column, or with periodic the periodic image column
    (f(-1) = f(nx-1), f(nx) = f(0)), wall-ghost rows included."""
    if not periodic:
        return _neumann_ghosts(a)
    a[0, :] = a[1, :]
    a[-1, :] = a[-2, :]
    a[:, 0] = a[:, -2]
    a[:, -1] = a[:, 1]
    return a


def _images(p: abi.Params, cx: float) -> list:
    """x positions of a shape centred at cx: itself, and with periodic x its
    images one domain length to either side."""
    if not abi.periodic_x(p):
        return [cx]
    L = p.nx * p.dx
    return [cx - L, cx, cx + L]


def _window(p: abi.Params, cx: float, cz: float, reach: float):
    """Ghosted index ranges (cols, rows) of the cells whose centres lie within
    `reach` of (cx, cz) in each coordinate, and those centres."""
    i0 = max(-1, int(math.floor((cx - reach) / p.dx - 0.5)) - 1)
    i1 = min(p.nx, int(math.ceil((cx + reach) / p.dx - 0.5)) + 1)
    k0 = max(-1, int(math.floor((cz - reach) / p.dz - 0.5)) - 1)
    k1 = min(p.ny, int(math.ceil((cz + reach) / p.dz - 0.5)) + 1)
    x = (np.arange(i0, i1 + 1) + 0.5) * p.dx
    z = (np.arange(k0, k1 + 1) + 0.5) * p.dz
    return slice(k0 + 1, k1 + 2), slice(i0 + 1, i1 + 2), x, z


# tanh(-a) is exactly -1.0 in double precision for a >= 19.1, so a droplet or
# disc changes no cell beyond TANH_SAT * sqrt2 * width from its rim and each
# is evaluated on that window only (the fields equal the full-grid forms).
TANH_SAT = 19.5


def droplet_phase(p: abi.Params, rng: SplitMix64, n: int, rmin: float, rmax: float, eps: float):
    """phi = max_d tanh((r_d - |x - c_d|)/(sqrt2 eps)), centres uniform in the domain."""
    X, Z = _cell_centres(p)
    L, H = p.nx * p.dx, p.ny * p.dz
    phi = -np.ones_like(X)
    for _ in range(n):
        cx, cz = rng.uniform(0, L), rng.uniform(0, H)
        r = rng.uniform(rmin, rmax)
        for xc in _images(p, cx):
            ks, is_, x, z = _window(p, xc, cz, r + TANH_SAT * math.sqrt(2.0) * eps)
            if x.size == 0:
                continue
            d = r - np.hypot(x[None, :] - xc, z[:, None] - cz)
            w = phi[ks, is_]
            np.maximum(w, np.tanh(d / (math.sqrt(2.0) * eps)), out=w)
    return phi


def disc_mask(p: abi.Params, rng: SplitMix64, n: int, rmin: float, rmax: float, width: float):
    """bdr_ph = 1/2 (1 + tanh(-d_s/(sqrt2 w))) for the union of n seeded discs."""
    X, Z = _cell_centres(p)
    L, H = p.nx * p.dx, p.ny * p.dz
    out = np.zeros_like(X)
    for _ in range(n):
        cx, cz = rng.uniform(0, L), rng.uniform(0, H)
        r = rng.uniform(rmin, rmax)
        for xc in _images(p, cx):
            ks, is_, x, z = _window(p, xc, cz, r + TANH_SAT * math.sqrt(2.0) * width)
            if x.size == 0:
                continue
            ds = np.hypot(x[None, :] - xc, z[:, None] - cz) - r
            w = out[ks, is_]
            np.maximum(w, 0.5 * (1.0 + np.tanh(-ds / (math.sqrt(2.0) * width))), out=w)
    return out


def serpentine_mask(p: abi.Params, amp: float, waves: float, half_width: float, width: float):
    """Solid outside the channel |z - (0.5 + amp sin(waves pi x))| < half_width."""
    X, Z = _cell_centres(p)
    zc = 0.5 + amp * np.sin(waves * math.pi * X)
    ds = half_width - np.abs(Z - zc)  # > 0 inside the channel
    return 0.5 * (1.0 + np.tanh(-ds / (math.sqrt(2.0) * width)))


def derived_fields(p: abi.Params, phi: np.ndarray, bdr: np.ndarray, lambda_eta: float,
                   sharp_penalty: bool = False) -> dict:
    """Secondary fields exactly as TP/finit.f90 derives them from bdr_ph and phi
    (ghost columns periodic with periodic x)."""
    per = abi.periodic_x(p)
    bdr = _ghosts(bdr.copy(), per)
    ph = _ghosts((1.0 - bdr) * phi, per)
    eps = p.I_thickness
    dr2, dz2 = p.dx * p.dx, p.dz * p.dz
    rmu = np.zeros_like(ph)
    c = ph[1:-1, 1:-1]
    lap = (ph[1:-1, 2:] + ph[1:-1, :-2] - 2 * c) / dr2 + (ph[2:, 1:-1] + ph[:-2, 1:-1] - 2 * c) / dz2
    rmu[1:-1, 1:-1] = -eps * lap - c / eps + c ** 3 / eps
    rmu = _ghosts(rmu, per)
    grad = np.zeros_like(bdr)
    gx = (bdr[1:-1, 2:] - bdr[1:-1, :-2]) / 2.0 / p.dx
    gz = (bdr[2:, 1:-1] - bdr[:-2, 1:-1]) / 2.0 / p.dz
    grad[1:-1, 1:-1] = np.sqrt(gx ** 2 + gz ** 2)
    if per:
        grad = _ghosts(grad, True)
    # viscosity of the phase field (phi in [-1,1]); solids keep eta of the fluid
    eta = 1.0 + (lambda_eta - 1.0) * (1.0 + ph) / 2.0
    rho = np.ones_like(ph)
    chi = (bdr > 0.5).astype(np.float64) if sharp_penalty else bdr.copy()
    ang = np.full(p.nx + 1, 90.0)
    nu = abi.field_shape(p, abi.F_U)
    nv = abi.field_shape(p, abi.F_V)
    return {
        abi.F_U: np.zeros(nu), abi.F_V: np.zeros(nv), abi.F_P: np.zeros_like(ph),
        abi.F_PH: ph, abi.F_RMU: rmu, abi.F_RHO: rho, abi.F_ETA: _ghosts(eta, per),
        abi.F_BDR_PH: bdr, abi.F_BDR_PH_USED: 1.0 - bdr, abi.F_BDR_PH_PENALTY: chi,
        abi.F_BDR_PH_GRAD: grad, abi.F_PH_M_INNER: 1.0 - bdr,
        abi.F_PH_M_0: 1.0 - bdr, abi.F_ANGLEB: ang, abi.F_ANGLET: ang.copy(),
        abi.F_U_PRE: np.zeros(nu), abi.F_V_PRE: np.zeros(nv), abi.F_PH_PRE: ph.copy(),
        abi.F_RMU_PRE: rmu.copy(), abi.F_PSI1: np.zeros_like(ph), abi.F_PSI2: np.zeros_like(ph),
    }


def two_phase_config(nx: int, ny: int, seed: int, *, n_drops: int, r_drop: tuple,
                     n_discs: int = 0, r_disc: tuple = (10, 16), serpentine: bool = False,
                     lambda_eta: float = 10.0, penalty: float = 1e4, dt: float = 1e-4,
                     tol: float = 1e-10, steps: int = 10, name: str = "custom",
                     periodic: bool = False) -> Config:
    """Two-phase CH + NS on an nx x ny grid of height 1 (dx = dz = 1/ny);
    periodic: periodic in x (droplets and discs wrap across x = 0)."""
    p = abi.default_params(nx, ny)
    if periodic:
        p.lr_bd_type = abi.LR_PERIODIC
    p.dx = p.dz = 1.0 / ny
    p.dt = dt
    p.re, p.ca, p.xld = 1.0, 1.0, 1.0e-3
    p.s_implicit = 1.5
    p.I_thickness = 1.5 * p.dx
    p.ggx, p.ggz = 1.0, 0.0
    p.tol = tol
    p.penalty_k = penalty
    p.angle = 90.0
    rng = SplitMix64(seed)
    eps = p.I_thickness
    phi = droplet_phase(p, rng, n_drops, r_drop[0] * p.dx, r_drop[1] * p.dx, eps)
    if serpentine:
        bdr = serpentine_mask(p, 0.3, 8.0, 0.125, 2 * p.dx)
    else:
        bdr = disc_mask(p, rng, n_discs, r_disc[0] * p.dx, r_disc[1] * p.dx, 2 * p.dx) if n_discs \
            else np.zeros_like(phi)
    cfg = Config(name, p, derived_fields(p, phi, bdr, lambda_eta), steps)
    return cfg


def tp_case(nx: int, ny: int | None = None, *, lambda_eta: float = 1.0, steps: int = 10,
            tol: float = 1e-8) -> Config:
    """The reference's shipped TP run (TP/input, TP/finit.f90 geometry type 4)
    on an nx x ny grid of the unit square: half-plane solid below z = 0.2
    (smoothed over I_thickness/2, sharp penalty mask), one droplet of radius
    0.2 resting on it, xld = 0.05, re = 3, ca = 12, eps = 0.01 * 128/nx,
    s = 1.5, dt = 5e-5, contact angle 60 (90 in the complex domain)."""
    ny = nx if ny is None else ny
    p = abi.default_params(nx, ny)
    p.dx, p.dz = 1.0 / nx, 1.0 / ny
    p.xld, p.re, p.ca, p.s_implicit, p.tol = 0.05, 3.0, 12.0, 1.5, tol
    p.I_thickness = 0.01 * 128.0 / max(nx, ny)
    p.dt = 5e-5 * (128.0 / max(nx, ny)) ** 2
    p.angle = 60.0
    p.penalty_k = 1e8
    X, Z = _cell_centres(p)
    eps_s = p.I_thickness / 2.0
    bdr = (np.tanh(-(Z - 0.2) / math.sqrt(2.0) / eps_s) + 1.0) / 2.0
    phi = np.tanh(-(np.sqrt((X - 0.5) ** 2 + (Z - 0.4) ** 2) - 0.2) / (math.sqrt(2.0) * p.I_thickness))
    return Config(f"tp{nx}x{ny}", p, derived_fields(p, phi, bdr, lambda_eta, sharp_penalty=True), steps,
                  "reference TP case")


def smooth_field(shape, x, z, rng: SplitMix64, modes: int = 16, kmax: float = 6.0) -> np.ndarray:
    """Sum of `modes` seeded cosines a cos(2 pi (kx x + kz z) + theta), wave
    numbers uniform in [-kmax, kmax] (per unit length), amplitudes ~ 1/|k|."""
    g = np.zeros(shape)
    for _ in range(modes):
        kx, kz = rng.uniform(-kmax, kmax), rng.uniform(-kmax, kmax)
        th = rng.uniform(0.0, 2.0 * math.pi)
        a = 1.0 / max(1.0, math.hypot(kx, kz))
        g += a * np.cos(2.0 * math.pi * (kx * x + kz * z) + th)
    return g


def velocity_system(n: int, seed: int, *, max_contrast: float = 1.0e3, n_discs: tuple = (3, 8),
                    r_disc: tuple = (0.02, 0.05), penalty: float = 1.0e6, dt: float = 5.0e-5):
    """One learned-preconditioner test system on an n x n grid of the unit
    square (the 1024^2 workload at n = 1024):

        (I - dt div(eta grad) + dt chi R) u = f

    in the reference's BDF2 velocity operator (TP/VeloSolve_2_combined.f90:
    536-676) with rho = 1 and re = 2/3, so the mass term 1.5 re rho is I.
    eta = exp(log(C) phi~), phi~ in [0, 1] a seeded random phase field
    (tanh of a smooth random field), contrast C log-uniform in [1, max_contrast];
    chi = smoothed union of seeded discs (radius r_disc of the height,
    smoothing 2 cells) with R = penalty; f a smooth random field on the u and v
    faces.  Returns (Config, f packed as the velocity Krylov vector)."""
    p = abi.default_params(n, n)
    p.dx = p.dz = 1.0 / n
    p.dt = dt
    p.re = 2.0 / 3.0
    p.penalty_k = penalty
    p.tol = 1.0e-10
    p.maxit_velo2 = 5000
    rng = SplitMix64(seed)
    X, Z = _cell_centres(p)
    contrast = 10.0 ** rng.uniform(0.0, math.log10(max_contrast))
    g = smooth_field(X.shape, X, Z, rng)
    phi = 0.5 * (1.0 + np.tanh(g / (0.15 * g.std())))
    eta = np.exp(math.log(contrast) * phi)
    nd = int(n_discs[0] + (n_discs[1] - n_discs[0] + 1) * rng.uniform())
    bdr = disc_mask(p, rng, nd, r_disc[0], r_disc[1], 2 * p.dx)
    f = derived_fields(p, np.zeros_like(phi), bdr, 1.0)
    f[abi.F_ETA] = _neumann_ghosts(eta)
    f[abi.F_RHO] = np.ones_like(eta)
    f[abi.F_BDR_PH_PENALTY] = bdr.copy()
    # right-hand side on the faces: u(i, k) at (i dx, (k+0.5) dz), v(i, k) at ((i+0.5) dx, k dz)
    iu, ku = np.meshgrid(np.arange(n + 1) * p.dx, (np.arange(n) + 0.5) * p.dz)
    iv, kv = np.meshgrid((np.arange(n) + 0.5) * p.dx, np.arange(1, n) * p.dz)
    fu = smooth_field(iu.shape, iu, ku, rng)
    fv = smooth_field(iv.shape, iv, kv, rng)
    b = np.concatenate([fu.ravel(), fv.ravel()])
    cfg = Config(f"velo{n}_{seed}", p, f, 1, f"contrast {contrast:.3g}, {nd} discs")
    return cfg, b


# nx of the batched workload: the multiples of 32 from 128 to 1024
C2_SIZES = tuple(range(128, 1025, 32))


def batch_configs(count: int = 256, seed: int = 2000, sizes=(128, 256, 512, 1024), r_ref: float = 1.0e6,
                  dt: float = 5.0e-5, tol: float = 1.0e-10, steps: int = 20):
    """The batched workload: `count` independent two-phase instances.

    nx log-uniform over `sizes` (the draw taken to the nearest size in log;
    C2_SIZES, the multiples of 32, for the workload as specified; powers of
    two for the fast-Poisson preconditioner, whose FFTs are radix-2^k), ny =
    nx/2, height 1; eta ratio
    uniform in [1, 100]; 0-8 seeded discs (radius 6-16 cells) whose
    resistance R is log-uniform in [1e4, 1e6]; 2-12 droplets (radius 4-16
    cells).  Every instance of one size shares its drd_params (penalty_k =
    r_ref); an instance's own R enters through bdr_ph_penalty = (R / r_ref)
    * bdr_ph.  FP32 velocity preconditioner, CH block Jacobi.  Returns a
    list of Config (instance i seeded with seed + i)."""
    out = []
    lo, hi = math.log(sizes[0]), math.log(sizes[-1])
    for i in range(count):
        rng = SplitMix64(seed + i)
        lx = lo + (hi - lo) * rng.uniform()
        nx = min(sizes, key=lambda s: abs(math.log(s) - lx))
        ny = nx // 2
        lam = rng.uniform(1.0, 100.0)
        n_discs = int(9 * rng.uniform())
        n_drops = 2 + int(11 * rng.uniform())
        r_pen = 10.0 ** rng.uniform(4.0, 6.0)
        cfg = two_phase_config(nx, ny, seed * 7 + i, n_drops=n_drops, r_drop=(4, 16), n_discs=n_discs,
                               r_disc=(6, 16), lambda_eta=lam, penalty=r_ref, dt=dt, tol=tol, steps=steps,
                               name=f"b{i}")
        cfg.params.velo_prec = abi.PREC_DCT32
        cfg.params.ch_prec = abi.CH_PREC_BLOCK
        cfg.fields[abi.F_BDR_PH_PENALTY] = cfg.fields[abi.F_BDR_PH_PENALTY] * (r_pen / r_ref)
        cfg.notes = f"{nx}x{ny}, eta ratio {lam:.1f}, {n_discs} discs R={r_pen:.3g}, {n_drops} droplets"
        out.append(cfg)
    return out


def add_body_mask(p: abi.Params, bdr: np.ndarray, table: np.ndarray, width: float) -> np.ndarray:
    """max(bdr, 1/2 (1 + tanh(-d/(sqrt2 w)))) for the discs (x, y, r) of table,
    each evaluated on its own window (the full-grid value beyond it is 0)."""
    out = bdr.copy()
    for x, y, r in table[:, :3]:
        ks, is_, xc, zc = _window(p, x, y, r + TANH_SAT * math.sqrt(2.0) * width)
        d = np.hypot(xc[None, :] - x, zc[:, None] - y) - r
        w = out[ks, is_]
        np.maximum(w, 0.5 * (1.0 + np.tanh(-d / (math.sqrt(2.0) * width))), out=w)
    return out


def c4_config(nx: int = 4096, ny: int = 2048, seed: int = 4000, n_particles: int = 400, n_fibers: int = 64,
              beads: int = 20, steps: int = 200):
    """The full combined workload: two-phase flow in the serpentine channel of
    c1 (centreline 0.5 + 0.3 sin(8 pi x), width 0.25, wall smoothing 2
    cells) on an nx x ny grid of height 1, 40 droplets of the physical sizes
    of c1 (radii 20-80 cells at 4096 x 2048), viscosity ratio 100, dt = 2e-5;
    n_particles rigid discs of radius 6-10 cells and n_fibers chains of
    `beads` beads of radius 4 cells (touching: spacing 8 cells), seeded
    inside the channel, all diffuse-resistance domains with R = 1e6
    (bdr_ph and the penalty mask are the union of the walls and the bodies
    at their starting positions, smoothed over 2 cells; the bodies move
    about a cell in 200 steps, and the mask is not rebuilt as they do).
    Returns (Config, bodies.Bodies)."""
    from .bodies import Bodies
    scale = nx / 2048.0
    cfg = two_phase_config(nx, ny, seed, n_drops=40, r_drop=(10 * scale, 40 * scale), serpentine=True,
                           lambda_eta=100.0, penalty=1e6, dt=2e-5, steps=steps, name="c4")
    p = cfg.params
    lx, dx = nx * p.dx, p.dx
    rng = SplitMix64(seed + 17)

    def in_channel(x):
        return 0.5 + 0.3 * math.sin(8.0 * math.pi * x) + rng.uniform(-0.08, 0.08)

    parts = np.zeros((n_particles, 3))
    for j in range(n_particles):
        x = rng.uniform(0.0, lx)
        parts[j] = (x, in_channel(x), rng.uniform(6.0, 10.0) * dx)
    rb, l0 = 4.0 * dx, 8.0 * dx
    fibers = []
    for _ in range(n_fibers):
        x = rng.uniform(0.0, lx)
        y = in_channel(x)
        a = rng.uniform(0.0, 2.0 * math.pi)
        s = (np.arange(beads) - (beads - 1) / 2.0) * l0
        fibers.append(np.stack([x + s * math.cos(a), y + s * math.sin(a)], axis=1))
    # Spring and bending constants for a stable explicit coupling at this dt:
    # a stretch d of one bond accelerates the fluid of a bead (area pi rb^2,
    # density 1) by ks d / (pi rb^2), so dt^2 ks / (pi rb^2) stays below 1e-2;
    # the bending force kb G / l0 / rd ~ kb d / l0^3 is held to a tenth of it.
    ks = 1e-2 * math.pi * rb * rb / (p.dt * p.dt)
    b = Bodies(parts, lx, fibers, bead_radius=rb, ks=ks, kb=0.1 * ks * l0 * l0, l0=l0, theta0=0.0)
    # the bodies join the walls in bdr_ph; the phase, mobilities and penalty follow
    table = b.table()
    walls = cfg.fields[abi.F_BDR_PH]
    bdr = add_body_mask(p, walls, np.concatenate([table[:, :1], table[:, 1:3]], axis=1), 2 * dx)
    phi = np.where(walls < 1.0, cfg.fields[abi.F_PH] / np.maximum(1.0 - walls, 1e-300), -1.0)
    cfg.fields = derived_fields(p, np.clip(phi, -1.0, 1.0), bdr, 100.0)
    cfg.notes = (f"{nx}x{ny} serpentine, 40 droplets, {n_particles} particles, {n_fibers} fibers of {beads} beads")
    return cfg, b


def read_domain_bin(path: str, nx: int, ny: int) -> np.ndarray:
    """A domain file of the reference (read_var TP/inout.f90:75-87): nx * ny
    FP64 values in a stream file, the x index outer, ((var(i,k), k=1,ny),
    i=1,nx).  Returns the (ny, nx) array a[k, i] = var(i+1, k+1), the
    interior of a cell field in the abi layout."""
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != nx * ny:
        raise ValueError(f"{path}: {raw.size} values, expected {nx} x {ny}")
    return raw.reshape(nx, ny).T.copy()


def mf_case(nx: int = 128, ny: int = 64, seed: int = 5, steps: int = 10, tol: float = 1e-10) -> Config:
    """The microfluidics variant (MF/, phi in [0, 1], PhSolve for the CH step,
    inflow through the left side): a serpentine channel (centreline 0.5 + 0.2
    sin(2 pi x), half width 0.2, wall smoothing 2 cells) on an nx x ny grid of
    height 1, a parabolic inflow of peak 1 between the channel walls at x = 0,
    3 seeded droplets (phi = (1 + tanh(d / (sqrt2 eps))) / 2), penalty 1e5
    (MF/global.f90:91), eta ratio 10, re 1, ca 1, xld 0.05, contact angle 90,
    dt 5e-5; the reference's MF iteration limits (800 for the BDF2 momentum
    solve, MF/VeloSolve_2_combined.f90:398)."""
    p = abi.default_params(nx, ny)
    p.variant = abi.VARIANT_MF
    p.phi_mode = 1
    p.dx = p.dz = 1.0 / ny
    p.dt, p.re, p.ca, p.xld = 5e-5, 1.0, 1.0, 0.05
    p.s_implicit, p.I_thickness = 1.5, 1.5 * p.dx
    p.tol, p.penalty_k, p.angle = tol, 1e5, 90.0
    p.maxit_velo1, p.maxit_velo2 = 800, 800
    p.UU, p.twall = 1.0, 0.01
    p.bdr_low, p.bdr_upp = 0.3, 0.7
    bdr = serpentine_mask(p, 0.2, 2.0, 0.2, 2 * p.dx)
    rng = SplitMix64(seed)
    X, Z = _cell_centres(p)
    phi = np.zeros_like(X)
    L = nx * p.dx
    for _ in range(3):
        cx = rng.uniform(0.2 * L, 0.8 * L)
        cz = 0.5 + 0.2 * math.sin(2.0 * math.pi * cx) + rng.uniform(-0.05, 0.05)
        r = rng.uniform(4, 8) * p.dx
        d = r - np.hypot(X - cx, Z - cz)
        phi = np.maximum(phi, 0.5 * (1.0 + np.tanh(d / (math.sqrt(2.0) * p.I_thickness))))
    f = derived_fields(p, phi, bdr, 10.0)
    # eta of the [0, 1] phase: 1 + (lambda - 1) phi
    f[abi.F_ETA] = _neumann_ghosts(1.0 + 9.0 * f[abi.F_PH])
    return Config(f"mf{nx}x{ny}", p, f, steps, "MF variant, serpentine channel with inflow")


def config(name: str) -> Config:
    """Named workloads.

    c0     256 x 128, 6 droplets r in [8,20] cells, 3 discs r in [10,16],
           eta ratio 10, R = 1e4, dt = 1e-4, tol 1e-10 (Neumann-x twin of the
           periodic case: the reference's TP path has no periodic x).
    c0p    the same, periodic in x (the workload's own boundary).
    c1     2048 x 1024 serpentine channel (0.5 + 0.3 sin(8 pi x), width 0.25),
           40 droplets r in [10,40], eta ratio 100, R = 1e6, dt = 5e-5, 20 steps.
    small  64 x 32 version of c0 for parity tests.
    tiny   32 x 16.
    """
    if name == "c0":
        return two_phase_config(256, 128, 1000, n_drops=6, r_drop=(8, 20), n_discs=3,
                                lambda_eta=10.0, penalty=1e4, dt=1e-4, steps=50, name=name)
    if name == "c0p":
        return two_phase_config(256, 128, 1000, n_drops=6, r_drop=(8, 20), n_discs=3,
                                lambda_eta=10.0, penalty=1e4, dt=1e-4, steps=50, name=name, periodic=True)
    if name == "smallp":
        return two_phase_config(64, 32, 7, n_drops=3, r_drop=(3, 6), n_discs=2, r_disc=(3, 5),
                                lambda_eta=10.0, penalty=1e4, dt=1e-4, steps=10, name=name, periodic=True)
    if name == "c1":
        return two_phase_config(2048, 1024, 1001, n_drops=40, r_drop=(10, 40), serpentine=True,
                                lambda_eta=100.0, penalty=1e6, dt=5e-5, steps=20, name=name)
    if name == "small":
        return two_phase_config(64, 32, 7, n_drops=3, r_drop=(3, 6), n_discs=2, r_disc=(3, 5),
                                lambda_eta=10.0, penalty=1e4, dt=1e-4, steps=10, name=name)
    if name == "tiny":
        return two_phase_config(32, 16, 3, n_drops=2, r_drop=(2, 4), n_discs=1, r_disc=(2, 3),
                                lambda_eta=10.0, penalty=1e4, dt=1e-4, steps=10, name=name)
    raise KeyError(name)


def load(solver, cfg: Config) -> None:
    """Pushes every field of cfg into a solver (GpuSolver or the oracle)."""
    for fid, a in cfg.fields.items():
        solver.set_field(fid, a)
