"""Particle coupling reductions on the GPU (drd_particle_force,
drd_particle_velocity; TP/particles_move.f90:758-920) against the oracle's
restatement, and the spread force entering drd_step's momentum equation
(TP/VeloSolve_2_combined.f90:242) as in the oracle's step.

The device's tanh and sqrt are not the host library's, and it sums each
body's weights as a tree of row chunks where the reference sums them in
order, so the bar is a relative 1e-12.  A step with the same force fields is
bit for bit the oracle's."""
import numpy as np
import pytest

from drd_application_amd import abi, geometry

pytestmark = pytest.mark.gpu


def _pair(cfg):
    from drd_application_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    g, o = GpuSolver(cfg.params), Oracle(cfg.params)
    geometry.load(g, cfg)
    geometry.load(o, cfg)
    return g, o


def _bodies(rng, n, lx, ly):
    # centres anywhere, the periodic edges included; radii of particles and swimmers
    b = np.stack([rng.uniform(0, lx, n), rng.uniform(0, ly, n), rng.choice([0.03, 0.05], n),
                  rng.normal(size=n), rng.normal(size=n)], axis=1)
    if n >= 8:  # bodies at the four periodic corners and across each edge
        b[:8, :2] = [[0.001, 0.001], [lx - 0.001, 0.001], [0.001, ly - 0.001], [lx - 0.001, ly - 0.001],
                     [0.0, ly / 2], [lx / 2, 0.0], [lx - 1e-9, ly / 3], [lx / 3, ly - 1e-9]]
    return b


# 520 bodies: the spread builds its per-tile lists in batches of 256 bodies,
# so three batches run; I_thickness 0.002 keeps each body's cutoff disc
# (20 sqrt2 I beyond its radius) small enough for the pruning to act.
@pytest.mark.parametrize("nx,ny,nbody,I", [(64, 32, 0, None), (96, 48, 7, None), (257, 130, 40, None),
                                           (200, 100, 520, 0.002)])
def test_particle_reductions_match_oracle(nx, ny, nbody, I):
    print(f"particle reductions {nx}x{ny} bodies {nbody}", flush=True)
    cfg = geometry.tp_case(nx, ny, lambda_eta=10.0, tol=1e-8)
    p = cfg.params
    if I is not None:
        p.I_thickness = I
    lx, ly = nx * p.dx, ny * p.dz
    g, o = _pair(cfg)
    rng = np.random.default_rng(nx + nbody)
    u = rng.normal(size=g.get_field(abi.F_U).shape)
    v = rng.normal(size=g.get_field(abi.F_V).shape)
    for s in (g, o):
        s.set_field(abi.F_U, u)
        s.set_field(abi.F_V, v)
    b = _bodies(rng, nbody, lx, ly)
    g.particle_force(b, lx, ly)
    o.particle_force(b, lx, ly)
    for fid in (abi.F_PFORCE1, abi.F_PFORCE2):
        a, r = g.get_field(fid), o.get_field(fid)
        if nbody == 0:
            assert not a.any()
        else:
            np.testing.assert_allclose(a, r, rtol=0, atol=1e-12 * np.abs(r).max())
    if nbody:
        ug, vg = g.particle_velocity(b, lx, ly)
        uo, vo = o.particle_velocity(b, lx, ly)
        np.testing.assert_allclose(ug, uo, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(vg, vo, rtol=1e-12, atol=1e-14)
    g.close()
    o.close()


def test_step_uses_particle_force():
    print("step with particle force 64x32", flush=True)
    cfg = geometry.tp_case(64, 32, lambda_eta=10.0, tol=1e-10)
    g, o = _pair(cfg)
    rng = np.random.default_rng(7)
    shape = g.get_field(abi.F_PFORCE1).shape
    for fid in (abi.F_PFORCE1, abi.F_PFORCE2):
        f = 50.0 * rng.normal(size=shape)
        g.set_field(fid, f)
        o.set_field(fid, f)
    t = 0.0
    for bdf in (1, 2):
        g.step(bdf, t)
        o.step(bdf, t)
        t += cfg.params.dt
    for fid in (abi.F_U, abi.F_V, abi.F_P):
        np.testing.assert_array_equal(g.get_field(fid), o.get_field(fid), err_msg=abi.FIELD_NAMES[fid])
    g.close()
    o.close()
