"""The learned correction of the velocity preconditioner on the GPU.

* network: forward output, loss, every parameter block's gradient and one
  Adam step of the hand-written HIP kernels against the CPU oracle
  (oracle/unet_ref.py, itself pinned to torch autograd and finite
  differences), the oracle mirroring where the kernels round to BF16;
* training on pairs the solver emits lowers the loss, and FGMRES with the
  trained correction converges on held-out systems to the FP64 true
  residual (checked with the oracle's operator).  The correction does not
  earn its place: on the full workload (20k pairs, 200 held-out 1024^2
  systems) it needs 2.7% more FGMRES iterations and 7.7% more time than the
  plain FP32 DCT preconditioner.  Here the iterations it needs are printed
  and held within 5% of plain FGMRES (the step alpha keeps each corrected
  residual within the plain one's, so it cannot stall the solve);
* drd_step with velo_prec = LEARNED: the judged first solve, and the
  automatic fallback to the plain method when the network is bad.
"""
import numpy as np
import pytest

from drd_application_amd import abi, geometry
from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _net_case(seed, H=16, W=32):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed, scale_out=0.3)
    _, bs, _, bo = U.split_params(p)
    for b in bs:
        b[:] = rng.normal(0, 0.1, b.shape)
    bo[:] = rng.normal(0, 0.1, bo.shape)
    img = U.bf16(rng.normal(size=(H, W, 8)).astype(np.float32))
    target = rng.normal(size=(H, W, 2)).astype(np.float32)
    return p, img, target


def _blocks():
    out, off = [], 0
    for l, (c1, _, c2, co, _) in enumerate(U.SPEC):
        k = co * 9 * (c1 + c2)
        out += [(f"W{l}", off, off + k), (f"b{l}", off + k, off + k + co)]
        off += k + co
    return out + [("Wo", off, off + 32), ("bo", off + 32, off + 34)]


def test_unet_forward_matches_oracle():
    from drd_application_amd.solver import unet_run
    print("unet forward 16x32", flush=True)
    p, img, _ = _net_case(1)
    out, _, _, _ = unet_run(p, U.bf16_bits(img))
    want, _ = U.forward(p, img, bf=True)
    assert rel_l2(out, want) < 1e-2


@pytest.mark.parametrize("H,W", [(16, 32), (32, 80)])
def test_unet_gradients_match_oracle(H, W):
    from drd_application_amd.solver import unet_run
    print(f"unet backward {H}x{W}", flush=True)
    p, img, t = _net_case(2, H, W)
    out, g, loss, _ = unet_run(p, U.bf16_bits(img), t)
    lw, ow, gw = U.loss_and_grads(p, img, t, bf=True)
    assert rel_l2(out, ow) < 1e-2
    assert abs(loss - lw) <= 1e-2 * lw
    for name, a, b in _blocks():
        assert rel_l2(g[a:b], gw[a:b]) < 3e-2, name
    assert rel_l2(g, gw) < 1e-2


def test_unet_adam_step_matches_oracle():
    from drd_application_amd.solver import unet_run
    print("unet adam", flush=True)
    p, img, t = _net_case(3)
    lr = 1e-3
    _, g, _, after = unet_run(p, U.bf16_bits(img), t, lr=lr)
    n = p.size
    want, _, _ = U.adam_step(p.astype(np.float64), g.astype(np.float64), np.zeros(n), np.zeros(n), 1, lr)
    # float32 arithmetic on parameters of size ~1: a few ulp of p plus a relative error of the step
    assert np.max(np.abs(after - want)) <= 2e-7 + 1e-3 * lr


N_C3 = 128


def _collect_and_train(g, seeds, iters=300, batch=8, lr=2e-3):
    for s in seeds:
        cfg, b = geometry.velocity_system(N_C3, s)
        geometry.load(g, cfg)
        x, added, its = g.learn_collect(b, tol=1e-12)
        assert added > 0 and its > 0
    return g.learn_train(iters, batch, lr)


def test_learned_correction_trains_and_converges_on_held_out_systems():
    from drd_application_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    cfg0, _ = geometry.velocity_system(N_C3, 3100)
    g = GpuSolver(cfg0.params)
    g.learn_init(capacity=256, crop_h=64, crop_w=64, seed=7)
    print("collect + train", flush=True)
    info = _collect_and_train(g, [3100 + i for i in range(8)])
    print("train", info, flush=True)
    assert info["pairs"] > 0 and info["train_iters"] == 300
    assert info["loss_last"] < 0.5 * info["loss_first"]
    results = []
    for s in (3990, 3991):
        cfg, b = geometry.velocity_system(N_C3, s)
        geometry.load(g, cfg)
        x0 = np.zeros_like(b)
        xp, itp, rp, ip = g.fgmres(abi.OP_VELO2, abi.PS_VELO_DCT32, b, x0, 5000, 1e-10)
        xl, itl, rl, il = g.fgmres(abi.OP_VELO2, abi.PS_VELO_LEARNED, b, x0, 5000, 1e-10)
        assert ip == 0 and il == 0
        o = Oracle(cfg.params)
        geometry.load(o, cfg)
        r = b - o.matvec(abi.OP_VELO2, 1.0, xl, 0.0, np.zeros_like(b))
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
        assert rel_l2(xl, xp) < 1e-6
        results.append((cfg.notes, itp, itl))
        print(f"held-out {cfg.notes}: plain {itp} its, learned {itl} its", flush=True)
    plain, learned = sum(r[1] for r in results), sum(r[2] for r in results)
    print(f"learned correction: {learned} vs {plain} FGMRES iterations ({100.0 * (plain - learned) / plain:+.1f}%)")
    # no gain to assert (see the module docstring): the correction stays
    # within 5% of the plain FGMRES iterations
    assert learned <= 1.05 * plain
    g.close()


def _learned_step_solver(bad_network=False):
    from drd_application_amd.solver import GpuSolver
    cfg = geometry.config("small")
    p = cfg.params.copy()
    p.velo_prec = abi.PREC_LEARNED
    p.ch_prec = abi.CH_PREC_BLOCK
    g = GpuSolver(p)
    geometry.load(g, cfg)
    g.learn_init(capacity=64, crop_h=16, crop_w=32, seed=3)
    g.learn_record(True)
    g.step(1, 0.0)
    for k in range(1, 4):
        g.step(2, k * p.dt)
    g.learn_record(False)
    info = g.learn_train(60, 4, 2e-3)
    assert info["pairs"] > 0
    if bad_network:  # a finite but useless correction: large random output layer
        q = g.learn_params()
        q[-34:] = np.random.default_rng(0).normal(0, 50.0, 34).astype(np.float32)
        g.learn_set_params(q)
    g.learn_enable(True)
    return g, p


def _converged(velo, tol):
    # info 0; RESID may exceed tol by the factor 1/|b| after the reference's
    # early exit, whose test is the absolute |s| <= tol (LIB/BiCGSTABREVCOM.f:311-316)
    return velo["info"] == 0 and velo["resid"] <= 10 * tol


def test_learned_step_is_judged_and_converges():
    print("learned step judge", flush=True)
    g, p = _learned_step_solver()
    s = g.step(2, 4 * p.dt)
    info = g.learn_info()
    assert info["state"] in (abi.LEARN_ON, abi.LEARN_FALLBACK)
    assert info["ref_iters"] > 0 and info["judged_iters"] > 0
    assert info["judged_ms"] > 0 and info["plain_ms"] > 0
    assert _converged(s["velo"], p.tol)
    s = g.step(2, 5 * p.dt)
    assert _converged(s["velo"], p.tol)
    g.close()


def test_bad_network_falls_back_to_the_plain_method():
    print("learned step fallback", flush=True)
    g, p = _learned_step_solver(bad_network=True)
    s = g.step(2, 4 * p.dt)
    info = g.learn_info()
    assert info["state"] == abi.LEARN_FALLBACK and info["fallbacks"] == 1
    assert _converged(s["velo"], p.tol)
    s = g.step(2, 5 * p.dt)  # stays on the plain preconditioner
    assert g.learn_info()["fallbacks"] == 1
    assert _converged(s["velo"], p.tol)
    g.close()


def test_bicgstab_rejects_the_learned_preconditioner():
    from drd_application_amd.solver import DrdError, GpuSolver
    cfg, b = geometry.velocity_system(64, 5)
    g = GpuSolver(cfg.params)
    geometry.load(g, cfg)
    with pytest.raises(DrdError):
        g.bicgstab(abi.OP_VELO2, abi.PS_VELO_LEARNED, b, np.zeros_like(b), 100, 1e-10)
    g.close()
