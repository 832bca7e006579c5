"""Benchmark: solver time steps/s of the DRD-Micro implicit-solve path on one MI355X.

One "step" is one full time step of the reference's driver sequence
(TP/mcl6.f90:395-502): Cahn-Hilliard phi/mu solve, BDF2 momentum solve,
pressure projection, every Krylov solve run to the configured relative
residual.  The timed region starts with every field resident in HBM.

Modes
  dropin  : the plain FP64 path with the reference's own preconditioners and
            stopping rule (what a Fortran caller of the drop-in gets).
  fp64dct : FP64 DCT/DST velocity preconditioner, block-Jacobi CH, BiCGSTAB,
            BDF2 solves started from the extrapolation 2x^n - x^{n-1} (all
            fast modes; the drop-in starts from x^n as the reference does).
  fp32    : the same with the velocity preconditioner in FP32.
  learned : fp32 plus the learned correction: the first BDF2 warmup steps
            record solver-emitted pairs, the network trains on them, one more
            warmup step judges it (learned FGMRES against plain FGMRES and
            the fp32 method) and the timed steps use what the judge kept.
  mg      : the BDF2 velocity solve preconditioned by one geometric
            multigrid V-cycle, the CH solve by sixteen red-black block SOR
            sweeps (weight 1.25), extrapolated start, FP64.
  mg32    : (the headline) the same with FP32 multigrid levels and FP32
            red-black sweeps.

The learned correction is also evaluated on its own workload (1024^2
variable-viscosity velocity systems, FGMRES to 1e-10 with and without it on
held-out systems) and its network's forward pass is timed against the BF16
MFMA peak.

Prints one JSON line (see the `result` dict below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0  # dense BF16 MFMA


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores():
    """(threads used, affinity count, os.cpu_count, CPU model) of this host.
    The threads used are the affinity count, capped by OMP_NUM_THREADS when
    the environment sets it (the share of the host the GPU's job owns)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    use = avail
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        use = min(use, int(env))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return use, avail, os.cpu_count(), model


def cpu_baseline(cfg, velo_iters: float, phmu_iters: float):
    """Times the CPU oracle (the FP64 C restatement of the reference's path,
    OpenMP loops whose results do not depend on the thread count) on bounded
    samples, at one thread (the reference is serial) and at the host's cores:

    * the workload itself: the step's right-hand-side assemblies and pressure
      projection once, plus k BiCGSTAB iterations of the momentum and of the
      CH solve with the reference's preconditioners, scaled by the GPU
      drop-in run's iteration counts per step;
    * whole steps, end to end, of c0 (256 x 128): a BDF1 and two BDF2 steps
      with the reference's velocity preconditioners and the block CH
      preconditioner (the reference's scalar Jacobi stops at its 5000
      iterations there without converging)."""
    import numpy as np

    from drd_application_amd import abi, geometry
    from oracle import oracle as O

    nthreads, avail, ncpu, model = host_cores()
    p = cfg.params
    k = 4
    nv = abi.n_uv(p)
    bv = np.random.default_rng(1).standard_normal(nv)
    nc = 2 * p.nx * p.ny
    cs = abi.cell_shape(p)
    res = {}
    for nt in sorted({1, nthreads}):
        O.set_threads(nt)
        o = O.Oracle(p)
        geometry.load(o, cfg)
        bc = o.ph_mu_rhs()
        o.bicgstab(abi.OP_VELO2, abi.PS_IDENTITY, bv, np.zeros(nv), 1, 1e-300)  # first touch, thread start
        o.bicgstab(abi.OP_PHMU, abi.PS_PHMU_JACOBI, bc, np.zeros(nc), 1, 1e-300)
        t0 = time.perf_counter()
        o.ph_mu_rhs()
        z = np.zeros(cs)
        o.velo_rhs(2, psi1=z, psi2=z, u_pre=o.get_field(abi.F_U), v_pre=o.get_field(abi.F_V),
                   ph_after=o.get_field(abi.F_PH), rmu_after=o.get_field(abi.F_RMU), t=p.dt)
        o.PreSolve_2(3 * p.re / 2.0 / p.dt)
        tfix = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, itv, _, _ = o.bicgstab(abi.OP_VELO2, abi.PS_IDENTITY, bv, np.zeros(nv), k, 1e-300)
        tv = (time.perf_counter() - t0) / max(itv, 1)
        t0 = time.perf_counter()
        _, itc, _, _ = o.bicgstab(abi.OP_PHMU, abi.PS_PHMU_JACOBI, bc, np.zeros(nc), k, 1e-300)
        tc = (time.perf_counter() - t0) / max(itc, 1)
        o.close()
        step_s = tfix + velo_iters * tv + phmu_iters * tc
        res[nt] = (1.0 / step_s if step_s > 0 else None, tv, tc, itv, itc, tfix)
    # c0 whole steps (the periodic workload at its own boundary)
    c0 = geometry.config("c0p")
    c0.params.ch_prec = abi.CH_PREC_BLOCK
    whole = {}
    for nt in sorted({1, nthreads}):
        O.set_threads(nt)
        o = O.Oracle(c0.params)
        geometry.load(o, c0)
        t, its = 0.0, []
        t0 = time.perf_counter()
        for s in range(3):
            info = o.step(1 if s == 0 else 2, t)
            its.append((info["phmu"]["iter"], info["velo"]["iter"]))
            t += c0.params.dt
        whole[nt] = {"steps_per_s": 3 / (time.perf_counter() - t0), "iters": its}
        o.close()
    O.set_threads(1)
    v1, vN = res[1], res[nthreads]
    return {
        "value": vN[0],
        "unit": "steps/s",
        "cores": nthreads,
        "kind": "port",
        "value_1thread": v1[0],
        "host": {"threads_used": nthreads, "affinity_cpus": avail, "os_cpus": ncpu, "model": model},
        "sample": (f"oracle (C, FP64, OpenMP) on the {cfg.name} fields at {nthreads} threads: the step's RHS "
                   f"assemblies and pressure projection ({vN[5]*1e3:.0f} ms; {v1[5]*1e3:.0f} ms at 1 thread) "
                   f"plus {vN[3]} momentum and {vN[4]} CH BiCGSTAB iterations with the reference's "
                   f"preconditioners ({vN[1]*1e3:.1f} and {vN[2]*1e3:.1f} ms each; at 1 thread "
                   f"{v1[1]*1e3:.1f} and {v1[2]*1e3:.1f} ms), scaled by the GPU drop-in run's "
                   f"{velo_iters:.1f} + {phmu_iters:.1f} iterations per step"),
        "c0_whole_steps": {
            "workload": "c0 256x128 periodic in x, BDF1 + 2 BDF2 steps end to end (RHS, CH, momentum, "
                        "pressure), reference velocity preconditioners, block CH preconditioner, tol 1e-10",
            "steps_per_s": whole[nthreads]["steps_per_s"], "steps_per_s_1thread": whole[1]["steps_per_s"],
            "iters": whole[nthreads]["iters"]},
    }


MODES = {
    # name: (velocity preconditioner, CH preconditioner)
    "dropin": ("PREC_REFERENCE", "CH_PREC_REFERENCE"),
    "fp64dct": ("PREC_DCT64", "CH_PREC_BLOCK"),
    "fp32": ("PREC_DCT32", "CH_PREC_BLOCK"),
    "learned": ("PREC_LEARNED", "CH_PREC_BLOCK"),
    "mg": ("PREC_MG", "CH_PREC_RBGS"),
    "mg32": ("PREC_MG32", "CH_PREC_RBGS32"),
}

GUESS = os.environ.get("DRD_BENCH_GUESS", "1") != "0"
# groups of launches of the roofline kernel timed with HIP events inside the
# timed steps (the first ones; each event pair adds a little to the steps
# it measures, so not every launch is timed)
TIMED_LAUNCHES = 160
LEARN = {"record_steps": 3, "capacity": 64, "crop": 256, "iters": 300, "batch": 4, "lr": 5e-4}


def learned_warmup(g, p, t):
    """BDF2 steps recording pairs, training, one judged step; returns (t, info)."""
    n = LEARN["crop"]
    g.learn_init(LEARN["capacity"], min(n, p.ny), min(n, p.nx), seed=1)
    g.learn_record(True)
    for _ in range(LEARN["record_steps"]):
        g.step(2, t)
        t += p.dt
    g.learn_record(False)
    info = g.learn_train(LEARN["iters"], LEARN["batch"], LEARN["lr"])
    log(f"[learned] {info['pairs']} pairs, {info['train_iters']} Adam steps in {info['train_ms']:.0f} ms, "
        f"loss {info['loss_first']:.3g} -> {info['loss_last']:.3g}")
    g.learn_enable(True)
    s = g.step(2, t)
    t += p.dt
    info = g.learn_info()
    log(f"[learned] judged step: learned FGMRES {info['judged_iters']} its {info['judged_ms']:.1f} ms, "
        f"plain FGMRES {info['ref_iters']} its, plain method {info['plain_ms']:.1f} ms -> state {info['state']}; "
        f"velo {s['velo']['iter']} ({s['velo']['info']})")
    return t, info


def run_mode(cfg, mode: str, steps: int, warmup: int, keep: bool = False, timer=None):
    """Times `steps` BDF2 steps after `warmup` steps (the first is BDF1), so
    every mode is timed on the same steps of the same run.  The learned mode
    spends the last record_steps + 1 warmup steps recording, training and
    judging.  With timer (an instrumented kernel id), every launch of that
    kernel inside the timed steps is bracketed by HIP events on the library's
    stream."""
    import numpy as np

    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuSolver

    p = cfg.params.copy()
    vp, cp = MODES[mode]
    p.velo_prec = getattr(abi, vp)
    p.ch_prec = getattr(abi, cp)
    g = GpuSolver(p)
    geometry.load(g, cfg)
    if mode != "dropin" and GUESS:  # the fast modes start each BDF2 solve from 2x^n - x^{n-1}
        g.set_initial_guess(1)
    if vp in ("PREC_DCT32", "PREC_DCT64", "PREC_LEARNED"):  # transpose tiles of the DCT preconditioner
        shape, ms = g.tune()
        log(f"[{mode}] transpose tiles: shape {shape} ({', '.join(f'{m*1e3:.1f}' for m in ms)} us per application)")
    g.synchronize()
    t = 0.0
    plain_warm = max(1, warmup - (LEARN["record_steps"] + 1 if mode == "learned" else 0))
    for s in range(plain_warm):
        info = g.step(1 if s == 0 else 2, t)
        t += p.dt
        log(f"[{mode}] warmup step {s}: phmu {info['phmu']['iter']} ({info['phmu']['info']}), "
            f"velo {info['velo']['iter']} ({info['velo']['info']})")
    learn = None
    if mode == "learned":
        t, learn = learned_warmup(g, p, t)
    g.synchronize()
    if timer is not None:
        g.kernel_timer(timer, TIMED_LAUNCHES)
    iters_v, iters_c, infos = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        info = g.step(2, t)
        t += p.dt
        iters_v.append(info["velo"]["iter"])
        iters_c.append(info["phmu"]["iter"])
        infos.append((info["phmu"]["info"], info["velo"]["info"]))
    g.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"[{mode}] {steps} steps in {elapsed:.3f} s ({steps/elapsed:.3f} steps/s); velo iters {iters_v}; "
        f"phmu iters {iters_c}; info {infos}")
    out = {"steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "warmup_steps": plain_warm + (LEARN["record_steps"] + 1 if learn is not None else 0),
           "velo_iters": sum(iters_v) / steps, "phmu_iters": sum(iters_c) / steps,
           "converged": all(a == 0 and b == 0 for a, b in infos)}
    if timer is not None:
        n, ms, by = g.kernel_timer_read()
        each = g.kernel_timer_launches()
        g.kernel_timer(timer, 0)
        # launches queued past the end of a solve return at once and move no
        # data: they take a small fraction of a working launch's time and are
        # counted apart, not averaged in
        work = each[each > 0.25 * np.median(each)] if each.size else each
        per = n // max(int(each.size), 1)  # launches per timed group (the mean of each is per launch)
        wsum = float(work.astype(np.float64).sum()) * per
        out["timed_launches"] = {"count": int(work.size) * per, "total_ms": wsum, "bytes_per_launch": by,
                                 "idle_launches": int(n - work.size * per), "idle_ms": float(ms - wsum)}
    if learn is not None:
        names = {abi.LEARN_ON: "learned FGMRES", abi.LEARN_FALLBACK: "plain fallback (fp32 BiCGSTAB)"}
        fin = g.learn_info()
        out["learned"] = {"pairs": learn["pairs"], "train_ms": learn["train_ms"],
                          "loss_first": learn["loss_first"], "loss_last": learn["loss_last"],
                          "judged": {"learned_fgmres_iters": learn["judged_iters"],
                                     "plain_fgmres_iters": learn["ref_iters"],
                                     "learned_fgmres_ms": learn["judged_ms"],
                                     "plain_method_ms": learn["plain_ms"]},
                          "timed_steps_used": names.get(learn["state"], str(learn["state"])),
                          "fallbacks": fin["fallbacks"]}
    if keep:
        return out, g
    g.close()
    return out, None


def learned_c3(n: int = 1024, pairs: int = 4000, heldout: int = 12, iters: int = 2000, batch: int = 8,
               crop: int = 256, crops_per_iterate: int = 16, lr: float = 5e-4):
    """The learned correction on its own workload: n x n variable-viscosity
    velocity systems (geometry.velocity_system).  Pairs are crop x crop crops
    of the iterates of recorded BiCGSTAB solves of training systems (16 per
    iterate) until `pairs` are held; held-out systems are solved to 1e-10 by
    FGMRES with and without the correction (which opens each restart cycle)
    and by the plain BiCGSTAB for reference.  Also times the n x n network's
    forward pass."""
    import numpy as np

    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import DrdError, GpuSolver

    cfg0, _ = geometry.velocity_system(n, 3000)
    g = GpuSolver(cfg0.params)
    g.learn_init_crops(pairs, crop, crop, crops_per_iterate, seed=11)
    t0 = time.perf_counter()
    systems, dropped = 0, 0
    while g.learn_info()["pairs"] < pairs:
        cfg, b = geometry.velocity_system(n, 3000 + systems)
        systems += 1
        geometry.load(g, cfg)
        try:
            g.learn_collect(b, tol=1e-11)
        except DrdError:  # not converged within maxit: its pairs are dropped
            dropped += 1
        if systems % 10 == 0:
            log(f"[c3] {systems} training systems, {g.learn_info()['pairs']} pairs")
    collect_s = time.perf_counter() - t0
    info = g.learn_train(iters, batch, lr)
    log(f"[c3] {info['pairs']} pairs from {systems} systems ({dropped} not converged, dropped; {collect_s:.1f} s "
        f"incl. generation), train {info['train_ms']:.0f} ms, loss {info['loss_first']:.3g} -> {info['loss_last']:.3g}")
    rows = []
    for s in range(heldout):
        cfg, b = geometry.velocity_system(n, 3900 + s)
        geometry.load(g, cfg)
        x0 = np.zeros_like(b)
        g.synchronize()
        r = {"system": cfg.notes}
        for key, fn in (("fgmres_plain", lambda: g.fgmres(abi.OP_VELO2, abi.PS_VELO_DCT32, b, x0, 8000, 1e-10)),
                        ("fgmres_learned", lambda: g.fgmres(abi.OP_VELO2, abi.PS_VELO_LEARNED, b, x0, 8000, 1e-10)),
                        ("bicgstab_plain", lambda: g.bicgstab(abi.OP_VELO2, abi.PS_VELO_DCT32, b, x0, 8000, 1e-10))):
            t1 = time.perf_counter()
            _, its, _, inf = fn()
            r[key] = {"iters": its, "ms": 1e3 * (time.perf_counter() - t1), "info": inf}
        rows.append(r)
        if s < 12 or s % 20 == 0:
            log(f"[c3] held-out {s} {cfg.notes}: FGMRES plain {r['fgmres_plain']['iters']} its "
                f"{r['fgmres_plain']['ms']:.0f} ms, learned {r['fgmres_learned']['iters']} its "
                f"{r['fgmres_learned']['ms']:.0f} ms; BiCGSTAB {r['bicgstab_plain']['iters']} its "
                f"{r['bicgstab_plain']['ms']:.0f} ms")
    ms, flops = g.bench_kernel(abi.K_UNET_FWD, 10)
    g.close()
    keys = ("fgmres_plain", "fgmres_learned", "bicgstab_plain")
    tot = {k: sum(r[k]["iters"] for r in rows) for k in keys}
    tms = {k: sum(r[k]["ms"] for r in rows) for k in keys}
    better = sum(1 for r in rows if r["fgmres_learned"]["iters"] < r["fgmres_plain"]["iters"])
    tflops = flops / (ms * 1e-3) / 1e12
    log(f"[c3] held-out totals: FGMRES plain {tot['fgmres_plain']} its {tms['fgmres_plain']:.0f} ms, learned "
        f"{tot['fgmres_learned']} its {tms['fgmres_learned']:.0f} ms (fewer iterations on {better} of {heldout}); "
        f"BiCGSTAB {tot['bicgstab_plain']} its {tms['bicgstab_plain']:.0f} ms")
    log(f"[c3] network forward {n}x{n}: {ms*1e3:.0f} us, {flops/1e9:.1f} GFLOP -> {tflops:.1f} TFLOP/s")
    return {
        "workload": f"{n}x{n} velocity systems (I - dt div(eta grad) + dt chi R) u = f, eta contrast "
                    f"log-uniform to 1e3, R = 1e6 discs, FGMRES(30) to 1e-10",
        "train_systems": systems, "train_systems_dropped": dropped, "pairs": info["pairs"],
        "pair": f"{crop}x{crop} crop, {crops_per_iterate} per recorded iterate",
        "adam_steps": info["train_iters"], "batch": batch, "train_ms": info["train_ms"],
        "loss_first": info["loss_first"], "loss_last": info["loss_last"],
        "heldout_systems": heldout, "iters": tot, "ms": tms, "heldout_fewer_iterations": better,
        "iteration_gain": (tot["fgmres_plain"] - tot["fgmres_learned"]) / tot["fgmres_plain"],
        "time_gain": (tms["fgmres_plain"] - tms["fgmres_learned"]) / tms["fgmres_plain"],
        "time_gain_vs_bicgstab": (tms["bicgstab_plain"] - tms["fgmres_learned"]) / tms["bicgstab_plain"],
        "per_system": rows,
        "roofline_network": {"bound": "mfma", "achieved": tflops, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                             "frac": tflops / PEAK_BF16_TFLOPS, "flops_per_launch": flops,
                             "us_per_launch": ms * 1e3},
    }


def particle_coupling(cfg_name: str = "c1", nbody: int = 400, reps: int = 10, cpu: bool = True):
    """The particle coupling reductions of one step (force_caculation_vel and
    update_position_vel, TP/particles_move.f90:758-920): `nbody` bodies spread
    over and averaged against the grid (the reference evaluates every body at
    every cell, O(nbody x cells)).  Each call includes its upload of the bodies and, for the
    velocity averages, the read-back.  The device evaluates each body only
    on the cells within its tanh cutoff; the CPU figure is the oracle (every
    body at every cell) on two bodies, one thread, scaled to `nbody`."""
    import numpy as np
    from drd_application_amd import geometry
    from drd_application_amd.solver import GpuSolver
    cfg = geometry.config(cfg_name)
    p = cfg.params
    lx, ly = p.nx * p.dx, p.ny * p.dz
    rng = np.random.default_rng(400)
    bodies = np.stack([rng.uniform(0, lx, nbody), rng.uniform(0, ly, nbody), np.full(nbody, 0.02 * ly),
                       rng.normal(size=nbody), rng.normal(size=nbody)], axis=1)
    g = GpuSolver(p)
    geometry.load(g, cfg)
    out = {"workload": f"particle force spread + velocity averages, {nbody} bodies on {p.nx}x{p.ny}",
           "bodies": nbody}
    try:
        for name, call in (("force_ms", lambda: g.particle_force(bodies, lx, ly)),
                           ("velocity_ms", lambda: g.particle_velocity(bodies, lx, ly))):
            call()
            call()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            out[name] = (time.perf_counter() - t0) / reps * 1e3
    finally:
        g.close()
    if cpu:
        from oracle.oracle import Oracle
        o = Oracle(p)
        geometry.load(o, cfg)
        t0 = time.perf_counter()
        o.particle_force(bodies[:2], lx, ly)
        dt = time.perf_counter() - t0
        o.close()
        out["cpu_force_ms"] = dt / 2 * nbody * 1e3
        out["cpu_sample"] = ("oracle, 2 bodies, 1 thread, scaled to the body count; the oracle evaluates every "
                             "body at every cell (no cutoff), the device skips cells beyond a body's tanh cutoff, "
                             "so the ratio mixes that pruning with the hardware")
    return out


def periodic_c0(steps: int = 20, warmup: int = 2):
    """The periodic workload (256 x 128, periodic in x, 6 droplets, 3 discs,
    eta ratio 10, R 1e4, dt 1e-4, tol 1e-10): BDF2 steps on the GPU with the
    reference's velocity preconditioners (Jacobi BDF1, identity BDF2) and the
    phi/mu block preconditioner, the ones periodic x takes."""
    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuSolver

    cfg = geometry.config("c0p")
    p = cfg.params.copy()
    p.ch_prec = abi.CH_PREC_BLOCK
    g = GpuSolver(p)
    geometry.load(g, cfg)
    t = 0.0
    for s in range(warmup):
        g.step(1 if s == 0 else 2, t)
        t += p.dt
    g.synchronize()
    its, infos = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        info = g.step(2, t)
        t += p.dt
        its.append((info["phmu"]["iter"], info["velo"]["iter"]))
        infos.append((info["phmu"]["info"], info["velo"]["info"]))
    g.synchronize()
    el = time.perf_counter() - t0
    g.close()
    log(f"[c0p] {steps} steps in {el:.3f} s ({steps/el:.1f} steps/s); iters {its[:4]} ...")
    return {"workload": "c0 256x128 periodic in x, reference velocity preconditioners, block CH, tol 1e-10",
            "steps": steps, "warmup": warmup, "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps,
            "phmu_iters": sum(a for a, _ in its) / steps, "velo_iters": sum(b for _, b in its) / steps,
            "converged": all(a == 0 and b == 0 for a, b in infos)}


def combined_c4(steps: int = 20, warmup: int = 2):
    """The full combined workload (geometry.c4_config): 4096 x 2048 serpentine
    channel, 40 droplets, 400 particles and 64 fibers of 20 beads as
    diffuse-resistance domains, headline preconditioners, dt 2e-5.  One
    coupled step = the bodies' forces spread on the device
    (drd_particle_force), the time step, and the bodies moved on the host
    with their device-averaged velocities (drd_particle_velocity)."""
    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuSolver

    t0 = time.perf_counter()
    cfg, bodies = geometry.c4_config()
    gen_s = time.perf_counter() - t0
    q = cfg.params.copy()
    vp, cp = MODES["mg32"]
    q.velo_prec, q.ch_prec = getattr(abi, vp), getattr(abi, cp)
    lx, ly = q.nx * q.dx, q.ny * q.dz
    g = GpuSolver(q)
    try:
        geometry.load(g, cfg)
        g.set_initial_guess(1)
        t = 0.0
        for s in range(warmup):
            bodies.spread(g, lx, ly)
            info = g.step(1 if s == 0 else 2, t)
            bodies.advance(g, q.dt, lx, ly)
            t += q.dt
        g.synchronize()
        its, ok = [], True
        t0 = time.perf_counter()
        for _ in range(steps):
            bodies.spread(g, lx, ly)
            info = g.step(2, t)
            bodies.advance(g, q.dt, lx, ly)
            t += q.dt
            its.append((info["phmu"]["iter"], info["velo"]["iter"]))
            ok = ok and info["phmu"]["info"] == 0 and info["velo"]["info"] == 0
        g.synchronize()
        el = time.perf_counter() - t0
    finally:
        g.close()
    log(f"[c4] {cfg.notes}: {steps} coupled steps in {el:.3f} s ({steps/el:.2f} steps/s); CH/velocity its {its}")
    return {"workload": cfg.notes + ", mg32 preconditioners, dt 2e-5, tol 1e-10",
            "bodies": bodies.count, "generated_s": gen_s, "warmup": warmup, "steps": steps,
            "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "converged": ok,
            "mean_iters": {"phmu": sum(a for a, _ in its) / steps, "velo": sum(b for _, b in its) / steps}}


def batched_c2(count: int = 256, steps: int = 3, warmup: int = 2, sample: int = 4, graphs: bool = True,
               velo: str = "PREC_MG32", sequential: bool = True, pow2: bool = False):
    """The batched workload: `count` independent instances of uneven sizes
    (geometry.batch_configs) stepped together by drd_batch_step, every size
    group's kernels launched once over all its instances, BiCGSTAB iteration
    chunks replayed from captured hipGraphs, transpose tiles autotuned.  The
    sequential figure steps up to `sample` instances of each size alone (each
    its own context, no graphs) and scales their mean by the number of
    instances of that size.  The CH preconditioner is the FP32 red-black one
    (the configs' own choice is block Jacobi); the velocity preconditioner
    `velo` (default the FP32 multigrid; PREC_DCT32 the FP32 fast-Poisson).
    nx is drawn from the multiples of 32 in [128, 1024] (geometry.C2_SIZES),
    or with pow2 from the powers of two (the fast-Poisson preconditioner's
    FFTs need them)."""
    import numpy as np

    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuBatch, GpuSolver

    t0 = time.perf_counter()
    cfgs = geometry.batch_configs(count, sizes=(128, 256, 512, 1024) if pow2 else geometry.C2_SIZES)
    for c in cfgs:
        c.params.ch_prec = abi.CH_PREC_RBGS32
        c.params.velo_prec = getattr(abi, velo)
    gen_s = time.perf_counter() - t0
    cells = sum(c.params.nx * c.params.ny for c in cfgs)
    dt = cfgs[0].params.dt
    b = GpuBatch([c.params for c in cfgs])

    def load():
        for i, c in enumerate(cfgs):
            for fid, a in c.fields.items():
                b.set_field(i, fid, a)

    load()
    b.set_initial_guess(1)
    tuned = b.tune()
    log(f"[c2] {velo}: {count} instances, {cells/1e6:.1f} M cells, groups {b.groups()} (generated in {gen_s:.1f} s); "
        f"tiles {[(t[0], t[2]) for t in tuned]}")
    t = 0.0

    def warm():
        nonlocal t
        t = 0.0
        for k in range(warmup):
            b.step(1 if k == 0 else 2, t)
            t += dt

    warm()

    def timed(n):
        nonlocal t
        its_v, its_c, bad = 0, 0, 0
        t1 = time.perf_counter()
        for _ in range(n):
            infos = b.step(2, t)
            t += dt
            its_v += sum(i["velo"]["iter"] for i in infos)
            its_c += sum(i["phmu"]["iter"] for i in infos)
            bad += sum(1 for i in infos if i["velo"]["info"] != 0 or i["phmu"]["info"] != 0)
        return time.perf_counter() - t1, its_v / n / count, its_c / n / count, bad
    el, iv, ic, bad = timed(steps)
    # the same warm-up and timed steps again from the same fields, with the
    # iteration chunks replayed from captured graphs (later steps need fewer
    # iterations, so only the same steps compare); skipped under a profiler,
    # whose kernel tracer has ended graph replays with a host SIGSEGV
    el_g = None
    if graphs:
        load()
        b.use_graphs(True)
        warm()
        el_g, iv_g, ic_g, _ = timed(steps)
        el_g /= steps
        if (iv_g, ic_g) != (iv, ic):
            log(f"[c2] the runs with and without graphs iterated differently: {iv}/{ic} and {iv_g}/{ic_g}")
    b.close()
    log(f"[c2] batched: {steps} steps in {el:.3f} s ({el/steps*1e3:.0f} ms/step, {count*steps/el:.0f} "
        f"instance-steps/s); with graphs {el_g * 1e3 if el_g else float('nan'):.0f} ms/step; mean its velo "
        f"{iv:.1f}, CH {ic:.1f}; non-converged solves {bad}")
    # sequential estimate: one instance per size, its own context
    seq = 0.0
    per_size = {}
    for nx in sorted({c.params.nx for c in cfgs}) if sequential else []:
        group = [c for c in cfgs if c.params.nx == nx]
        times = []
        for c in group[:sample]:
            g = GpuSolver(c.params)
            geometry.load(g, c)
            g.set_initial_guess(1)
            tt = 0.0
            for k in range(warmup):
                g.step(1 if k == 0 else 2, tt)
                tt += dt
            g.synchronize()
            t1 = time.perf_counter()
            for _ in range(steps):
                g.step(2, tt)
                tt += dt
            g.synchronize()
            times.append((time.perf_counter() - t1) / steps * 1e3)
            g.close()
        ms = float(np.mean(times))
        per_size[f"{nx}x{nx // 2}"] = {"instances": len(group), "sampled": len(times), "ms_per_step_alone": ms}
        seq += ms * len(group)
    log(f"[c2] sequential estimate {seq:.0f} ms/step ({per_size})")
    return {
        "workload": f"{count} instances, nx log-uniform over {'powers of two' if pow2 else 'multiples of 32'} "
                    f"128..1024, ny = nx/2, eta ratio "
                    f"U[1,100], 0-8 discs R log-uniform 1e4..1e6, 2-12 droplets, "
                    f"{'FP32 multigrid' if velo == 'PREC_MG32' else velo} velocity and FP32 red-black CH "
                    f"preconditioners, FP64 BiCGSTAB, tol 1e-10",
        "cells": cells, "groups": [list(gr) for gr in tuned],
        "ms_per_batch_step": el / steps * 1e3, "instance_steps_per_s": count * steps / el,
        "cell_steps_per_s": cells * steps / el, "ms_per_batch_step_graphs": el_g * 1e3 if el_g else None,
        "sequential_ms_per_step_estimate": seq if sequential else None,
        "speedup_vs_sequential": seq / (el / steps * 1e3) if sequential else None,
        "mean_iters": {"velo": iv, "phmu": ic}, "non_converged_solves": bad,
        "per_size_alone": per_size,
    }


# the kernel the roofline reports per mode: the one that takes the most time
# in its steps (kernel id, name, prefix of its counter summary in profiles/)
ROOF_SWEEP = ("K_PHMU_SWEEP32", "k_phmu_rb_sweep<FP32 middle launch, 2 sweeps>", "k_phmu_rb_sweep32")
ROOF_MATVEC = ("K_VELO_MATVEC", "k_velo_matvec<2,MV_DOT1>", "k_velo_matvec")


def roof_kernel(mode: str):
    return ROOF_SWEEP if MODES[mode][1] == "CH_PREC_RBGS32" else ROOF_MATVEC


def measured_traffic(prefix: str, cfg_name: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary for this workload, if there is one."""
    path = os.path.join(ROOT, "profiles", f"{prefix}_{cfg_name}_counters.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def profile_kernels(cfg, mode: str, reps: int = 10) -> int:
    """The velocity and CH matvecs, the BiCGSTAB x/r update, a plain copy of
    the velocity vector's length, the mode's velocity preconditioner and
    (FP32 red-black modes) one middle CH sweep,
    each launched `reps` times after one BDF1 step has filled the fields;
    under rocprofv3 --pmc the copy's FETCH_SIZE/WRITE_SIZE calibrate the
    counters."""
    from drd_application_amd import abi
    from drd_application_amd.solver import GpuSolver
    from drd_application_amd import geometry

    p = cfg.params.copy()
    vp, cp = MODES[mode]
    p.velo_prec, p.ch_prec = getattr(abi, vp), getattr(abi, cp)
    g = GpuSolver(p)
    geometry.load(g, cfg)
    g.step(1, 0.0)
    out = {}
    prec_k = (abi.K_VELO_MG, "velo_mg_vcycle") if mode.startswith("mg") else (abi.K_VELO_PREC, "velo_prec")
    sweep = [(abi.K_PHMU_SWEEP32, "phmu_rb_sweep32")] if MODES[mode][1] == "CH_PREC_RBGS32" else []
    for which, name in [(abi.K_VELO_MATVEC, "velo_matvec"), (abi.K_PHMU_MATVEC, "phmu_matvec"),
                        (abi.K_BICG_XR, "bicg_xr"), (abi.K_COPY, "copy"), prec_k] + sweep:
        ms, by = g.bench_kernel(which, reps)
        out[name] = {"us": ms * 1e3, "bytes": by, "launches": reps + 1}
        log(f"{name}: {ms*1e3:.1f} us, {by/1e6:.1f} MB algorithmic")
    g.close()
    print(json.dumps(out))
    return 0


def profile_network(n: int = 1024, reps: int = 10) -> int:
    """The learned correction's n x n network forward pass launched `reps`
    times (for rocprofv3 --kernel-trace; the parameters are the initial ones)."""
    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuSolver

    cfg, _ = geometry.velocity_system(n, 3000)
    g = GpuSolver(cfg.params)
    g.learn_init(4, 256, 256, seed=1)
    ms, flops = g.bench_kernel(abi.K_UNET_FWD, reps)
    log(f"network forward {n}x{n}: {ms*1e3:.0f} us, {flops/1e9:.1f} GFLOP -> {flops/(ms*1e-3)/1e12:.1f} TFLOP/s")
    g.close()
    print(json.dumps({"us": ms * 1e3, "flops": flops, "launches": reps + 1}))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps (BDF1 first); the learned mode records, trains and judges in the last "
                         "4 of them")
    ap.add_argument("--config", default="c1")
    ap.add_argument("--mode", default="mg32", choices=sorted(MODES))
    ap.add_argument("--compare", default="mg,learned,fp32,fp64dct,dropin",
                    help="comma-separated modes timed beside the headline mode ('' for none)")
    ap.add_argument("--compare-steps", type=int, default=None,
                    help="timed steps of the comparison modes (default: --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the learned correction's own workload")
    ap.add_argument("--no-c2", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-particles", action="store_true", help="skip the particle coupling reductions")
    ap.add_argument("--no-c4", action="store_true", help="skip the full combined 4096x2048 workload")
    ap.add_argument("--no-c0", action="store_true", help="skip the periodic 256x128 workload")
    ap.add_argument("--no-graphs", action="store_true",
                    help="skip the batched workload's run with captured graphs (for profiled runs)")
    ap.add_argument("--c3-full", action="store_true",
                    help="only the learned correction's workload at its full size: 20k training pairs, "
                         "200 held-out systems (several minutes)")
    ap.add_argument("--profile-network", action="store_true",
                    help="only launch the learned correction's 1024^2 network forward pass")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the stand-alone kernel loops on the config's fields (for rocprofv3 --pmc)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solve path "
              f"is single-GPU, see DESIGN.md for how it would shard", file=sys.stderr)
        return 2
    import numpy as np

    from drd_application_amd import abi, geometry

    cfg = geometry.config(args.config)
    p = cfg.params
    log(f"config {cfg.name}: {p.nx}x{p.ny}, tol {p.tol:g}, dt {p.dt:g}")
    if args.profile_kernels:
        return profile_kernels(cfg, args.mode)
    if args.profile_network:
        return profile_network()
    if args.c3_full:
        print(json.dumps({"learned_c3": learned_c3(pairs=20000, heldout=200, iters=6000)}))
        return 0
    rk_id, rk_name, rk_prefix = roof_kernel(args.mode)
    rk = getattr(abi, rk_id)
    head, g = run_mode(cfg, args.mode, args.steps, args.warmup, keep=True, timer=rk)
    tl = head.pop("timed_launches")
    source = "in-solve HIP events"
    if tl["count"] == 0:  # FGMRES launches the plain matvec: time it stand-alone
        ms_k, bytes_k = g.bench_kernel(rk, 20)
        achieved = bytes_k / (ms_k * 1e-3) / 1e9
        source = "stand-alone loop (the timed steps launched no instrumented kernel)"
        log(f"{rk_name} stand-alone: {ms_k*1e3:.1f} us -> {achieved:.0f} GB/s")
    else:
        ms_k = tl["total_ms"] / tl["count"]
        bytes_k = tl["bytes_per_launch"]
        achieved = bytes_k / (ms_k * 1e-3) / 1e9
        log(f"{rk_name} in the timed steps: {tl['count']} launches, {ms_k*1e3:.1f} us each, "
            f"{bytes_k/1e6:.1f} MB -> {achieved:.0f} GB/s ({tl['idle_launches']} launches after "
            f"convergence, {tl['idle_ms']*1e3:.0f} us in all, not counted)")
    # stand-alone timings of other kernels on the final state (same launch shapes)
    kern = {}
    prec_k = (abi.K_VELO_MG, "velo_mg_vcycle") if args.mode.startswith("mg") else (abi.K_VELO_PREC, "velo_prec")
    sweep = [(abi.K_PHMU_SWEEP32, "phmu_rb_sweep32")] if rk_prefix == "k_phmu_rb_sweep32" else []
    for which, name in [(abi.K_VELO_MATVEC, "velo_matvec"), (abi.K_PHMU_MATVEC, "phmu_matvec"),
                        (abi.K_BICG_XR, "bicg_xr"), (abi.K_PRESSURE, "pressure"), prec_k] + sweep:
        ms, by = g.bench_kernel(which, 20)
        kern[name] = {"us": ms * 1e3, "GB/s": by / (ms * 1e-3) / 1e9}
        log(f"{name} (stand-alone): {ms*1e3:.1f} us -> {kern[name]['GB/s']:.0f} GB/s")
    modes = {args.mode: head}
    field_err = None
    for m in [s for s in args.compare.split(",") if s and s != args.mode]:
        same_steps = (args.compare_steps or args.steps) == args.steps
        modes[m], g64 = run_mode(cfg, m, args.compare_steps or args.steps, args.warmup,
                                 keep=(m == "fp64dct" and same_steps))
        if g64 is not None:  # the headline run against the FP64-preconditioned run, same steps
            field_err = {}
            for fid in (abi.F_U, abi.F_V, abi.F_P, abi.F_PH, abi.F_RMU):
                a, b = g.get_field(fid), g64.get_field(fid)
                field_err[abi.FIELD_NAMES[fid]] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
            g64.close()
            log(f"relative L2 difference of the fields, {args.mode} vs fp64dct after {args.warmup} + "
                f"{args.steps} steps: {field_err}")
    g.close()
    c2 = None
    if not args.no_c2:
        try:
            c2 = batched_c2(graphs=not args.no_graphs)
            # the FP32 fast-Poisson and the multigrid velocity preconditioners on a
            # batch of power-of-two sizes (the fast-Poisson FFTs need them), for comparison
            for key, velo in (("pow2_fp32_fast_poisson", "PREC_DCT32"), ("pow2_fp32_multigrid", "PREC_MG32")):
                r = batched_c2(graphs=False, velo=velo, sequential=False, pow2=True)
                c2[key] = {k: r[k] for k in ("cells", "groups", "ms_per_batch_step", "instance_steps_per_s",
                                             "mean_iters", "non_converged_solves")}
        except Exception as e:  # reported beside the metric, never required for it
            log(f"batched workload failed: {e}")
    pc = None
    if not args.no_particles:
        try:
            pc = particle_coupling(cfg.name, cpu=not args.no_cpu)
        except Exception as e:  # reported beside the metric, never required for it
            log(f"particle coupling failed: {e}")
    c4 = None
    if not args.no_c4:
        try:
            c4 = combined_c4()
        except Exception as e:  # reported beside the metric, never required for it
            log(f"combined workload failed: {e}")
    c0p = None
    if not args.no_c0:
        try:
            c0p = periodic_c0()
        except Exception as e:  # reported beside the metric, never required for it
            log(f"periodic workload failed: {e}")
    c3 = None
    if not args.no_c3:
        try:
            c3 = learned_c3()
        except Exception as e:  # reported beside the metric, never required for it
            log(f"learned-correction workload failed: {e}")
    cpu = None
    if not args.no_cpu:
        ref = modes.get("dropin", head)
        try:
            cpu = cpu_baseline(cfg, ref["velo_iters"], ref["phmu_iters"])
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
    result = {
        "metric": "solver steps/s at 1e-10 rel-res, learned+FP32 vs plain FP64 precond; HBM GB/s",
        "value": head["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": head["warmup_steps"],
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name} {p.nx}x{p.ny} two-phase CH+NS, tol {p.tol:g}",
                   "mode": args.mode, "modes": modes,
                   "speedup_vs_fp64dct": (head["steps_per_s"] / modes["fp64dct"]["steps_per_s"]
                                          if "fp64dct" in modes and args.mode != "fp64dct" else None),
                   "speedup_vs_dropin": (head["steps_per_s"] / modes["dropin"]["steps_per_s"]
                                         if "dropin" in modes and args.mode != "dropin" else None),
                   # the metric's own ratio: learned + FP32-preconditioned against the plain FP64 DCT
                   "speedup_learned_vs_fp64dct": (modes["learned"]["steps_per_s"] / modes["fp64dct"]["steps_per_s"]
                                                  if "learned" in modes and "fp64dct" in modes else None),
                   "field_rel_l2_vs_fp64dct": field_err,
                   "kernels": kern},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": measured_traffic(rk_prefix, cfg.name),
                     "kernel": rk_name, "bytes_per_launch": bytes_k,
                     "us_per_launch": ms_k * 1e3 if ms_k else None,
                     "launches_timed": tl["count"], "idle_launches": tl.get("idle_launches", 0),
                     "source": source},
        "cpu_baseline": cpu,
        "learned_c3": c3,
        "batched_c2": c2,
        "combined_c4": c4,
        "periodic_c0": c0p,
        "particle_coupling": pc,
    }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
