"""Benchmark: solver time steps/s of the DRD-Micro implicit-solve path on one MI355X.

One "step" is one full time step of the reference's driver sequence
(TP/mcl6.f90:395-502): Cahn-Hilliard phi/mu solve, BDF2 momentum solve,
pressure projection, every Krylov solve run to the configured relative
residual.  The timed region starts with every field resident in HBM.

Modes
  dropin : the plain FP64 path with the reference's own preconditioners and
           stopping rule (what a Fortran caller of the drop-in gets).

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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(cfg, velo_iters: float, phmu_iters: float, budget_s: float = 20.0):
    """Times the CPU oracle (serial FP64 restatement of the reference) on a
    bounded sample: k BiCGSTAB iterations of the momentum and of the CH solve
    on the same fields, scaled by the GPU run's iteration counts per step."""
    import numpy as np

    from drd_application_amd import abi, geometry
    from oracle.oracle import Oracle

    o = Oracle(cfg.params)
    geometry.load(o, cfg)
    p = cfg.params
    k = 4
    nv = abi.n_uv(p)
    bv = np.random.default_rng(1).standard_normal(nv)
    t0 = time.perf_counter()
    _, itv, _, _ = o.bicgstab(abi.OP_VELO2, abi.PS_IDENTITY, bv, np.zeros(nv), k, 1e-300)
    tv = (time.perf_counter() - t0) / max(itv, 1)
    nc = 2 * p.nx * p.ny
    bc = o.ph_mu_rhs()
    t0 = time.perf_counter()
    _, itc, _, _ = o.bicgstab(abi.OP_PHMU, abi.PS_PHMU_JACOBI, bc, np.zeros(nc), k, 1e-300)
    tc = (time.perf_counter() - t0) / max(itc, 1)
    o.close()
    step_s = velo_iters * tv + phmu_iters * tc
    return {
        "value": 1.0 / step_s if step_s > 0 else None,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (serial C, FP64) on the {cfg.name} fields: {itv} momentum and {itc} CH "
                   f"BiCGSTAB iterations timed ({tv*1e3:.1f} ms and {tc*1e3:.1f} ms each), scaled by "
                   f"the GPU run's {velo_iters:.1f} + {phmu_iters:.1f} iterations per step; RHS "
                   f"assembly and the pressure solve are not included"),
    }


MODES = {
    # name: (velocity preconditioner, CH preconditioner)
    "dropin": ("PREC_REFERENCE", "CH_PREC_REFERENCE"),
    "fp64dct": ("PREC_DCT64", "CH_PREC_BLOCK"),
    "fp32": ("PREC_DCT32", "CH_PREC_BLOCK"),
}


def run_mode(cfg, mode: str, steps: int, warmup: int, keep: bool = False, timer: bool = False):
    """Times `steps` BDF2 steps after `warmup` steps (the first is BDF1).
    With timer, every launch of the BDF2 velocity matvec (MV_DOT1) inside
    the timed steps is bracketed by HIP events on the library's stream."""
    from drd_application_amd import abi, geometry
    from drd_application_amd.solver import GpuSolver

    p = cfg.params.copy()
    vp, cp = MODES[mode]
    p.velo_prec = getattr(abi, vp)
    p.ch_prec = getattr(abi, cp)
    g = GpuSolver(p)
    geometry.load(g, cfg)
    g.synchronize()
    t = 0.0
    for s in range(max(1, warmup)):
        info = g.step(1 if s == 0 else 2, t)
        t += p.dt
        log(f"[{mode}] warmup step {s}: phmu {info['phmu']['iter']} ({info['phmu']['info']}), "
            f"velo {info['velo']['iter']} ({info['velo']['info']})")
    g.synchronize()
    if timer:
        g.kernel_timer(abi.K_VELO_MATVEC, 1 << 14)
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
           "velo_iters": sum(iters_v) / steps, "phmu_iters": sum(iters_c) / steps,
           "converged": all(a == 0 and b == 0 for a, b in infos)}
    if timer:
        n, ms, by = g.kernel_timer_read()
        g.kernel_timer(abi.K_VELO_MATVEC, 0)
        out["timed_launches"] = {"count": n, "total_ms": ms, "bytes_per_launch": by}
    if keep:
        return out, g
    g.close()
    return out, None


def measured_traffic(cfg_name: str):
    """HBM bytes per launch of the velocity matvec from the committed
    rocprofv3 counter summary for this workload, if there is one."""
    path = os.path.join(ROOT, "profiles", f"k_velo_matvec_{cfg_name}_counters.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def profile_kernels(cfg, mode: str, reps: int = 10) -> int:
    """The velocity matvec and a plain copy of the same length, each launched
    `reps` times after one BDF1 step has filled the fields; under
    rocprofv3 --pmc the copy's FETCH_SIZE/WRITE_SIZE calibrate the counters."""
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
    for which, name in ((abi.K_VELO_MATVEC, "velo_matvec"), (abi.K_COPY, "copy")):
        ms, by = g.bench_kernel(which, reps)
        out[name] = {"us": ms * 1e3, "bytes": by, "launches": reps + 1}
        log(f"{name}: {ms*1e3:.1f} us, {by/1e6:.1f} MB algorithmic")
    g.close()
    print(json.dumps(out))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--mode", default="fp32", choices=sorted(MODES))
    ap.add_argument("--compare", default="fp64dct,dropin",
                    help="comma-separated modes timed beside the headline mode ('' for none)")
    ap.add_argument("--compare-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the stand-alone kernel loops on the config's fields (for rocprofv3 --pmc)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solve path "
              f"is single-GPU, see DESIGN.md for how it would shard", file=sys.stderr)
        return 2
    from drd_application_amd import abi, geometry

    cfg = geometry.config(args.config)
    p = cfg.params
    log(f"config {cfg.name}: {p.nx}x{p.ny}, tol {p.tol:g}, dt {p.dt:g}")
    if args.profile_kernels:
        return profile_kernels(cfg, args.mode)
    head, g = run_mode(cfg, args.mode, args.steps, args.warmup, keep=True, timer=True)
    tl = head.pop("timed_launches")
    if tl["count"] == 0:
        log("the headline mode launched no BDF2 velocity matvec; no roofline")
        ms_k, bytes_k, achieved = None, tl["bytes_per_launch"], None
    else:
        ms_k = tl["total_ms"] / tl["count"]
        bytes_k = tl["bytes_per_launch"]
        achieved = bytes_k / (ms_k * 1e-3) / 1e9
        log(f"velocity matvec in the timed steps: {tl['count']} launches, {ms_k*1e3:.1f} us each, "
            f"{bytes_k/1e6:.1f} MB -> {achieved:.0f} GB/s")
    # stand-alone timings of other kernels on the final state (same launch shapes)
    kern = {}
    for which, name in ((abi.K_VELO_MATVEC, "velo_matvec"), (abi.K_PHMU_MATVEC, "phmu_matvec"),
                        (abi.K_BICG_XR, "bicg_xr"), (abi.K_PRESSURE, "pressure")):
        ms, by = g.bench_kernel(which, 20)
        kern[name] = {"us": ms * 1e3, "GB/s": by / (ms * 1e-3) / 1e9}
        log(f"{name} (stand-alone): {ms*1e3:.1f} us -> {kern[name]['GB/s']:.0f} GB/s")
    g.close()
    modes = {args.mode: head}
    for m in [s for s in args.compare.split(",") if s and s != args.mode]:
        modes[m], _ = run_mode(cfg, m, args.compare_steps, 1)
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
        "warmup": max(1, args.warmup),
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
                   "kernels": kern},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": measured_traffic(cfg.name),
                     "kernel": "k_velo_matvec<2,MV_DOT1>", "bytes_per_launch": bytes_k,
                     "us_per_launch": ms_k * 1e3 if ms_k else None,
                     "launches_timed": tl["count"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
