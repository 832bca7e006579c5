"""Batched stepping (drd_batch) and captured iteration chunks on the GPU.

Every instance of a batch runs the same kernels as a single-instance context,
with its own state machine and its own canonical reductions, so each
instance's fields and Krylov reports must be bit for bit those of the same
instance stepped alone (drd_init/drd_step).  Batches replay captured hipGraph
chunks when asked (drd_batch_use_graphs); the single-instance runs they are compared with do not,
and a single context with graphs on is compared step by step with one
without (every replay after eager work and synchronisation).
"""
import numpy as np
import pytest

from drd_application_amd import abi, geometry

pytestmark = pytest.mark.gpu

FIELDS = (abi.F_U, abi.F_V, abi.F_P, abi.F_PH, abi.F_RMU)


def _configs(n=5, seed=77, sizes=(64, 128)):
    return geometry.batch_configs(n, seed=seed, sizes=sizes)


def _load_batch(b, cfgs):
    for i, cfg in enumerate(cfgs):
        for fid, a in cfg.fields.items():
            b.set_field(i, fid, a)


def test_batch_groups_and_field_round_trip():
    from drd_application_amd.solver import GpuBatch
    cfgs = _configs()
    with GpuBatch([c.params for c in cfgs]) as b:
        sizes = sorted((c.params.nx, c.params.ny) for c in cfgs)
        groups = b.groups()
        assert sum(g[2] for g in groups) == len(cfgs)
        assert sorted(set(sizes)) == sorted((g[0], g[1]) for g in groups)
        _load_batch(b, cfgs)
        for i, cfg in enumerate(cfgs):
            for fid in (abi.F_ETA, abi.F_BDR_PH_PENALTY, abi.F_U):
                np.testing.assert_array_equal(b.get_field(i, fid), cfg.fields[fid])


@pytest.mark.parametrize("velo,graphs,sizes", [("PREC_MG32", False, (128, 256)), ("PREC_MG32", True, (128, 256)),
                                               ("PREC_MG", False, (128, 256)), ("PREC_MG32", False, (160, 224, 288))])
def test_batch_multigrid_steps_are_bitwise_the_single_instance_steps(velo, graphs, sizes):
    """Batch groups with the multigrid velocity preconditioner (each
    instance's BDF1 and BDF2 hierarchies inside its block of the group's
    memory, every level's launches over all instances) and the FP32
    red-black CH preconditioner: three steps of every instance equal the same
    instance stepped alone, fields and Krylov reports bit for bit.  The size
    groups step at once from host threads of their own; sizes that are
    multiples of 32 but not powers of two take the dense-transform coarsest
    solve."""
    from drd_application_amd.solver import GpuBatch, GpuSolver
    cfgs = _configs(n=6, seed=5, sizes=sizes)
    assert len({c.params.nx for c in cfgs}) >= 2
    for c in cfgs:
        c.params.velo_prec = getattr(abi, velo)
        c.params.ch_prec = abi.CH_PREC_RBGS32
    print("multigrid batch of", [c.notes for c in cfgs], velo, "graphs", graphs, flush=True)
    steps = 3
    with GpuBatch([c.params for c in cfgs]) as b:
        _load_batch(b, cfgs)
        b.use_graphs(graphs)
        binfo = [b.step(1 if k == 0 else 2, k * cfgs[0].params.dt) for k in range(steps)]
        for i, cfg in enumerate(cfgs):
            g = GpuSolver(cfg.params)
            geometry.load(g, cfg)
            for k in range(steps):
                s = g.step(1 if k == 0 else 2, k * cfg.params.dt)
                for key in ("phmu", "velo"):
                    assert s[key]["info"] == 0 and binfo[k][i][key]["info"] == 0, (i, k, key)
                    assert s[key]["iter"] == binfo[k][i][key]["iter"], (i, k, key)
                    assert s[key]["resid"] == binfo[k][i][key]["resid"]
            for fid in FIELDS:
                np.testing.assert_array_equal(b.get_field(i, fid), g.get_field(fid), err_msg=f"instance {i} field {fid}")
            g.close()
        print("velocity iterations per step", [[x["velo"]["iter"] for x in bi] for bi in binfo], flush=True)


@pytest.mark.parametrize("steps,ch,graphs", [(4, None, True), (3, "CH_PREC_RBGS32", False),
                                            (3, "CH_PREC_RBGS", True)])
def test_batch_steps_are_bitwise_the_single_instance_steps(steps, ch, graphs):
    """ch: the CH preconditioner for every instance (None: the configs' own);
    the red-black forms keep their iterates in a per-instance scratch.
    graphs: the batch replays captured iteration chunks (opt-in)."""
    from drd_application_amd.solver import GpuBatch, GpuSolver
    cfgs = _configs()
    if ch is not None:
        for c in cfgs:
            c.params.ch_prec = getattr(abi, ch)
    print("batch of", [c.notes for c in cfgs], "ch", ch, flush=True)
    with GpuBatch([c.params for c in cfgs]) as b:
        _load_batch(b, cfgs)
        b.use_graphs(graphs)
        tuned = b.tune()
        print("tuned", tuned, flush=True)
        assert all(0 <= t[2] < 3 and all(m > 0 for m in t[3]) for t in tuned)
        binfo = []
        for k in range(steps):
            binfo.append(b.step(1 if k == 0 else 2, k * cfgs[0].params.dt))
        for i, cfg in enumerate(cfgs):
            g = GpuSolver(cfg.params)
            geometry.load(g, cfg)
            for k in range(steps):
                s = g.step(1 if k == 0 else 2, k * cfg.params.dt)
                for key in ("phmu", "velo"):
                    assert s[key]["iter"] == binfo[k][i][key]["iter"], (i, k, key)
                    assert s[key]["info"] == binfo[k][i][key]["info"]
                    assert s[key]["resid"] == binfo[k][i][key]["resid"]
                assert s["div_before"] == binfo[k][i]["div_before"]
            for fid in FIELDS:
                np.testing.assert_array_equal(b.get_field(i, fid), g.get_field(fid), err_msg=f"instance {i} field {fid}")
            g.close()


def test_batch_instances_converge():
    from drd_application_amd.solver import GpuBatch
    cfgs = _configs(8, seed=5, sizes=(64, 128, 256))
    with GpuBatch([c.params for c in cfgs]) as b:
        _load_batch(b, cfgs)
        for k in range(3):
            infos = b.step(1 if k == 0 else 2, k * cfgs[0].params.dt)
            for inf in infos:
                assert inf["phmu"]["info"] == 0
                if k > 0:  # BDF2 with the DCT preconditioner
                    assert inf["velo"]["info"] == 0


def test_batch_rejects_mixed_parameters_and_nonlinear_preconditioners():
    from drd_application_amd.solver import DrdError, GpuBatch
    cfgs = _configs(3, sizes=(64,))
    ps = [c.params.copy() for c in cfgs]
    ps[1].dt = 2 * ps[1].dt
    with pytest.raises(DrdError):
        GpuBatch(ps)
    ps = [c.params.copy() for c in cfgs]
    for q in ps:
        q.velo_prec = abi.PREC_LEARNED
    with pytest.raises(DrdError):
        GpuBatch(ps)


@pytest.mark.parametrize("prec", [abi.PREC_DCT32, abi.PREC_JACOBI])
def test_graph_replays_match_eager_steps(prec):
    from drd_application_amd.solver import GpuSolver
    cfg = geometry.config("small")
    p = cfg.params.copy()
    p.velo_prec = prec
    p.ch_prec = abi.CH_PREC_BLOCK
    a, b = GpuSolver(p), GpuSolver(p)
    geometry.load(a, cfg)
    geometry.load(b, cfg)
    b.use_graphs(True)
    for k in range(6):
        sa = a.step(1 if k == 0 else 2, k * p.dt)
        sb = b.step(1 if k == 0 else 2, k * p.dt)
        assert sa == sb, k
        for fid in FIELDS:
            np.testing.assert_array_equal(a.get_field(fid), b.get_field(fid), err_msg=f"step {k} field {fid}")
    a.close()
    b.close()
