"""Ensemble and batching entry points: rp_step_batch (one HIP graph per
external step), rp_step_scenario (per-ensemble launches), rp_set_tuning.
None of them may change a result: FP64 mode stays exact against the oracle,
and with counter-based deviates the forms agree bit for bit with rp_step.
"""
import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun, injected_deviates, ndev_of
from romspath_amd.schedule import schedule

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "age", "status", "gid", "isin")


def _exact_vs_oracle(g, o, cfg):
    assert np.array_equal(g["status"], o.status)
    assert np.array_equal(g["gid"], o.gid)
    assert np.array_equal(g["isin"], o.isin)
    layer = np.maximum(o.bath, 1.0) / cfg.grids[0].N
    assert np.abs(g["x"] - o.x).max() <= 1e-9
    assert np.abs(g["y"] - o.y).max() <= 1e-9
    assert (np.abs(g["z"] - o.z) / layer).max() <= 1e-9


def test_batch_graph_c1_matches_oracle():
    cfg = synth.config1(n_particles=2000)
    print("case c1 batch graph fp64 sort=7", flush=True)
    run = GpuRun(cfg, precision="fp64", sort_interval=7)
    g = run.run(max_steps=120, mode="batch", use_graph=True, diag_every_step=True)
    act = run.ctx.active_steps()
    run.close()
    o = O.run_oracle(cfg, max_steps=120)
    _exact_vs_oracle(g, o, cfg)
    assert act == o.active


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_step_mode_with_turbulence(use_graph):
    cfg = synth.config_mini_nested(n_particles=600)
    print(f"case nested mini fp32 batch graph={use_graph} vs step", flush=True)
    outs = []
    for mode in ("step", "batch"):
        run = GpuRun(cfg, precision="fp32", sort_interval=5)
        outs.append(run.run(max_steps=60, mode=mode, use_graph=use_graph))
        run.close()
    for k in KEYS:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_scenarios_match_oracle():
    cfg = synth.config_mini(n_particles=400, days=2 / 24.0)
    counts = [100, 50, 200, 50]
    print("case mini fp64 4 scenarios sort=3 injected", flush=True)
    run = GpuRun(cfg, precision="fp64", sort_interval=3)
    g = run.run(max_steps=60, inject=True, mode="scenario", counts=counts, diag_every_step=True)
    act = run.ctx.active_steps()
    run.close()
    o = O.run_oracle(cfg, max_steps=60, inject=True)
    _exact_vs_oracle(g, o, cfg)
    assert act == o.active


def test_tuning_knobs_do_not_change_results():
    cfg = synth.config_mini(n_particles=400, days=2 / 24.0)
    print("case mini fp64 block=64 sort=2 tile=4x2 injected", flush=True)
    run = GpuRun(cfg, precision="fp64")
    run.ctx.set_tuning(64, 2, (4, 2))
    g = run.run(max_steps=60, inject=True, diag_every_step=True)
    run.close()
    o = O.run_oracle(cfg, max_steps=60, inject=True)
    _exact_vs_oracle(g, o, cfg)


def test_switch_from_global_sort_to_scenarios():
    # global re-sorts mix the scenarios' stored ranges; the first scenario
    # step restores release order
    cfg = synth.config_mini(n_particles=400, days=2 / 24.0)
    counts = [150, 250]
    prm = cfg.params
    n = cfg.particles.n
    ndev = ndev_of(prm)
    print("case mini fp64 global sort then scenarios", flush=True)
    run = GpuRun(cfg, precision="fp64", sort_interval=2)
    ctx = run.ctx
    ctx.set_scenarios(counts)
    for st in schedule(prm.dt, prm.idt, prm.days, prm.iprint, 60):
        run.load(st)
        ctx.set_deviates(injected_deviates(prm.seed, st.step_index, n, ndev))
        if st.step_index < 25 or st.step_index >= 45:
            ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=True)
        else:
            for s in range(len(counts)):
                ctx.step_scenario(s, st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index,
                                  want_diag=True)
    g = ctx.get_particles()
    run.close()
    o = O.run_oracle(cfg, max_steps=60, inject=True)
    _exact_vs_oracle(g, o, cfg)


def test_scenario_arguments_rejected():
    from romspath_amd.abi import RpError
    cfg = synth.config_mini(n_particles=100)
    run = GpuRun(cfg, precision="fp32")
    with pytest.raises(RpError):
        run.ctx.set_scenarios([50, 40])          # does not add up to 100
    with pytest.raises(RpError):
        run.ctx.step_scenario(0, (0.0, 1.0, 2.0), (0.0, 1.0, 2.0), 1, 1, 2, 3, 0)
    with pytest.raises(RpError):
        run.ctx.set_tuning(96, 4, (1, 1))         # block not a multiple of 64
    run.close()
