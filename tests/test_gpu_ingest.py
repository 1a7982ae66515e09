"""Hydro ingest (rp_set_hydro_slot): the caller's arrays are copied before the
call returns, the host-to-device copies run on a second stream while queued
steps execute, and each field's conversion is ordered on the step stream.
None of that may change a result: FP64 mode stays exact against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun, injected_deviates, ndev_of
from romspath_amd.schedule import schedule

pytestmark = pytest.mark.gpu

STEPS = 120      # 4 h at idt 120 s: six record loads after the initial three


def _exact_vs_oracle(g, o, cfg):
    assert np.array_equal(g["status"], o.status)
    assert np.array_equal(g["gid"], o.gid)
    assert np.array_equal(g["isin"], o.isin)
    layer = np.maximum(o.bath, 1.0) / cfg.grids[0].N
    assert np.abs(g["x"] - o.x).max() <= 1e-9
    assert np.abs(g["y"] - o.y).max() <= 1e-9
    assert (np.abs(g["z"] - o.z) / layer).max() <= 1e-9


def _drive(cfg, upload):
    prm = cfg.params
    n = cfg.particles.n
    ndev = ndev_of(prm)
    run = GpuRun(cfg, precision="fp64", sort_interval=5)
    for st in schedule(prm.dt, prm.idt, prm.days, prm.iprint, STEPS):
        for ld in st.loads:
            for g in range(1, cfg.ngrid + 1):
                upload(run, g, ld.slot, run.records.get(g, ld.record))
        run.ctx.set_deviates(injected_deviates(prm.seed, st.step_index, n, ndev))
        run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=True)
    out = run.ctx.get_particles()
    run.close()
    return out


def test_host_arrays_reusable_when_upload_returns():
    cfg = synth.config_mini(n_particles=500, days=STEPS * 120 / 86400.0)
    print("case mini fp64, host arrays overwritten after every upload", flush=True)

    def upload(run, g, slot, rec):
        tmp = {k: np.array(v, dtype=np.float64, order="F", copy=True) for k, v in rec.items()}
        run.ctx.set_hydro_slot(g, slot, tmp)
        for v in tmp.values():
            v.fill(np.nan)

    g = _drive(cfg, upload)
    o = O.run_oracle(cfg, max_steps=STEPS, inject=True)
    _exact_vs_oracle(g, o, cfg)


def test_uploads_to_one_slot_apply_in_call_order():
    # a decoy record goes into the slot first; the real one, uploaded right
    # after it with no step in between, must be what the next steps read
    cfg = synth.config_mini(n_particles=500, days=STEPS * 120 / 86400.0)
    print("case mini fp64, decoy upload before every real one", flush=True)

    def upload(run, g, slot, rec):
        decoy = {k: 3.0 * np.asarray(v) + 0.25 for k, v in rec.items()}
        run.ctx.set_hydro_slot(g, slot, decoy)
        run.ctx.set_hydro_slot(g, slot, rec)

    g = _drive(cfg, upload)
    o = O.run_oracle(cfg, max_steps=STEPS, inject=True)
    _exact_vs_oracle(g, o, cfg)


def test_nested_fp32_async_uploads_match_synchronised_ones():
    # the same run with the stream synchronised after every upload
    cfg = synth.config_mini_nested(n_particles=600)
    print("case nested mini fp32, async uploads vs synchronised", flush=True)
    outs = []
    for sync in (False, True):
        run = GpuRun(cfg, precision="fp32", sort_interval=5)
        prm = cfg.params
        for st in schedule(prm.dt, prm.idt, prm.days, prm.iprint, 60):
            for ld in st.loads:
                for g in range(1, cfg.ngrid + 1):
                    run.ctx.set_hydro_slot(g, ld.slot, run.records.get(g, ld.record))
                    if sync:
                        run.ctx.synchronize()
            run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)
        outs.append(run.ctx.get_particles())
        run.close()
    for k in ("x", "y", "z", "status", "gid", "isin"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
