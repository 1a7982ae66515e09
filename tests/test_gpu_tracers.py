"""rp_sample_tracers (SaltTempOn, ROMSPath.f90:958-981) on the GPU against
the oracle: bit for bit in FP64 mode (the affine case and nested grids with
an island after 30 steps), within FP32 rounding in the FP32-field form.
"""
import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun

from lockstep import Lockstep
from romspath_amd.schedule import schedule

from test_tracers import A0, BX, CZ, DT, tracer_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_affine_tracer_matches_oracle(precision):
    cfg, slot_recs, ex, x, y, z, zk = tracer_case()
    parts = synth.Particles(x, y, z, np.zeros(len(x)))
    cfg2 = synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, parts, cfg.refine)
    print(f"case tracers affine C1 {precision}", flush=True)
    run = GpuRun(cfg2, precision=precision)
    for s, rec in enumerate(slot_recs):
        run.ctx.set_hydro_slot(1, s + 1, rec)
        run.ctx.set_hydro_tracers(1, s + 1, rec["salt"], rec["temp"])
    ix = (0.0, 600.0, 720.0)
    salt, temp = run.ctx.sample_tracers(ex, ix, 1, 2, 3)
    run.close()
    og = O.OracleGrid(cfg.grids[0])
    for s, rec in enumerate(slot_recs):
        og.set_slot(s + 1, rec)
    os_, ot = O.sample_tracers([og], ex, ix, 1, 2, 3, x, y, z, np.ones(len(x), np.int32))
    if precision == "fp64":
        assert np.array_equal(salt, os_) and np.array_equal(temp, ot)
    else:
        assert np.abs(salt - os_).max() < 1e-4 and np.abs(temp - ot).max() < 2e-4


def test_nested_tracers_after_steps_match_oracle():
    cfg = synth.config_mini_nested(n_particles=600)
    prm = cfg.params
    steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, 30))
    rng = np.random.default_rng(9)
    tr = {}
    for gi, g in enumerate(cfg.grids):
        for s in range(3):
            salt = 34.0 + rng.standard_normal((g.N, g.M, g.L))
            tr[gi, s] = (salt, 10.0 + 3.0 * rng.standard_normal((g.N, g.M, g.L)))
    print("case tracers nested mini fp64 after 30 steps", flush=True)
    run = GpuRun(cfg, precision="fp64")
    for st in steps:
        run.load(st)
        run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)
    for gi in range(cfg.ngrid):
        for s in range(3):
            run.ctx.set_hydro_tracers(gi + 1, s + 1, *tr[gi, s])
    last = steps[-1]
    p = run.ctx.get_particles()
    salt, temp = run.ctx.sample_tracers(last.ex, last.ix, last.tb, last.tc, last.tf)
    # the oracle grids with the slots the last step used
    from romspath_amd.nesting import scale_offsets
    scl, off = scale_offsets(cfg.grids, cfg.refine)
    ogs = [O.OracleGrid(g, scl[i], off[i]) for i, g in enumerate(cfg.grids)]
    loaded = {}
    for st in steps:
        for ld in st.loads:
            loaded[ld.slot] = ld.record
    for gi, og in enumerate(ogs):
        for slot, rec_no in loaded.items():
            rec = dict(run.records.get(gi + 1, rec_no))
            rec["salt"], rec["temp"] = tr[gi, slot - 1]
            og.set_slot(slot, rec)
    run.close()
    os_, ot = O.sample_tracers(ogs, last.ex, last.ix, last.tb, last.tc, last.tf, p["x"], p["y"],
                               p["z"], p["gid"])
    assert set(np.unique(p["gid"])) == {1, 2}
    assert np.array_equal(salt, os_) and np.array_equal(temp, ot)


@pytest.mark.parametrize("mean", [False, True])
def test_per_step_tracers_match_oracle(mean):
    """rp_tracers_begin/update/get against the oracle's restatement, the GPU
    and the oracle stepping side by side with the same injected deviates,
    each sampler on its own code's state, bit for bit in FP64: release-time
    values, values kept for particles that did not step, rejected moves
    sampled with the attempted point's depth, and SaltTempMean's mean / mI
    with a reset after a read; re-sorting on, no output steps."""
    cfg = synth.config_mini_nested(n_particles=600)
    prm = cfg.params
    steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, 24))
    print(f"case per-step tracers mini-nested fp64 mean={mean}", flush=True)

    def extra(gi, rec_no):
        g = cfg.grids[gi]
        rng = np.random.default_rng(1000 * gi + rec_no)
        salt = 34.0 + rng.standard_normal((g.N, g.M, g.L))
        return dict(salt=salt, temp=10.0 + 3.0 * rng.standard_normal((g.N, g.M, g.L)))

    ls = Lockstep(cfg, precision="fp64", sort_interval=4, extra=extra)
    rejected = 0
    try:
        ls.load(steps[0])
        ls.ctx.tracers_begin(mean, steps[0].tc)
        la, ps, pt, ms, mt, mI = O.tracers_begin(ls.ogs, steps[0].tc, ls.state, mean)
        assert ps.any()
        checked = 0
        for k, st in enumerate(steps):
            if k > 0:
                ls.load(st)
            ls.step(st)
            ls.ctx.tracers_update(st.ex, st.ix, st.tb, st.tc, st.tf)
            O.tracers_update(ls.ogs, st, ls.state, la, ps, pt, ms, mt)
            rejected += int((ls.state.status == -9.0).sum())
            mI += 1
            if k in (9, len(steps) - 1):
                gs, gt = ls.ctx.tracers_get(reset_mean=True)
                if mean:
                    assert np.array_equal(gs, ms / mI) and np.array_equal(gt, mt / mI)
                    ms[:] = 0.0
                    mt[:] = 0.0
                    mI = 0
                else:
                    assert np.array_equal(gs, ps) and np.array_equal(gt, pt)
                checked += 1
        assert checked == 2
        assert rejected > 0, "the case has no rejected moves"
    finally:
        ls.close()
