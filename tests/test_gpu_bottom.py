"""WriteBottom on the GPU (rp_set_hydro_bstress, rp_bottom_begin /
rp_bottom_update / rp_bottom_get; ROMSPath.f90:935-951, 2056-2065) against
the oracle's update applied to the GPU's own particle state after every
step: bit for bit in FP64 mode, and in the FP32-field form the same minimum
heights with the stresses within FP32 rounding of the fields.  Nested grids
with an island, re-sorting every 4 steps, P_HOB reset after a read.
"""
import types

import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun
from romspath_amd.schedule import schedule

from test_bottom import stress_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,sort_interval", [("fp64", 0), ("fp64", 4), ("fp32", 4)])
def test_bottom_matches_oracle_each_step(precision, sort_interval):
    cfg = synth.config_mini_nested(n_particles=600)
    prm = cfg.params
    steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, 24))
    print(f"case bottom mini-nested {precision} sort {sort_interval}", flush=True)
    run = GpuRun(cfg, precision=precision, sort_interval=sort_interval)
    ogs = [O.OracleGrid(g) for g in cfg.grids]
    dummy = {k: np.zeros(1) for k in ("zeta", "u", "v", "w", "aks")}
    run.ctx.bottom_begin()
    n = cfg.particles.n
    la = np.ascontiguousarray(cfg.particles.age, dtype=np.float64).copy()
    hob, bu, bv = np.full(n, 9999.0), np.zeros(n), np.zeros(n)
    tol = 0.0 if precision == "fp64" else 2e-7
    checked = 0
    try:
        for k, st in enumerate(steps):
            run.load(st)
            for ld in st.loads:
                for gi, g in enumerate(cfg.grids):
                    f = stress_fields(g, 3600.0 * ld.record, np.random.default_rng(100 * gi + ld.record))
                    run.ctx.set_hydro_bstress(gi + 1, ld.slot, *f)
                    ogs[gi].set_slot(ld.slot, dict(dummy, bustr=f[0], bvstr=f[1]))
            run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)
            run.ctx.bottom_update(st.ex, st.ix, st.tb, st.tc, st.tf)
            p = run.ctx.get_particles()
            state = types.SimpleNamespace(n=n, x=p["x"], y=p["y"], z=p["z"], gid=p["gid"],
                                          age=p["age"], bath=p["bath"])
            T = O.make_time(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf)
            O.bottom_update(ogs, T, state, la, hob, bu, bv)
            if k in (9, len(steps) - 1):
                reset = k == 9
                gh, gu, gv = run.ctx.bottom_get(reset_hob=reset)
                assert np.array_equal(gh, hob)
                if tol == 0.0:
                    assert np.array_equal(gu, bu) and np.array_equal(gv, bv)
                else:
                    assert np.abs(gu - bu).max() < tol and np.abs(gv - bv).max() < tol
                if reset:
                    hob[:] = 9999.0
                checked += 1
        assert set(np.unique(p["gid"])) == {1, 2}
        assert (hob < 9999.0).sum() > n // 2 and (bu != 0.0).sum() > n // 2
    finally:
        run.close()
    assert checked == 2


def test_bottom_errors():
    cfg = synth.config1(n_particles=100)
    run = GpuRun(cfg, precision="fp64")
    try:
        st = next(iter(schedule(cfg.params.dt, cfg.params.idt, cfg.params.days, cfg.params.iprint, 1)))
        from romspath_amd.abi import RpError
        with pytest.raises(RpError) as e:
            run.ctx.bottom_update(st.ex, st.ix, 1, 2, 3)
        assert e.value.code == -2                       # RP_E_STATE: no rp_bottom_begin
        run.ctx.bottom_begin()
        with pytest.raises(RpError) as e:
            run.ctx.bottom_update(st.ex, st.ix, 1, 2, 3)
        assert e.value.code == -2                       # stress slots not loaded
        with pytest.raises(RpError) as e:
            run.ctx.bottom_update(st.ex, st.ix, 0, 2, 3)
        assert e.value.code == -1
        hob, bu, bv = run.ctx.bottom_get()
        assert np.all(hob == 9999.0) and np.all(bu == 0.0) and np.all(bv == 0.0)
    finally:
        run.close()


def test_samplers_with_no_particles():
    """Every sampler entry point accepts an empty particle set (n = 0) and
    returns empty columns without launching anything."""
    cfg = synth.config1(n_particles=0)
    g = cfg.grids[0]
    print("case samplers with no particles", flush=True)
    run = GpuRun(cfg, precision="fp64")
    try:
        st = next(iter(schedule(cfg.params.dt, cfg.params.idt, cfg.params.days, cfg.params.iprint, 1)))
        run.load(st)
        rng = np.random.default_rng(0)
        for slot in (1, 2, 3):
            run.ctx.set_hydro_bstress(1, slot, *stress_fields(g))
            run.ctx.set_hydro_tracers(1, slot, 34.0 + rng.standard_normal((g.N, g.M, g.L)),
                                      rng.standard_normal((g.N, g.M, g.L)))
        run.ctx.bottom_begin()
        run.ctx.tracers_begin(True, 2)
        run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)
        run.ctx.bottom_update(st.ex, st.ix, st.tb, st.tc, st.tf)
        run.ctx.tracers_update(st.ex, st.ix, st.tb, st.tc, st.tf)
        for cols in (run.ctx.bottom_get(reset_hob=True), run.ctx.tracers_get(reset_mean=True),
                     run.ctx.sample_tracers(st.ex, st.ix, st.tb, st.tc, st.tf)):
            assert all(len(c) == 0 for c in cols)
    finally:
        run.close()
