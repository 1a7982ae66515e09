"""rp_step_noalloc, the step that allocates nothing, against rp_step through
the C-ABI: two inputs stepped one after the other on the same GPU, the
second smaller than the first, each equal to rp_step bit for bit in both
field precisions.  The first input (300,000 particles) spans three launch
pairs of the shared scratch array; the second (a nested case) reuses the
array after it, so anything left over from the first would show."""
import numpy as np
import pytest

from romspath_amd import synth
from romspath_amd.runner import GpuRun
from romspath_amd.schedule import schedule

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "age", "status", "gid", "isin", "wd", "zeta", "bath")


def _step_both(cfg, precision, nsteps):
    prm = cfg.params
    ref = GpuRun(cfg, precision=precision)
    na = GpuRun(cfg, precision=precision)
    try:
        steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, nsteps))
        for k, st in enumerate(steps):
            last = k == len(steps) - 1
            for r in (ref, na):
                r.load(st)
            ref.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=last)
            na.ctx.step_noalloc(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index,
                                want_diag=last)
        a, b = ref.ctx.get_particles(), na.ctx.get_particles()
        acts = ref.ctx.active_steps(), na.ctx.active_steps()
    finally:
        ref.close()
        na.close()
    return a, b, acts


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_step_equals_step_large_then_small(precision):
    big = synth.config_mini(n_particles=300_000, dob_spread=0.0, seed=31)
    small = synth.config_mini_nested(n_particles=20_000, seed=32)
    for cfg, nsteps in ((big, 4), (small, 6)):
        print(f"case noalloc {precision} {cfg.name} n={cfg.particles.n} steps={nsteps}", flush=True)
        a, b, (act_a, act_b) = _step_both(cfg, precision, nsteps)
        moved = np.isin(a["status"], (1.0, float(np.float32(1.1)), float(np.float32(1.2))))
        assert moved.sum() > cfg.particles.n // 2
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), (cfg.name, k)
        assert act_a == act_b > 0


def test_noalloc_rejects_recorded_run():
    from romspath_amd.abi import RpError
    cfg = synth.config_mini(n_particles=100)
    run = GpuRun(cfg, precision="fp64")
    try:
        st = next(iter(schedule(cfg.params.dt, cfg.params.idt, cfg.params.days, cfg.params.iprint, 1)))
        run.load(st)
        run.ctx.adjoint_record(4)
        with pytest.raises(RpError) as e:
            run.ctx.step_noalloc(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)
        assert e.value.code == -2
    finally:
        run.close()
