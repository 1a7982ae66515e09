"""The autotuner (romspath_amd/autotune.py) on a simulated context: it picks
the fastest of the candidates, every candidate replays the same window from
the restored state, and windows too short to tell the re-sort intervals
apart are refused.  No GPU."""
import pytest

from romspath_amd.autotune import Trial, autotune
from romspath_amd.schedule import schedule, slots_before


class FakeCtx:
    """Device time per step depends on the knobs: a re-sort costs `sort_ms`
    and a step costs less the more recently the particles were sorted."""

    def __init__(self, cost):
        self.cost = cost
        self.knobs = None
        self.ms = 0.0
        self.launches = 0
        self.active = 0
        self.since_sort = 0

    def set_tuning(self, block, sort_interval, tile):
        self.knobs = (block, sort_interval, tuple(tile))
        self.since_sort = 0

    def synchronize(self):
        pass

    def kernel_timing(self):
        ms, n = self.ms, self.launches
        self.ms, self.launches = 0.0, 0
        return (ms / n if n else 0.0), n

    def active_steps(self):
        return self.active

    def step(self, n_active):
        self.ms += self.cost(self.knobs, self.since_sort)
        self.launches += 1
        self.active += n_active
        self.since_sort += 1
        if self.since_sort >= self.knobs[1]:
            self.since_sort = 0


def _window(n):
    return list(schedule(3600, 120, 5.0, max_steps=n))


def test_picks_the_fastest_of_two_choices():
    # block 256 is faster per step; everything else equal
    ctx = FakeCtx(lambda k, since: 1.0 if k[0] == 256 else 1.5)
    log = []
    trials = [Trial(128, 8, (1, 1)), Trial(256, 8, (1, 1))]
    restores = []
    res = autotune(ctx, _window(20), lambda st: ctx.step(1000), lambda: restores.append(1),
                   trials=trials, log=log.append)
    assert res.best.block == 256 and ctx.knobs == (256, 8, (1, 1))
    assert len(restores) == len(trials) + 1             # before each trial and at the end
    assert all(t.steps == 20 and t.active == 20 * 1000 for t in res.trials)
    assert len(log) == 2


def test_sort_interval_is_told_apart():
    # a re-sort costs 5 ms; steps slow down by 0.1 ms per step since the last sort:
    # over 64 steps interval 8 costs 8 sorts, interval 32 only 2 but slower steps
    def cost(k, since):
        return (5.0 if since == 0 else 0.0) + 1.0 + 0.1 * since

    ctx = FakeCtx(cost)
    trials = [Trial(256, 8, (1, 1)), Trial(256, 32, (1, 1))]
    res = autotune(ctx, _window(64), lambda st: ctx.step(1), lambda: None, trials=trials)
    t8, t32 = res.trials
    want8 = 8 * (5.0 + sum(1.0 + 0.1 * s for s in range(8)))
    want32 = 2 * (5.0 + sum(1.0 + 0.1 * s for s in range(32)))
    assert t8.ms == pytest.approx(want8) and t32.ms == pytest.approx(want32)
    assert res.best.sort_interval == (8 if want8 < want32 else 32)


def test_short_window_refused():
    ctx = FakeCtx(lambda k, since: 1.0)
    with pytest.raises(ValueError):
        autotune(ctx, _window(40), lambda st: ctx.step(1), lambda: None,
                 trials=[Trial(256, 8, (1, 1)), Trial(256, 32, (1, 1))])


def test_unequal_replays_detected():
    ctx = FakeCtx(lambda k, since: 1.0)
    calls = {"n": 0}

    def run_step(st):
        ctx.step(1 + calls["n"])      # a restore that does not restore
        calls["n"] += 1

    with pytest.raises(RuntimeError):
        autotune(ctx, _window(16), run_step, lambda: None,
                 trials=[Trial(256, 8, (1, 1)), Trial(128, 8, (1, 1))])


def test_slots_before_a_window():
    steps = list(schedule(3600, 120, 5.0, max_steps=200))
    # before step 100 (external step 4, its 11): records 1-3 at start, record 3
    # into slot 1 at ets 3 (step 60), record 4 into slot 2 at ets 4 (step 90)
    s = slots_before(steps, 100)
    assert {k: v.record for k, v in s.items()} == {1: 3, 2: 4, 3: 3}
    assert {k: v.record for k, v in slots_before(steps, 1).items()} == {1: 1, 2: 2, 3: 3}
