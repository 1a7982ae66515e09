"""Autotuner for the step's launch knobs: threads per block, re-sort interval
and sort tile shape (rp_set_tuning).

None of the knobs changes a result (each particle's step reads only its own
columns and its deviates are keyed by its release id), so the tuner times
them on the run itself: every candidate replays the same window of steps
from the same state (the caller's restore() puts the particles and hydro
slots back to the window's start), starting with a fresh re-sort, and is
scored by device time of those steps (HIP events on the context stream,
rp_kernel_timing).  The window spans at least twice the largest re-sort
interval among the candidates, so each candidate's interval shows in its
time; the active particle-steps are the same for every candidate, which the
tuner checks.  After the trials the state is restored once more and the
fastest candidate's knobs are left set.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import itertools

BLOCKS = (128, 256)
SORT_INTERVALS = (8, 32)
TILES = ((1, 1), (4, 4), (16, 4))


@dataclass
class Trial:
    block: int
    sort_interval: int
    tile: tuple
    steps: int = 0
    active: int = 0
    ms: float = 0.0

    @property
    def ns_per_particle_step(self):
        return 1e6 * self.ms / self.active if self.active else float("inf")

    def as_dict(self):
        return {"block": self.block, "sort_interval": self.sort_interval, "tile": list(self.tile),
                "steps": self.steps, "active": self.active, "ms": self.ms,
                "ns_per_particle_step": self.ns_per_particle_step}


@dataclass
class TuneResult:
    best: Trial
    trials: list = field(default_factory=list)

    def as_dict(self):
        return {"best": self.best.as_dict(), "trials": [t.as_dict() for t in self.trials]}


def candidates(blocks=BLOCKS, intervals=SORT_INTERVALS, tiles=TILES):
    return [Trial(b, s, t) for b, s, t in itertools.product(blocks, intervals, tiles)]


def autotune(ctx, window, run_step, restore, trials=None, log=None) -> TuneResult:
    """Time each candidate on the same window of steps.

    ctx: abi.Context (or any object with set_tuning, synchronize,
    kernel_timing and active_steps).  window: the schedule steps every
    candidate replays; run_step(st) performs one of them (hydro loads
    included); restore() returns particles and hydro slots to the state
    before window[0].  Leaves the state restored and ctx tuned to the
    fastest candidate."""
    trials = trials or candidates()
    longest = max(t.sort_interval for t in trials)
    if len(window) < 2 * longest:
        raise ValueError(f"a tuning window of {len(window)} steps cannot tell sort intervals up to "
                         f"{longest} apart: it needs at least {2 * longest}")
    done = []
    for t in trials:
        restore()
        ctx.set_tuning(t.block, t.sort_interval, t.tile)
        ctx.synchronize()
        ctx.kernel_timing()
        a0 = ctx.active_steps()
        for st in window:
            run_step(st)
        ms, launches = ctx.kernel_timing()
        t.steps = len(window)
        t.active = ctx.active_steps() - a0
        t.ms = ms * launches
        done.append(t)
        if log:
            log(f"  tune block={t.block} sort={t.sort_interval} tile={t.tile}: "
                f"{t.ms:.2f} ms over {t.steps} steps, {t.ns_per_particle_step:.3f} ns/particle-step")
    if len({t.active for t in done}) != 1:
        raise RuntimeError("tuning trials stepped different particle counts: restore() did not "
                           "return to the same state")
    best = min(done, key=lambda t: t.ms)
    restore()
    ctx.set_tuning(best.block, best.sort_interval, best.tile)
    return TuneResult(best, done)
