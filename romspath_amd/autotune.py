"""Autotuner for the step's launch knobs: threads per block, re-sort interval
and sort tile shape (rp_set_tuning).

None of the knobs changes a result (each particle's step reads only its own
columns and its deviates are keyed by its release id), so the tuner times
them on the run itself: every candidate takes the next steps of the
schedule, starting with a fresh re-sort, and is scored by device time per
active particle-step (HIP events on the context stream, rp_kernel_timing).
The steps it spends are ordinary steps of the run.
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


def autotune(ctx, next_step, trial_steps=8, trials=None, log=None) -> TuneResult:
    """Time each candidate on the next trial_steps steps of the run.

    ctx: abi.Context with grids, particles and the first hydro records loaded.
    next_step(): performs the run's next internal step (hydro loads included)
    and returns False when the schedule is exhausted.
    Leaves ctx tuned to the fastest candidate."""
    trials = trials or candidates()
    done = []
    for t in trials:
        ctx.set_tuning(t.block, t.sort_interval, t.tile)
        ctx.kernel_timing()
        a0 = ctx.active_steps()
        k = 0
        while k < trial_steps and next_step():
            k += 1
        ms, launches = ctx.kernel_timing()
        t.steps = k
        t.active = ctx.active_steps() - a0
        t.ms = ms * launches
        done.append(t)
        if log:
            log(f"  tune block={t.block} sort={t.sort_interval} tile={t.tile}: "
                f"{t.ns_per_particle_step:.3f} ns/particle-step over {k} steps")
        if k < trial_steps:
            break
    best = min(done, key=lambda t: t.ns_per_particle_step)
    ctx.set_tuning(best.block, best.sort_interval, best.tile)
    return TuneResult(best, done)
