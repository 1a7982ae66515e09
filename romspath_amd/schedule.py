"""Host-side time schedule around the step (ROMSPath.f90:151-546).

The reference keeps three hydro slots t_b, t_c, t_f.  Records 1, 2, 3 fill
slots 1, 2, 3 at start-up (ROMSPath.f90:210-213, 307-309).  From the third
external step on, the slots rotate and record `tstep` is read into the new
forward slot (ROMSPath.f90:488-493); `tstep` advances after every external
step (ROMSPath.f90:510).  Record r holds model time (r-1)*dt.  The step itself
only sees the slot numbers and the times ex(3), ix(3).
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass
class Load:
    slot: int       # 1-based slot to fill
    record: int     # 1-based record number (tstep: within the file when read from files)
    filenum: int | None = None   # history file number (HydroFiles), None for in-memory records


@dataclass
class InternalStep:
    ets: int
    its: int
    ex: tuple
    ix: tuple
    tb: int
    tc: int
    tf: int
    step_index: int           # 0-based count of internal steps
    loads: list               # records to read before this step
    output: bool              # dataOutput after this step (printdt >= iprint)


def schedule(dt: int, idt: int, days: float, iprint: int | None = None, max_steps=None,
             hydro=None):
    """Yield the internal steps of a run in order (run_ROMSPath and below).

    hydro: records read from history files (anything with filenum0,
    multifile and tdim(filenum), e.g. roms.HydroFiles).  Loads then name the
    file, and at the top of each external step a tstep past the records of
    the current file moves to record 1 of the next file, or stops the run
    when multifile is off (ROMSPath.f90:473-480)."""
    t_b, t_c, t_f = 1, 2, 3
    tstep = 1
    filenum = hydro.filenum0 if hydro is not None else None
    pending = [Load(1, 1, filenum), Load(2, 2, filenum), Load(3, 3, filenum)]
    seconds = int(days * 86400.0)
    stepT = seconds // dt
    stepIT = int(dt / idt)
    printdt = 0
    iprint = iprint if iprint is not None else dt
    k = 0
    for ets in range(1, stepT + 1):
        if hydro is not None and tstep > hydro.tdim(filenum):
            tstep = 1
            filenum += 1
            if not hydro.multifile:
                raise RuntimeError("NOT ENOUGH TIMES STEP TO CONTINUE (multifile is off)")
        if ets > 2:
            t_b = t_b % 3 + 1
            t_c = t_c % 3 + 1
            t_f = t_f % 3 + 1
            pending.append(Load(t_f, tstep, filenum))
        ex = (float((ets - 2) * dt), float((ets - 1) * dt), float(ets * dt))
        for its in range(1, stepIT + 1):
            ix = (ex[1] + float((its - 2) * idt), ex[1] + float((its - 1) * idt),
                  ex[1] + float(its * idt))
            printdt += idt
            out = printdt >= iprint
            if out:
                printdt = 0
            yield InternalStep(ets, its, ex, ix, t_b, t_c, t_f, k, pending, out)
            pending = []
            k += 1
            if max_steps is not None and k >= max_steps:
                return
        tstep += 1


def slots_before(steps, k):
    """The load that last filled each slot before steps[k] runs its own
    loads: {slot: Load}.  Replaying steps[k:] from a saved particle state
    needs the hydro slots put back to these records first."""
    last = {}
    for st in steps[:k]:
        for ld in st.loads:
            last[ld.slot] = ld
    return last


def record_time(record: int, dt: int) -> float:
    return float((record - 1) * dt)
