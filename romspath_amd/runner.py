"""Drive librp through a run: the host loop of ROMSPath.f90:151-546 around
rp_step, with hydro records produced by the synthetic models."""
from __future__ import annotations

import numpy as np

from . import abi
from .nesting import scale_offsets
from .schedule import record_time, schedule


class RecordSource:
    """Hydro records by (grid, record number), cached."""

    def __init__(self, cfg, cache=True):
        self.cfg = cfg
        self.cache = {} if cache else None

    def get(self, ng, record, filenum=None):
        assert filenum is None, "in-memory records have no files"
        key = (ng, record)
        if self.cache is not None and key in self.cache:
            return self.cache[key]
        rec = self.cfg.hydro[ng - 1].record(record_time(record, self.cfg.params.dt))
        if self.cache is not None:
            self.cache[key] = rec
        return rec


def ndev_of(prm):
    return 2 + (prm.loop if prm.VTurbOn else 0)


def injected_deviates(seed, step_index, n, ndev):
    """Deterministic N(0,1) deviates for one step, [n][ndev] (test hook input)."""
    rng = np.random.default_rng([seed, step_index])
    return rng.standard_normal((n, ndev))


class GpuRun:
    def __init__(self, cfg, precision="fp32", block=0, records=None, sort_interval=0):
        self.cfg = cfg
        ng = cfg.ngrid
        self.ctx = abi.Context(abi.make_config(cfg.params, ngrid=ng, precision=precision,
                                               block=block, sort_interval=sort_interval))
        scl, off = scale_offsets(cfg.grids, cfg.refine) if ng > 1 else (None, None)
        for i, g in enumerate(cfg.grids):
            self.ctx.set_grid(i + 1, g, None if scl is None else scl[i], None if off is None else off[i])
            if hasattr(g, "lon"):            # a grid read from a history file (roms.RomsGrid)
                self.ctx.set_lonlat(i + 1, g.lon, g.lat)
                self.ctx.set_grid_angle(i + 1, g.spherical, g.angle)
            else:
                # Cartesian grids: InitGrid reads x_rho / y_rho into lon_rho / lat_rho
                self.ctx.set_lonlat(i + 1, g.x_rho, g.y_rho)
        self.ctx.set_particles(cfg.particles)
        self.records = records or RecordSource(cfg)

    def load_initial(self):
        """Records 1, 2, 3 into slots 1, 2, 3, as ini_ROMSPath does (ROMSPath.f90:307-309)."""
        for slot in (1, 2, 3):
            for g in range(1, self.cfg.ngrid + 1):
                rec = self.records.get(g, slot)
                self.ctx.set_hydro_slot(g, slot, rec)
                self._set_wetdry(g, slot, rec)

    def _set_wetdry(self, g, slot, rec):
        if "wetdry_mask_rho" in rec:
            self.ctx.set_hydro_wetdry(g, slot, rec["wetdry_mask_rho"], rec["wetdry_mask_u"],
                                      rec["wetdry_mask_v"])

    def load(self, st):
        """updateHydro for the step's loads: the slot's hydro fields, and salt/temp
        (SaltTempOn), bustr/bvstr (WriteBottom) and the wet/dry masks (-DWETDRY)
        when the record has them."""
        for ld in st.loads:
            for g in range(1, self.cfg.ngrid + 1):
                rec = self.records.get(g, ld.record, ld.filenum)
                self.ctx.set_hydro_slot(g, ld.slot, rec)
                if "salt" in rec and "temp" in rec:
                    self.ctx.set_hydro_tracers(g, ld.slot, rec["salt"], rec["temp"])
                if "bustr" in rec and "bvstr" in rec:
                    self.ctx.set_hydro_bstress(g, ld.slot, rec["bustr"], rec["bvstr"])
                self._set_wetdry(g, ld.slot, rec)

    def run(self, max_steps=None, inject=False, on_step=None, diag_every_step=False,
            mode="step", counts=None, use_graph=True, output=None, salttemp_mean=False,
            release_lonlat=None):
        """mode "step": one rp_step per internal step; "batch": the steps of each
        external step in one rp_step_batch (one HIP graph when use_graph);
        "scenario": rp_step_scenario for each of the scenarios given by counts.
        output: a TrajectoryFile that gets a record whenever the reference's
        printdt reaches iprint (ROMSPath.f90:532-543), at model time ix(3).
        When the file has the salinity/temperature (SaltTempOn; salttemp_mean
        for SaltTempMean) or HOB/bustr/bvstr (WriteBottom) variables, the
        per-particle samplers start after the first loads, as ini_ROMSPath
        starts them, and are updated after every step (mode "step" only).
        The file gets the release record at time 0 before the first step
        (writeNetCDF(0, pLon, pLat), ROMSPath.f90:453), with release_lonlat
        as its lon/lat when given (the release file's values) and sampled
        at the release positions otherwise; every record's model_time is
        time + reftime, reftime being the last grid's (InitGrid,
        grid_module.f90:374-376; 0 for grids not read from a file)."""
        from .output import FLAGS
        prm = self.cfg.params
        n = self.cfg.particles.n
        ndev = ndev_of(prm)
        want_tr = output is not None and bool(output.flags & FLAGS["salttemp"])
        want_bot = output is not None and bool(output.flags & FLAGS["bottom"])
        if (want_tr or want_bot) and mode != "step":
            raise ValueError("the SaltTempOn / WriteBottom samplers update after every rp_step: "
                             "use mode='step'")
        hydro = self.records if hasattr(self.records, "tdim") else None
        steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, max_steps, hydro=hydro))
        if diag_every_step:
            for st in steps:
                st.output = True
        if mode == "scenario":
            self.ctx.set_scenarios(counts)
        if mode == "batch":
            if inject:
                raise ValueError("injected deviates are per step: use mode='step' or 'scenario'")
            i = 0
            while i < len(steps):
                j = i + 1
                while j < len(steps) and not steps[j].loads:
                    j += 1
                self.load(steps[i])
                self.ctx.step_batch(steps[i:j], use_graph=use_graph)
                if on_step is not None:
                    for st in steps[i:j]:
                        on_step(st, self)
                i = j
            return self.ctx.get_particles()
        reftime = float(getattr(self.cfg.grids[-1], "reftime", 0.0))
        for k, st in enumerate(steps):
            self.load(st)
            if k == 0:
                if want_bot:
                    self.ctx.bottom_begin()
                if want_tr:
                    self.ctx.tracers_begin(salttemp_mean, st.tc)
                if output is not None:
                    from .output import data_output
                    data_output(self.ctx, output, 0.0 + reftime, lonlat=release_lonlat)
            if inject:
                self.ctx.set_deviates(injected_deviates(prm.seed, st.step_index, n, ndev))
            if mode == "scenario":
                for s_ in range(len(counts)):
                    self.ctx.step_scenario(s_, st.ex, st.ix, st.ets, st.tb, st.tc, st.tf,
                                           st.step_index, want_diag=st.output)
            else:
                self.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index,
                              want_diag=st.output)
                if want_bot:
                    self.ctx.bottom_update(st.ex, st.ix, st.tb, st.tc, st.tf)
                if want_tr:
                    self.ctx.tracers_update(st.ex, st.ix, st.tb, st.tc, st.tf)
            if output is not None and st.output:
                from .output import data_output
                data_output(self.ctx, output, float(int(st.ix[2])) + reftime)
            if on_step is not None:
                on_step(st, self)
        return self.ctx.get_particles()

    def snapshot(self):
        """The particle columns now, as a Particles set rp_set_particles takes
        back (release order)."""
        from .synth import Particles
        p = self.ctx.get_particles()
        return Particles(p["x"], p["y"], p["z"], self.cfg.particles.dob.copy(), gid=p["gid"],
                         status=p["status"], age=p["age"], isin=p["isin"])

    def restore(self, snap, steps, k):
        """Particles from a snapshot taken before steps[k], and the hydro slots
        as they were then (the records that last filled them)."""
        from .schedule import slots_before
        self.ctx.set_particles(snap)
        for ld in slots_before(steps, k).values():
            for g in range(1, self.cfg.ngrid + 1):
                self.ctx.set_hydro_slot(g, ld.slot, self.records.get(g, ld.record, ld.filenum))

    def close(self):
        self.ctx.close()
