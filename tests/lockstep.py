"""Step the HIP library and the oracle side by side on the same inputs and the
same injected deviates, each on its own particle state: the harness of the
GPU tests that compare per-step samplers and long windows with the oracle."""
import numpy as np

from oracle import oracle as O
from romspath_amd.nesting import scale_offsets
from romspath_amd.runner import GpuRun, RecordSource, injected_deviates, ndev_of


class Lockstep:
    def __init__(self, cfg, precision="fp64", sort_interval=0, records=None, particles=None,
                 extra=None, block=0):
        """extra(gi, record) -> dict of additional fields of that record
        (salt/temp, bustr/bvstr), loaded into both codes."""
        self.cfg = cfg
        self.records = records or RecordSource(cfg)
        self.run = GpuRun(cfg, precision=precision, sort_interval=sort_interval,
                          records=self.records, block=block)
        if particles is not None:
            self.run.ctx.set_particles(particles)
        ng = cfg.ngrid
        if ng > 1:
            scl, off = scale_offsets(cfg.grids, cfg.refine)
            self.ogs = [O.OracleGrid(g, scl[i], off[i]) for i, g in enumerate(cfg.grids)]
        else:
            self.ogs = [O.OracleGrid(cfg.grids[0])]
        self.P = O.make_params(cfg.params, ng)
        self.state = O.OracleState(particles if particles is not None else cfg.particles)
        self.ndev = ndev_of(cfg.params)
        self.extra = extra
        self.active = 0

    @property
    def ctx(self):
        return self.run.ctx

    def load(self, st):
        for ld in st.loads:
            for gi in range(self.cfg.ngrid):
                rec = dict(self.records.get(gi + 1, ld.record, ld.filenum))
                if self.extra is not None:
                    rec.update(self.extra(gi, ld.record))
                self.ctx.set_hydro_slot(gi + 1, ld.slot, rec)
                if "salt" in rec:
                    self.ctx.set_hydro_tracers(gi + 1, ld.slot, rec["salt"], rec["temp"])
                if "bustr" in rec:
                    self.ctx.set_hydro_bstress(gi + 1, ld.slot, rec["bustr"], rec["bvstr"])
                if "wetdry_mask_rho" in rec:
                    self.ctx.set_hydro_wetdry(gi + 1, ld.slot, rec["wetdry_mask_rho"],
                                              rec["wetdry_mask_u"], rec["wetdry_mask_v"])
                self.ogs[gi].set_slot(ld.slot, rec)

    def step(self, st, want_diag=False):
        n = self.state.n
        dev = injected_deviates(self.cfg.params.seed, st.step_index, n, self.ndev)
        self.ctx.set_deviates(dev)
        self.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=want_diag)
        T = O.make_time(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf)
        self.active += O.update_particles(self.ogs, self.P, T, self.state, dev, self.ndev)
        return T

    def close(self):
        self.run.close()


def check_same_state(g, o, what=""):
    """FP64 parity: discrete columns identical, positions within 1e-9 cell."""
    assert np.array_equal(g["status"], o.status), what
    assert np.array_equal(g["gid"], o.gid), what
    assert np.array_equal(g["isin"], o.isin), what
    assert np.abs(g["x"] - o.x).max() <= 1e-9 and np.abs(g["y"] - o.y).max() <= 1e-9, what
    assert np.abs(g["z"] - o.z).max() <= 1e-9, what
