"""SaltTempOn's salt/temp at the particles (ROMSPath.f90:958-981) in the
oracle, pinned by a field that is affine in (x index, level depth, time): the
bilinear weights, the linear vertical blend between rho levels and polintd's
quadratic in time all reproduce it, and above the top (below the bottom) rho
level the value is that level's.  No GPU.
"""
import numpy as np

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import RecordSource

A0, BX, CZ, DT = 30.0, 0.01, -0.02, 1e-5


def tracer_case(n=300, seed=3):
    cfg = synth.config1(n_particles=10)           # flat 50 m, all water, zeta 0
    g = cfg.grids[0]
    S = (g.hc * g.s_rho + 50.0 * g.cs_r) / (g.hc + 50.0)
    zk = 50.0 * S                                 # rho level depths, zeta = 0
    i = np.arange(1, g.L + 1, dtype=np.float64)
    ex = (-3600.0, 0.0, 3600.0)
    recs = RecordSource(cfg)
    slot_recs = []
    for s in range(3):
        rec = dict(recs.get(1, s + 1))
        f = A0 + BX * i[None, None, :] + CZ * zk[:, None, None] + DT * ex[s]
        rec["salt"] = np.broadcast_to(f, (g.N, g.M, g.L)).copy()
        rec["temp"] = 2.0 * rec["salt"] - 40.0
        slot_recs.append(rec)
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(2.0, g.L - 2.0, n), rng.uniform(2.0, g.M - 2.0, n)
    z = rng.uniform(zk[0] + 0.01, zk[-1] - 0.01, n)
    z[:10] = zk[-1] + 0.3                          # above the top rho level
    return cfg, slot_recs, ex, x, y, z, zk


def test_affine_tracer_is_reproduced():
    cfg, slot_recs, ex, x, y, z, zk = tracer_case()
    og = O.OracleGrid(cfg.grids[0])
    for s, rec in enumerate(slot_recs):
        og.set_slot(s + 1, rec)
    ix = (0.0, 600.0, 720.0)
    salt, temp = O.sample_tracers([og], ex, ix, 1, 2, 3, x, y, z, np.ones(len(x), np.int32))
    zz = np.where(z > zk[-1], zk[-1], z)
    want = A0 + BX * x + CZ * zz + DT * ix[1]
    assert np.abs(salt - want).max() < 1e-9
    assert np.abs(temp - (2.0 * want - 40.0)).max() < 1e-8


def test_per_step_state_and_mean():
    """rp_tracers_* restated: release-time values at t_c for the particles in a
    grid, per-step values only for the particles that stepped, and the
    SaltTempMean sums."""
    import types
    cfg, slot_recs, ex, x, y, z, zk = tracer_case()
    og = O.OracleGrid(cfg.grids[0])
    for s, rec in enumerate(slot_recs):
        og.set_slot(s + 1, rec)
    n = len(x)
    rng = np.random.default_rng(4)
    isin = rng.random(n) < 0.8
    st = types.SimpleNamespace(n=n, x=x, y=y, z=z, gid=np.ones(n, np.int32), isin=isin,
                               age=rng.uniform(0, 1e4, n), bath=np.full(n, 50.0))
    zz = np.where(z > zk[-1], zk[-1], z)
    la, ps, pt, ms, mt, mI = O.tracers_begin([og], 2, st, mean=True)
    want0 = A0 + BX * x + CZ * zz + DT * ex[1]
    assert mI == 1 and np.array_equal(la, st.age)
    assert np.abs(ps[isin] - want0[isin]).max() < 1e-9 and np.all(ps[~isin] == 0.0)
    assert np.array_equal(ms, ps) and np.abs(pt[isin] - (2 * want0[isin] - 40)).max() < 1e-8
    step = types.SimpleNamespace(ex=ex, ix=(0.0, 600.0, 720.0), tb=1, tc=2, tf=3)
    act = np.arange(n) % 3 != 0
    st.age = st.age + np.where(act, 120.0, 0.0)
    p0, m0 = ps.copy(), ms.copy()
    O.tracers_update([og], step, st, la, ps, pt, ms, mt)
    want1 = A0 + BX * x + CZ * zz + DT * 600.0
    assert np.abs(ps[act] - want1[act]).max() < 1e-9 and np.array_equal(ps[~act], p0[~act])
    assert np.array_equal(ms[act], m0[act] + ps[act]) and np.array_equal(ms[~act], m0[~act])
    assert np.array_equal(la, st.age)
