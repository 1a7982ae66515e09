"""The reference built with -DWETDRY, in the oracle: the wet/dry masks of
each hydro slot (hydrodynamic_module.f90:510-570) take the place of the
static masks in getInterp3d's u, v (dry corners zeroed), w, salt, temp
(weights renormalised over the wet corners) and getInterpStr's bustr, bvstr
(interpolation_module.f90:628-923, 1022-1140); getInterpAKs keeps the static
mask (:1175-1229).

Pinned three ways: hand-computed values for a corner that is dry in one slot
only (the other slots interpolate all four corners); substitution, i.e. each
slot's value equals the default build's on a grid whose static masks are
that slot's wet masks; and a full step whose wet masks equal the static
masks in every slot, which must equal the default build bit for bit.
No GPU."""
import ctypes as C

import numpy as np

from oracle import oracle as O
from romspath_amd import synth

L = O.lib()


def _grid(N=6, L_=10, M=9, mask=None):
    return synth.Grid(L_, M, N, 1000.0, 1000.0, 0.0, 0.0, h=np.full((M, L_), 30.0),
                      mask_rho=np.ones((M, L_), np.int32) if mask is None else mask)


def _uv_masks(m):
    return (np.ascontiguousarray(m[:, :-1] * m[:, 1:], dtype=np.int32),
            np.ascontiguousarray(m[:-1, :] * m[1:, :], dtype=np.int32))


def _record(g, u, v, w, wet=None, salt=None, bu=None, bv=None):
    rec = dict(zeta=np.zeros((g.M, g.L)), u=u, v=v, w=w, aks=np.full((g.N + 1, g.M, g.L), 1e-3))
    if salt is not None:
        rec["salt"] = salt
        rec["temp"] = 2.0 * salt
    if bu is not None:
        rec["bustr"], rec["bvstr"] = bu, bv
    if wet is not None:
        mu, mv = _uv_masks(wet)
        rec["wetdry_mask_rho"], rec["wetdry_mask_u"], rec["wetdry_mask_v"] = wet, mu, mv
    return rec


def _u3(og, slot, X, Y, Z=-15.0, var=0):
    return L.og_interp3d(C.byref(og.s), var, X, Y, Z, slot, 1 if var != 2 else 2, 0.0, -30.0)


def test_u_corner_dry_in_one_slot_only():
    g = _grid()
    u = np.full((g.N, g.M, g.L - 1), 2.0)
    u[:, 4, 5] = 7.0                               # u point (6, 5) (1-based i, j)
    v = np.zeros((g.N, g.M - 1, g.L))
    w = np.zeros((g.N + 1, g.M, g.L))
    dry = np.ones((g.M, g.L), np.int32)
    dry[4, 6] = 0                                  # rho (7, 5) dry -> u point (6, 5) dry
    og = O.OracleGrid(g)
    og.set_slot(1, _record(g, u, v, w, wet=np.ones((g.M, g.L), np.int32)))
    og.set_slot(2, _record(g, u, v, w, wet=dry))
    og.set_slot(3, _record(g, u, v, w))            # slot without masks: the static ones
    X, Y = 5.8, 4.6                                # uvI 5.3: Inode 5, x 0.3; Jnode 4, y 0.6
    x, y = 0.3, 0.6
    c = [(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y]
    full = c[0] * 2.0 + c[1] * 2.0 + c[2] * 7.0 + c[3] * 2.0
    zeroed = c[0] * 2.0 + c[1] * 2.0 + c[3] * 2.0   # corner (6, 5) zeroed, weights kept
    assert abs(_u3(og, 1, X, Y) - full) < 1e-14
    assert abs(_u3(og, 2, X, Y) - zeroed) < 1e-14
    assert abs(_u3(og, 3, X, Y) - full) < 1e-14
    # a cell dry in all four corners in slot 2 gives 0
    dry2 = np.ones((g.M, g.L), np.int32)
    dry2[3:5, 5:8] = 0
    og.set_slot(2, _record(g, u, v, w, wet=dry2))
    assert _u3(og, 2, X, Y) == 0.0
    assert abs(_u3(og, 1, X, Y) - full) < 1e-14


def test_w_and_salt_renormalise_over_the_slots_wet_corners():
    g = _grid()
    rng = np.random.default_rng(3)
    w = rng.uniform(-1e-3, 1e-3, (g.N + 1, g.M, g.L))
    salt = rng.uniform(30.0, 35.0, (g.N, g.M, g.L))
    u = np.zeros((g.N, g.M, g.L - 1))
    v = np.zeros((g.N, g.M - 1, g.L))
    dry = np.ones((g.M, g.L), np.int32)
    dry[4, 5] = 0                                  # rho (6, 5)
    og = O.OracleGrid(g)
    og.set_slot(1, _record(g, u, v, w, wet=np.ones((g.M, g.L), np.int32), salt=salt))
    og.set_slot(2, _record(g, u, v, w, wet=dry, salt=salt))
    X, Y = 5.25, 4.75                              # corners (5,4) (6,4) (6,5) (5,5)
    x, y = 0.25, 0.75
    m = np.array([1.0, 1.0, 0.0, 1.0])
    cff = np.zeros(4)
    L.og_getcoeff(x, y, O.dptr(m), O.dptr(cff))
    assert cff[2] == 0.0 and abs(cff.sum() - 1.0) < 1e-15
    bil = np.array([(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y])
    corners = [(4, 3), (5, 3), (5, 4), (4, 4)]     # 0-based (i, j) of the four rho corners
    Z = -15.0
    for slot, wts in ((1, bil), (2, cff)):
        # w at the w level bracket of Z and salt at the rho level bracket:
        # check against the same-level blend of the corners with these weights
        for var, f, nz in ((2, w, g.N + 1), (3, salt, g.N)):
            got = _u3(og, slot, X, Y, Z, var)
            ref = O.OracleGrid(_grid(mask=np.ones((g.M, g.L), np.int32) if slot == 1 else dry))
            ref.set_slot(1, _record(g, u, v, w, salt=salt))
            want = _u3(ref, 1, X, Y, Z, var)
            assert got == want, (slot, var, got, want)
            # the weights act on the corners: a field equal to 1 at the dry
            # corner and 0 elsewhere gives exactly that corner's weight
            spike = np.zeros_like(f)
            spike[:, corners[2][1], corners[2][0]] = 1.0
            og2 = O.OracleGrid(g)
            rec = _record(g, u, v, spike if var == 2 else w, wet=dry if slot == 2 else np.ones((g.M, g.L), np.int32),
                          salt=spike if var == 3 else salt)
            og2.set_slot(1, rec)
            assert abs(_u3(og2, 1, X, Y, Z, var) - wts[2]) < 1e-15


def test_each_slot_equals_the_default_build_with_its_masks_as_static():
    """Substitution over random masks, positions and every masked field."""
    rng = np.random.default_rng(11)
    g0 = _grid(N=5, L_=12, M=10)
    u = rng.standard_normal((g0.N, g0.M, g0.L - 1))
    v = rng.standard_normal((g0.N, g0.M - 1, g0.L))
    w = rng.standard_normal((g0.N + 1, g0.M, g0.L))
    salt = rng.uniform(30, 35, (g0.N, g0.M, g0.L))
    bu = 0.1 * rng.standard_normal((g0.M, g0.L - 1))
    bv = 0.1 * rng.standard_normal((g0.M - 1, g0.L))
    wets = [(rng.uniform(size=(g0.M, g0.L)) > 0.25).astype(np.int32) for _ in range(3)]
    og = O.OracleGrid(g0)
    for s in range(3):
        og.set_slot(s + 1, _record(g0, u, v, w, wet=wets[s], salt=salt, bu=bu, bv=bv))
    refs = []
    for s in range(3):
        r = O.OracleGrid(_grid(N=5, L_=12, M=10, mask=wets[s]))
        r.set_slot(1, _record(g0, u, v, w, salt=salt, bu=bu, bv=bv))
        refs.append(r)
    n_masked = 0
    for X, Y, Z in zip(rng.uniform(1.0, g0.L - 0.01, 400), rng.uniform(1.0, g0.M - 0.01, 400),
                       rng.uniform(-29.0, -1.0, 400)):
        for s in range(3):
            for var in (0, 1, 2, 3, 4):
                got = _u3(og, s + 1, X, Y, Z, var)
                want = _u3(refs[s], 1, X, Y, Z, var)
                assert got == want, (s, var, X, Y, Z)
            for var in (0, 1):
                got = L.og_interp_str(C.byref(og.s), var, s + 1, X, Y)
                want = L.og_interp_str(C.byref(refs[s].s), var, 1, X, Y)
                assert got == want
        n_masked += int(any(_u3(og, s + 1, X, Y, Z, 0) != _u3(og, 1 + (s + 1) % 3, X, Y, Z, 0)
                            for s in range(3)))
    assert n_masked > 50       # the slots' masks do differ where the particles are


def test_vertical_walk_aks_keeps_the_static_mask():
    g = _grid()
    rng = np.random.default_rng(5)
    aks = rng.uniform(1e-4, 1e-2, (g.N + 1, g.M, g.L))
    u = np.zeros((g.N, g.M, g.L - 1))
    v = np.zeros((g.N, g.M - 1, g.L))
    w = np.zeros((g.N + 1, g.M, g.L))
    og = O.OracleGrid(g)
    rec = _record(g, u, v, w, wet=np.zeros((g.M, g.L), np.int32))   # all dry
    rec["aks"] = aks
    og.set_slot(1, rec)
    plain = O.OracleGrid(g)
    rec2 = _record(g, u, v, w)
    rec2["aks"] = aks
    plain.set_slot(1, rec2)
    for X, Y in zip(rng.uniform(1, g.L - 0.1, 50), rng.uniform(1, g.M - 0.1, 50)):
        for k in range(1, g.N + 2):
            a = L.og_interp_aks(C.byref(og.s), X, Y, 1, k)
            b = L.og_interp_aks(C.byref(plain.s), X, Y, 1, k)
            assert a == b and a > 0.0


class _WetRecords:
    """A record source whose records carry wet/dry masks: mask_of(grid, rec)."""

    def __init__(self, base, mask_of):
        self.base = base
        self.mask_of = mask_of

    def get(self, ng, record, filenum=None):
        rec = dict(self.base.get(ng, record, filenum))
        wet = self.mask_of(ng, record)
        mu, mv = _uv_masks(wet)
        rec["wetdry_mask_rho"], rec["wetdry_mask_u"], rec["wetdry_mask_v"] = wet, mu, mv
        return rec


def test_step_with_static_masks_as_wet_masks_equals_default_build():
    from romspath_amd.runner import RecordSource
    cfg = synth.config_mini(n_particles=300, days=40 * 120 / 86400.0, seed=9)
    base = RecordSource(cfg)
    plain = O.run_oracle(cfg, max_steps=40, records=base)
    wet = O.run_oracle(cfg, max_steps=40,
                       records=_WetRecords(RecordSource(cfg), lambda ng, r: cfg.grids[ng - 1].mask_rho))
    for k in ("x", "y", "z", "status"):
        assert np.array_equal(getattr(plain, k), getattr(wet, k)), k
