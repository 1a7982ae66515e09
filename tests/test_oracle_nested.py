"""Nested-grid pieces of the oracle: the affine grid maps (grid_module.f90:437-481)
and the grid hand-off in update_particles (ROMSPath.f90:791-822), pinned by closed
forms: the maps of a 3:1 child placed at a known parent point, their composition,
and straight-line tracks in a uniform current that cross the child's edge both ways.
"""
import numpy as np

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.nesting import scale_offsets


class _Uniform:
    def __init__(self, u, v):
        self.u, self.v = u, v

    def uv(self, x, y, t):
        return np.full(np.shape(x), self.u), np.full(np.shape(y), self.v)


def _pair(i0=11, j0=9, refine=3, L=40, M=30, Lc=24, Mc=24, N=6, h=30.0):
    dx = 1000.0
    parent = synth.Grid(L, M, N, dx, dx, 0.0, 0.0, h=np.full((M, L), h),
                        mask_rho=np.ones((M, L), np.int32))
    child = synth.Grid(Lc, Mc, N, dx / refine, dx / refine, (i0 - 1) * dx, (j0 - 1) * dx,
                       h=np.full((Mc, Lc), h), mask_rho=np.ones((Mc, Lc), np.int32), den=refine)
    return child, parent


def test_child_points_coincide_with_parent_points():
    c, p = _pair()
    # every third child point is a parent point, bit for bit
    assert np.array_equal(c.x_rho[0, ::3], p.x_rho[0, 10:10 + 8])
    assert np.array_equal(c.y_rho[::3, 0], p.y_rho[8:8 + 8, 0])


def test_scale_offsets_closed_form():
    i0, j0, r = 11, 9, 3
    c, p = _pair(i0, j0, r)
    scl, off = scale_offsets([c, p], (float(r),))
    # child (1,1) sits on parent (i0,j0): X_p = (X_c - 1)/r + i0
    assert np.allclose(scl[0, 1], 1.0 / r) and np.allclose(off[0, 1], [i0 - 1.0 / r, j0 - 1.0 / r])
    assert np.allclose(scl[1, 0], r) and np.allclose(off[1, 0], [1 - r * i0, 1 - r * j0])
    assert np.allclose(scl[0, 0], 1.0) and np.allclose(off[0, 0], 0.0)
    assert np.allclose(scl[1, 1], 1.0) and np.allclose(off[1, 1], 0.0)
    X = np.linspace(1.0, 24.0, 17)
    back = scl[1, 0, 0] * (scl[0, 1, 0] * X + off[0, 1, 0]) + off[1, 0, 0]
    assert np.allclose(back, X, atol=1e-12)


def _uniform_run(U, V, starts, steps):
    c, p = _pair()
    hc = synth.HydroModel(c, _Uniform(U, V))
    hp = synth.HydroModel(p, _Uniform(U, V))
    X = np.array([s[0] for s in starts], float)
    Y = np.array([s[1] for s in starts], float)
    gid = np.array([s[2] for s in starts], np.int32)
    part = synth.Particles(X, Y, np.full(X.size, -10.0), np.zeros(X.size), gid=gid)
    prm = synth.RunParams(dt=3600, idt=120, days=1.0)
    cfg = synth.Config("nested-uniform", [c, p], [hc, hp], prm, part, refine=(3.0,))
    return cfg, O.run_oracle(cfg, max_steps=steps)


def _metres(gid, X, Y):
    # grid 1 (child): x = 10 km + (X-1)*1000/3; grid 2 (parent): x = (X-1)*1000
    xm = np.where(gid == 1, 10e3 + (X - 1.0) * 1000.0 / 3.0, (X - 1.0) * 1000.0)
    ym = np.where(gid == 1, 8e3 + (Y - 1.0) * 1000.0 / 3.0, (Y - 1.0) * 1000.0)
    return xm, ym


def test_handoff_child_to_parent_and_back_in_uniform_current():
    U, V = 0.5, 0.2
    starts = [(20.0, 12.0, 1),     # in the child, leaves through its east edge
              (8.0, 10.5, 2),      # in the parent west of the child, enters it
              (35.0, 20.0, 2)]     # stays in the parent
    cfg, o = _uniform_run(U, V, starts, 60)
    x0, y0 = _metres(np.array([1, 2, 2]), np.array([s[0] for s in starts]),
                     np.array([s[1] for s in starts]))
    assert np.all(o.status == 1.0)
    assert list(o.gid) == [2, 1, 2]
    xm, ym = _metres(o.gid, o.x, o.y)
    assert np.allclose(xm, x0 + U * o.age, atol=1e-6)
    assert np.allclose(ym, y0 + V * o.age, atol=1e-6)


def test_exit_through_open_boundary_from_child():
    # a current carrying a child particle straight out of both grids through the
    # parent's open boundary ends with status -8 and grid id Ngrid
    cfg, o = _uniform_run(0.0, -2.0, [(12.0, 6.0, 1)], 60)
    assert o.status[0] == -8.0 and o.gid[0] == 2 and o.isin[0] == 0
