"""The FP32-field vertical walk's two forms across resampling sizes: the
eigenbasis form where n = np - 2 is even and 64 <= n <= 128, the band form
elsewhere (and under RP_VWALK_FORM=band).  np = int((N+1) * sub)
(ver_turb_module.f90:52): N 16 gives n 66 (eigenbasis, 31 padded modes per
symmetry half), N 31 gives n 126 (eigenbasis, one padded mode), N 15 gives
n 62 and N 32 gives n 130 (band), N 30 with sub 4.1 gives n 125 (odd: band)
and with sub 3.3 gives n 100 (eigenbasis).  Three more cases on n 122 take
the smoothing weight's ends: smth 0 (smooth's p = 1 branch, the
interpolating spline), smth 30 (p near 1e-10: the knot values rebuilt from
their second differences) and smth 1e6 (the p = 0 branch: the least-squares
line of the column, pppack.f90:5236-5298).

Each case steps 256 particles on the island grid with the vertical walk and
injected deviates for 30 steps and is compared with the FP64 oracle: the
largest position error within 1e-4 cell over the particles whose status
agrees (at least 99% of them).  The forced band form must give the same
results as the default where the eigenbasis does not apply, and different
(also within 1e-4 cell of the oracle) where it does, which shows that the
eigenbasis form ran.  At p = 0 the knots are the line whatever u holds, so
there the two forms must agree exactly.
"""
import dataclasses
import os

import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun

pytestmark = pytest.mark.gpu

CASES = [(16, 4.0, True), (31, 4.0, True), (15, 4.0, False), (32, 4.0, False), (30, 4.1, False),
         (30, 3.3, True)]


def _gpu(cfg, steps, form=None):
    old = os.environ.get("RP_VWALK_FORM")
    if form is None:
        os.environ.pop("RP_VWALK_FORM", None)
    else:
        os.environ["RP_VWALK_FORM"] = form
    try:
        run = GpuRun(cfg, precision="fp32")
    finally:
        if old is None:
            os.environ.pop("RP_VWALK_FORM", None)
        else:
            os.environ["RP_VWALK_FORM"] = old
    try:
        out = run.run(max_steps=steps, inject=True, diag_every_step=True)
        capped = run.ctx.secant_capped()
    finally:
        run.close()
    return out, capped


def _check(g, o, cfg, what):
    layer = np.maximum(o.bath, 1.0) / cfg.grids[0].N
    ex, ey, ez = np.abs(g["x"] - o.x), np.abs(g["y"] - o.y), np.abs(g["z"] - o.z) / layer
    same = g["status"] == o.status
    print(f"{what}: max err x {ex[same].max():.2e} y {ey[same].max():.2e} z {ez[same].max():.2e}, "
          f"status agree {same.mean():.4f}", flush=True)
    assert same.mean() > 0.99
    assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


@pytest.mark.parametrize("N,sub,eigen,smth", [c + (0.6,) for c in CASES] + [
    (30, 4.0, True, 0.0), (30, 4.0, True, 30.0), (30, 4.0, False, 1e6)])
def test_vertical_walk_forms_match_oracle(N, sub, eigen, smth):
    """eigen: the two forms' results must differ (the eigenbasis form ran and
    its u enters the knots)."""
    cfg = synth.config_mini(n_particles=256, N=N, days=1 / 24.0, dob_spread=600.0)
    cfg = dataclasses.replace(cfg, params=dataclasses.replace(cfg.params, sub=sub, smth=smth))
    np_ = int((N + 1) * sub)
    steps = 30
    print(f"case vwalk forms N={N} sub={sub} np={np_} n={np_ - 2} smth={smth} eigenbasis={eigen}",
          flush=True)
    o = O.run_oracle(cfg, max_steps=steps, inject=True)
    g, capped = _gpu(cfg, steps)
    assert capped == 0
    _check(g, o, cfg, f"N={N} sub={sub} default form")
    b, _ = _gpu(cfg, steps, form="band")
    _check(b, o, cfg, f"N={N} sub={sub} band form")
    same = all(np.array_equal(g[k], b[k]) for k in ("x", "y", "z", "status"))
    assert same != eigen
