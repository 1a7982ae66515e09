"""Parity of the HIP step (librp.so) with the oracle on the same seeded inputs.

FP64 mode: status, grid id and isIn must match exactly and positions to
1e-9 of a cell (x, y in cell units; z in units of the local mean layer
thickness h/N).  FP32-field mode: the largest position error against the
FP64 oracle within 1e-4 of a cell, over the particles whose status (and grid)
agree, which must be at least 99% of them.  Random walks use identical
injected deviates.
"""
import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun

pytestmark = pytest.mark.gpu


def _run_both(cfg, steps, precision, inject, sort_interval=0):
    print(f"case {cfg.name} n={cfg.particles.n} steps={steps} {precision} sort={sort_interval}",
          flush=True)
    run = GpuRun(cfg, precision=precision, sort_interval=sort_interval)
    g = run.run(max_steps=steps, inject=inject, diag_every_step=True)
    act = run.ctx.active_steps()
    capped = run.ctx.secant_capped()
    run.close()
    o = O.run_oracle(cfg, max_steps=steps, inject=inject)
    return g, o, act, capped


def _layer(cfg, o):
    return np.maximum(o.bath, 1.0) / cfg.grids[0].N


def _check_exact_discrete(g, o):
    assert np.array_equal(g["status"], o.status)
    assert np.array_equal(g["gid"], o.gid)
    assert np.array_equal(g["isin"], o.isin)


def _err(g, o, cfg):
    ex = np.abs(g["x"] - o.x)
    ey = np.abs(g["y"] - o.y)
    ez = np.abs(g["z"] - o.z) / _layer(cfg, o)
    return ex, ey, ez


def test_c1_fp64_matches_oracle():
    cfg = synth.config1(n_particles=3000)
    g, o, act, _ = _run_both(cfg, 240, "fp64", inject=False)
    _check_exact_discrete(g, o)
    ex, ey, ez = _err(g, o, cfg)
    print("c1 fp64 max err cells", ex.max(), ey.max(), ez.max(), flush=True)
    assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
    assert np.array_equal(g["age"], o.age)
    assert act == o.active


def test_mini_fp64_turbulence_injected_matches_oracle():
    cfg = synth.config_mini(n_particles=400, days=2 / 24.0)
    g, o, act, capped = _run_both(cfg, 60, "fp64", inject=True)
    _check_exact_discrete(g, o)
    ex, ey, ez = _err(g, o, cfg)
    print("mini fp64 max err", ex.max(), ey.max(), ez.max(), "capped", capped, flush=True)
    assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
    m = np.isin(o.status, (1.0, 1.1, 1.2, -8.0, -9.0))
    for k in ("wd", "zeta", "bath"):
        assert np.max(np.abs(g[k][m] - getattr(o, k)[m])) <= 1e-9
    assert act == o.active
    assert capped == 0


def test_c1_fp32_fields_error_bound():
    cfg = synth.config1(n_particles=3000)
    g, o, _, _ = _run_both(cfg, 240, "fp32", inject=False)
    ex, ey, ez = _err(g, o, cfg)
    print("c1 fp32 err max/p95", ex.max(), np.percentile(ex, 95), ez.max(), flush=True)
    same = g["status"] == o.status
    assert same.mean() > 0.999
    assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


def test_mini_fp32_turbulence_error():
    cfg = synth.config_mini(n_particles=400, days=2 / 24.0)
    g, o, _, _ = _run_both(cfg, 60, "fp32", inject=True)
    ex, ey, ez = _err(g, o, cfg)
    same = g["status"] == o.status
    print("mini fp32 err max", ex[same].max(), ey[same].max(), ez[same].max(),
          "status agree", same.mean(), flush=True)
    assert same.mean() > 0.99
    assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


def test_rng_matches_philox_reference():
    from rng_ref import normals
    cfg = synth.config_mini(n_particles=64)
    run = GpuRun(cfg, precision="fp64")
    got = run.ctx.sample_deviates(5, 64, 12)
    run.close()
    want = np.stack([normals(cfg.params.seed, p, 5, 12) for p in range(64)])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_error_codes_not_crashes():
    from romspath_amd.abi import RpError
    cfg = synth.config1(n_particles=10)
    run = GpuRun(cfg, precision="fp64")
    with pytest.raises(RpError) as e:
        run.ctx.step((0.0, 1.0, 2.0), (0.0, 1.0, 2.0), 1, 1, 2, 3, 0)
    assert e.value.code == -2            # slots not loaded yet
    with pytest.raises(RpError) as e:
        run.ctx.step((0.0, 1.0, 2.0), (0.0, 1.0, 2.0), 1, 0, 2, 3, 0)
    assert e.value.code == -1
    run.close()


def test_nested_fp64_handoff_matches_oracle():
    cfg = synth.config_mini_nested(n_particles=600)
    g, o, act, capped = _run_both(cfg, 60, "fp64", inject=True)
    _check_exact_discrete(g, o)
    ex, ey, ez = _err(g, o, cfg)
    print("nested fp64 max err", ex.max(), ey.max(), ez.max(), "gid counts", np.bincount(o.gid),
          flush=True)
    assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
    assert act == o.active and capped == 0
    # the case really hands particles over in both directions
    assert np.any(o.gid != cfg.particles.gid)


def test_nested_fp32_fields_error():
    cfg = synth.config_mini_nested(n_particles=600)
    g, o, _, _ = _run_both(cfg, 60, "fp32", inject=True)
    same = (g["status"] == o.status) & (g["gid"] == o.gid)
    ex, ey, ez = _err(g, o, cfg)
    print("nested fp32 err max", ex[same].max(), ey[same].max(), ez[same].max(),
          "agree", same.mean(), flush=True)
    assert same.mean() > 0.99
    assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_thirty_levels_spline_path(precision):
    # N=30, sub=4: 124 spline knots per column, the level count of the large grids
    cfg = synth.config_mini(n_particles=256, N=30, days=1 / 24.0, dob_spread=600.0)
    g, o, act, capped = _run_both(cfg, 30, precision, inject=True)
    ex, ey, ez = _err(g, o, cfg)
    print("N30", precision, "max err", ex.max(), ez.max(), "p95 z", np.percentile(ez, 95), flush=True)
    if precision == "fp64":
        _check_exact_discrete(g, o)
        assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
        assert act == o.active and capped == 0
    else:
        same = g["status"] == o.status
        assert same.mean() > 0.99
        assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


@pytest.mark.parametrize("case", ["c1", "mini", "nested"])
def test_cell_sorted_order_gives_identical_results(case):
    # re-sorting the stored columns by cell changes which lane steps which
    # particle, never the result: FP64 mode stays exact against the oracle
    if case == "c1":
        cfg, steps, inject, every = synth.config1(n_particles=3000), 120, False, 7
    elif case == "mini":
        cfg, steps, inject, every = synth.config_mini(n_particles=400, days=2 / 24.0), 60, True, 3
    else:
        cfg, steps, inject, every = synth.config_mini_nested(n_particles=600), 60, True, 1
    g, o, act, _ = _run_both(cfg, steps, "fp64", inject, sort_interval=every)
    _check_exact_discrete(g, o)
    ex, ey, ez = _err(g, o, cfg)
    print(case, "sorted fp64 max err", ex.max(), ey.max(), ez.max(), flush=True)
    assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
    assert np.array_equal(g["age"], o.age)
    assert act == o.active
