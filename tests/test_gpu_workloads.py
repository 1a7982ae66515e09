"""The benchmark workloads through the HIP library against the oracle.

* C2: the 400 x 300 x 30 shelf-and-island grid with eddies and tide, a seeded
  10,000-particle subsample of the 1M release, 150 steps (external steps
  1-5: the first-step slot rule, the lagged rotation and the steady
  (k-2, k-1, k) pattern, with a record load at every boundary).
* C3: the nested 600 x 500 x 30 parent and 300 x 300 x 30 child with Stokes
  drift, a 20,000-particle subsample of the 4M release (both grids), 75
  steps across two external boundaries, hand-offs both ways.
* C4: the scenarios of the 256-scenario draw released in the first 2.5 h,
  500 particles of each, 90 steps: per-scenario launches with injected
  deviates equal the oracle; the batched HIP-graph form equals them with
  counter-based deviates.
* C5: the adjoint on the C2 grid with FP32 fields and the C5 loss, against
  central differences of the same recorded forward in release x, y, s and
  three velocity-perturbation coefficients.
* The well-mixed condition (Visser 1997) for the FP32-field vertical walk.

FP64 mode: status, grid id and isIn identical, positions within 1e-9 cell
(z in layers).  FP32-field mode: max error within 1e-4 cell over the
particles whose status and grid agree (at least 99.9% of them).
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from romspath_amd import synth
from romspath_amd.runner import GpuRun, RecordSource

pytestmark = pytest.mark.gpu


def _subsample(cfg, idx):
    p = cfg.particles
    sub = synth.Particles(p.x[idx].copy(), p.y[idx].copy(), p.z[idx].copy(), p.dob[idx].copy(),
                          gid=p.gid[idx].copy(), status=p.status[idx].copy(),
                          age=p.age[idx].copy(), isin=p.isin[idx].copy())
    return synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, sub, cfg.refine)


@functools.lru_cache(maxsize=None)
def _case(name):
    if name == "C2":
        cfg = synth.config2(n_particles=1_000_000)
        return _subsample(cfg, np.arange(0, 1_000_000, 100)), 150
    if name == "C3":
        cfg = synth.config3(n_particles=4_000_000)
        return _subsample(cfg, np.arange(0, 4_000_000, 200)), 75
    raise KeyError(name)


@functools.lru_cache(maxsize=None)
def _records(name):
    return RecordSource(_case(name)[0])


@functools.lru_cache(maxsize=None)
def _oracle(name):
    cfg, steps = _case(name)
    o = O.run_oracle(cfg, max_steps=steps, inject=True, records=_records(name))
    print(f"oracle {name}: {cfg.particles.n} particles x {steps} steps, "
          f"{o.step_seconds:.1f} s", flush=True)
    return o


def _layer(o, cfg):
    return np.maximum(o.bath, 1.0) / cfg.grids[0].N


@pytest.mark.parametrize("name", ["C2", "C3"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_workload_window_matches_oracle(name, precision):
    cfg, steps = _case(name)
    print(f"case {name} {cfg.particles.n} particles x {steps} steps {precision} sort 16", flush=True)
    run = GpuRun(cfg, precision=precision, records=_records(name), sort_interval=16)
    try:
        g = run.run(max_steps=steps, inject=True, diag_every_step=True)
        act = run.ctx.active_steps()
        capped = run.ctx.secant_capped()
    finally:
        run.close()
    o = _oracle(name)
    ex, ey = np.abs(g["x"] - o.x), np.abs(g["y"] - o.y)
    ez = np.abs(g["z"] - o.z) / _layer(o, cfg)
    if name == "C3":
        assert set(np.unique(o.gid)) == {1, 2} and np.any(o.gid != cfg.particles.gid)
    if precision == "fp64":
        assert np.array_equal(g["status"], o.status)
        assert np.array_equal(g["gid"], o.gid)
        assert np.array_equal(g["isin"], o.isin)
        print(f"{name} fp64 max err {ex.max():.2e} {ey.max():.2e} {ez.max():.2e}", flush=True)
        assert ex.max() <= 1e-9 and ey.max() <= 1e-9 and ez.max() <= 1e-9
        assert np.array_equal(g["age"], o.age)
        assert act == o.active and capped == 0
    else:
        same = (g["status"] == o.status) & (g["gid"] == o.gid)
        print(f"{name} fp32 max err {ex[same].max():.2e} {ey[same].max():.2e} {ez[same].max():.2e}"
              f" agree {same.mean():.5f}", flush=True)
        assert same.mean() >= 0.999
        assert ex[same].max() <= 1e-4 and ey[same].max() <= 1e-4 and ez[same].max() <= 1e-4


# ---------------------------------------------------------------------------
# C4: release ensembles
# ---------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def _c4_subset():
    cfg, counts = synth.config4()
    starts = np.r_[0, np.cumsum(counts)[:-1]]
    dob0 = cfg.particles.dob[starts]
    keep = np.nonzero(dob0 < 2.5 * 3600.0)[0]
    idx = np.concatenate([np.arange(starts[s], starts[s] + min(int(counts[s]), 500)) for s in keep])
    sub_counts = [min(int(counts[s]), 500) for s in keep]
    return _subsample(cfg, idx), sub_counts


def test_c4_subset_scenarios_batch_and_oracle():
    cfg, counts = _c4_subset()
    steps = 90
    recs = RecordSource(cfg)
    print(f"case C4 subset: {len(counts)} scenarios, {cfg.particles.n} particles, {steps} steps",
          flush=True)
    # per-scenario launches, injected deviates, FP64: equal to the oracle
    run = GpuRun(cfg, precision="fp64", records=recs, sort_interval=8)
    try:
        g = run.run(max_steps=steps, inject=True, mode="scenario", counts=counts,
                    diag_every_step=True)
        act = run.ctx.active_steps()
    finally:
        run.close()
    o = O.run_oracle(cfg, max_steps=steps, inject=True, records=recs)
    assert o.active > 0.3 * cfg.particles.n * steps / 2
    assert np.array_equal(g["status"], o.status) and np.array_equal(g["isin"], o.isin)
    assert np.abs(g["x"] - o.x).max() <= 1e-9 and np.abs(g["y"] - o.y).max() <= 1e-9
    assert (np.abs(g["z"] - o.z) / _layer(o, cfg)).max() <= 1e-9
    assert act == o.active
    # counter-based deviates: the batched graph form equals the per-scenario
    # launches bit for bit, in both field precisions
    for precision in ("fp64", "fp32"):
        outs = []
        for mode in ("scenario", "batch"):
            r = GpuRun(cfg, precision=precision, records=recs, sort_interval=8)
            try:
                outs.append(r.run(max_steps=steps, mode=mode, counts=counts, use_graph=True))
            finally:
                r.close()
        for k in ("x", "y", "z", "age", "status", "gid", "isin"):
            assert np.array_equal(outs[0][k], outs[1][k]), (precision, k)


# ---------------------------------------------------------------------------
# C5: the adjoint on the C2 grid, FP32 fields
# ---------------------------------------------------------------------------
def test_c5_adjoint_fp32_fields_matches_finite_differences():
    from scipy.special import ndtr
    full = synth.config2(n_particles=262_144, seed=5, days=2.0)
    cfg = _subsample(full, np.arange(0, 262_144, 128))          # 2048 particles
    n = cfg.particles.n
    steps, ckpt = 128, 64
    g = cfg.grids[0]
    recs = RecordSource(cfg)
    rng = np.random.default_rng(55)
    tx, ty = rng.uniform(5.0, g.L - 5.0, n), rng.uniform(5.0, g.M - 5.0, n)
    bx0, by0 = rng.uniform(20.0, g.L - 40.0), rng.uniform(20.0, g.M - 40.0)
    sig, beta = 2.0, 10.0
    coef = 0.02 * np.random.default_rng(56).standard_normal((2, 16, 16))

    def soft(u, lo):
        a, b = (u - lo) / sig, (u - lo - 20.0) / sig
        return ndtr(a) - ndtr(b), (np.exp(-0.5 * a * a) - np.exp(-0.5 * b * b)) / (sig * np.sqrt(2 * np.pi))

    tz = -rng.uniform(2.0, 18.0, n)

    def loss(o):
        # the C5 loss plus the squared depth error in units of 10 m: the C2
        # currents are depth-uniform, so without it the loss would not
        # depend on the release depth at all
        dx, dy, dz = o["x"] - tx, o["y"] - ty, (o["z"] - tz) / 10.0
        px, dpx = soft(o["x"], bx0)
        py, dpy = soft(o["y"], by0)
        return (dx * dx + dy * dy + dz * dz - beta * px * py,
                (2 * dx - beta * dpx * py, 2 * dy - beta * px * dpy, 2 * dz / 10.0))

    def forward(parts=None, c=coef):
        c2 = cfg if parts is None else synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params,
                                                   parts, cfg.refine)
        run = GpuRun(c2, precision="fp32", records=recs)
        run.ctx.set_perturbation(c)
        run.ctx.adjoint_record(ckpt)       # the recorded forward is the map differentiated
        return run, run.run(max_steps=steps)

    print(f"case C5 adjoint fp32 fields, {n} particles x {steps} steps, ckpt {ckpt}", flush=True)
    run, out = forward()
    L, lam = loss(out)
    (gx, gy, gz), gp = run.ctx.adjoint(lam, want_pert=True, pert_shape=coef.shape)
    run.close()
    P = cfg.particles
    # release s: z0 = s h(x0, y0) (zeta 0 at release); dL/ds = dL/dz0 h
    hh = np.array([O.interp2d_rho(g.h, x, y) for x, y in zip(P.x, P.y)])
    s0 = P.z / hh
    grads = {"x": gx, "y": gy, "s": gz * hh}
    for coord in ("x", "y", "s"):
        fds, ok = [], out["status"] == out["status"]
        for eps in (1e-4, 1e-5):
            vals = []
            for sgn in (1.0, -1.0):
                X, Y, Z = P.x.copy(), P.y.copy(), P.z.copy()
                if coord == "x":
                    X += sgn * eps
                elif coord == "y":
                    Y += sgn * eps
                else:
                    Z = (s0 + sgn * eps) * hh
                r, o = forward(synth.Particles(X, Y, Z, P.dob.copy()))
                r.close()
                ok &= o["status"] == out["status"]
                vals.append(loss(o)[0])
            fds.append((vals[0] - vals[1]) / (2 * eps))
        floor = 1e-3 * np.median(np.abs(fds[0][ok]))
        trusted = ok & (np.abs(fds[0] - fds[1]) <= 1e-5 * np.maximum(np.abs(fds[0]), floor))
        d = (grads[coord] - fds[1])[trusted]
        nrm = np.linalg.norm(d) / np.linalg.norm(fds[1][trusted])
        print(f"C5 d/d{coord}0: {trusted.sum()}/{ok.sum()} trusted, norm-wise rel {nrm:.2e}",
              flush=True)
        assert trusted.sum() >= 0.9 * ok.sum() and nrm <= 1e-5
    # three perturbation coefficients: the largest of each component and the
    # second largest of u
    picks = [(0, *np.unravel_index(np.argmax(np.abs(gp[0])), gp[0].shape)),
             (1, *np.unravel_index(np.argmax(np.abs(gp[1])), gp[1].shape)),
             (0, *np.unravel_index(np.argsort(np.abs(gp[0]).ravel())[-2], gp[0].shape))]
    for comp, j, i in picks:
        vals = []
        for sgn in (1.0, -1.0):
            c2 = coef.copy()
            c2[comp, j, i] += sgn * 1e-5
            r, o = forward(c=c2)
            r.close()
            vals.append(loss(o)[0].sum())
        fd = (vals[0] - vals[1]) / 2e-5
        rel = abs(gp[comp, j, i] - fd) / max(abs(fd), 1e-12)
        print(f"C5 coef[{comp},{j},{i}]: adjoint {gp[comp, j, i]:.8e} fd {fd:.8e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-5


# ---------------------------------------------------------------------------
# well-mixed condition of the FP32-field vertical walk
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_well_mixed_condition_on_gpu(precision):
    """Uniform particles stay uniform under the random displacement model
    with a sheared diffusivity AKs = 4e-2 sigma (1 - sigma) (Visser 1997;
    ver_turb_module.f90:240-293), counter-based deviates, no advection."""
    L, M, N, h = 10, 10, 12, 20.0
    g = synth.Grid(L, M, N, 1000.0, 1000.0, 0.0, 0.0, h=np.full((M, L), h),
                   mask_rho=np.ones((M, L), np.int32))
    zw = g.z_w(np.zeros((M, L)))
    sig = (zw + h) / h
    aks = 4.0 * 1e-2 * sig * (1 - sig)

    class Rec:
        def record(self, t):
            z = lambda *s: np.zeros(s)
            return dict(zeta=z(M, L), u=z(N, M, L - 1), v=z(N, M - 1, L), w=z(N + 1, M, L),
                        aks=aks.copy())

    n = 20_000
    rng = np.random.default_rng(21)
    parts = synth.Particles(rng.uniform(4.1, 5.9, n), rng.uniform(4.1, 5.9, n),
                            rng.uniform(-h, 0.0, n), np.zeros(n))
    prm = synth.RunParams(dt=3600, idt=120, days=150 * 120 / 86400.0, deltat=12.0, VTurbOn=True,
                          scheme=0, seed=9)
    cfg = synth.Config("well-mixed", [g], [Rec()], prm, parts)
    print(f"case well-mixed {precision} {n} particles x 150 steps", flush=True)
    run = GpuRun(cfg, precision=precision)
    try:
        out = run.run()
    finally:
        run.close()
    z = out["z"]
    assert np.all((z > -h) & (z < 0.0))
    counts, _ = np.histogram(z, bins=8, range=(-h, 0.0))
    exp = n / 8.0
    chi2 = ((counts - exp) ** 2 / exp).sum()
    print(f"well-mixed {precision}: chi2 {chi2:.1f} counts {counts}", flush=True)
    assert chi2 < 24.3, counts       # 7 dof, p = 0.001
