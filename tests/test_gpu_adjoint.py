"""Adjoint of the particle step (rp_adjoint_record / rp_adjoint) against
central finite differences of the forward run, in FP64 mode.

Loss: L = sum_i w_i . (x_i, y_i, z_i) at the end, w seeded, so lambda = w.
Central differences use eps = 1e-4 and 1e-5 (cells for x, y, metres for z).
Every particle's final state depends only on its own release, so one pair
of runs with every release shifted by +-eps gives the finite-difference
gradient of all particles at once.  Particles whose status differs between
the shifted runs (a land rejection or an exit in one of them) are left out.
"""
import numpy as np
import pytest

from romspath_amd import synth
from romspath_amd.runner import GpuRun

pytestmark = pytest.mark.gpu

STEPS = 60
CKPT = 16        # 60 steps: three full segments and a partial one


def _case(n=256, seed=11):
    cfg = synth.config_mini(n_particles=n, days=STEPS * 120 / 86400.0, dob_spread=0.0, seed=seed)
    return cfg


def _with_release(cfg, X, Y, Z):
    p = cfg.particles
    parts = synth.Particles(X.copy(), Y.copy(), Z.copy(), p.dob.copy())
    return synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, parts, cfg.refine)


def _forward(cfg, precision="fp64", pert=None, record=0):
    run = GpuRun(cfg, precision=precision)
    if pert is not None:
        run.ctx.set_perturbation(pert)
    if record:
        run.ctx.adjoint_record(record)
    out = run.run(max_steps=STEPS)
    return run, out


def _loss_terms(out, w):
    return w[:, 0] * out["x"] + w[:, 1] * out["y"] + w[:, 2] * out["z"]


def _check(g_adj, fd1, fd2, ok, name, tol=1e-5):
    """fd1, fd2: central differences at eps and eps/10.  The trajectory map is
    piecewise smooth: a shifted pair whose trajectories straddle a kink (a
    cell edge of the bilinear stencils, a level of the vertical brackets, a
    reflection) gives a difference quotient that is not the derivative, and
    where the vertical walk has mixed a particle's depth the derivative is
    tiny and the quotient is FP64 rounding noise.  A quotient is trusted
    where the two step sizes agree; the adjoint must match the trusted ones
    norm-wise, and nearly every particle must be trusted."""
    floor = 1e-3 * np.median(np.abs(fd1[ok]))
    trusted = ok & (np.abs(fd1 - fd2) <= tol * np.maximum(np.abs(fd1), floor))
    d = (g_adj - fd2)[trusted]
    nrm = np.linalg.norm(d) / np.linalg.norm(fd2[trusted])
    rel = np.abs(d) / np.maximum(np.abs(fd2[trusted]), floor)
    print(f"{name}: {trusted.sum()}/{ok.sum()} particles trusted, norm-wise rel {nrm:.2e}, "
          f"per-particle median {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    assert trusted.sum() >= 0.9 * ok.sum()
    assert nrm <= tol


def test_recorded_forward_equals_plain_forward():
    cfg = _case()
    print("case adjoint recorded forward vs plain fp64", flush=True)
    r0, a = _forward(cfg)
    r0.close()
    r1, b = _forward(cfg, record=CKPT)
    # the reverse sweep leaves the state as it was
    w = np.ones((cfg.particles.n, 3))
    r1.ctx.adjoint((w[:, 0], w[:, 1], w[:, 2]))
    c = r1.ctx.get_particles()
    r1.close()
    for k in ("x", "y", "z", "status", "gid"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(b[k], c[k]), k


@pytest.mark.parametrize("coord", ["x", "y", "z"])
def test_release_gradient_matches_finite_differences(coord):
    cfg = _case()
    n = cfg.particles.n
    rng = np.random.default_rng(5)
    w = rng.standard_normal((n, 3))
    w[:, 2] *= 0.1
    print(f"case adjoint release gradient d/d{coord}, {n} particles, {STEPS} steps", flush=True)
    run, out = _forward(cfg, record=CKPT)
    (gx, gy, gz), _ = run.ctx.adjoint((w[:, 0], w[:, 1], w[:, 2]))
    run.close()
    g = {"x": gx, "y": gy, "z": gz}[coord]
    P = cfg.particles
    X, Y, Z = P.x.copy(), P.y.copy(), P.z.copy()
    fds = []
    ok = np.ones(n, dtype=bool)
    for eps in (1e-4, 1e-5):     # cells for x, y; metres for z
        outs = []
        for sgn in (1.0, -1.0):
            Xs, Ys, Zs = X.copy(), Y.copy(), Z.copy()
            {"x": Xs, "y": Ys, "z": Zs}[coord][:] += sgn * eps
            r, o = _forward(_with_release(cfg, Xs, Ys, Zs))
            r.close()
            outs.append(o)
            ok &= o["status"] == out["status"]
        fds.append((_loss_terms(outs[0], w) - _loss_terms(outs[1], w)) / (2 * eps))
    _check(g, fds[0], fds[1], ok, f"d/d{coord}")


def test_perturbation_gradient_matches_finite_differences():
    cfg = _case()
    n = cfg.particles.n
    rng = np.random.default_rng(6)
    w = rng.standard_normal((n, 3))
    w[:, 2] *= 0.1
    coef = 0.02 * rng.standard_normal((2, 16, 16))
    print("case adjoint perturbation gradient, 16x16 net", flush=True)
    run, _ = _forward(cfg, pert=coef, record=CKPT)
    _, gp = run.ctx.adjoint((w[:, 0], w[:, 1], w[:, 2]), want_pert=True, pert_shape=coef.shape)
    run.close()
    # the most influential coefficients of each component
    picks = [tuple(np.unravel_index(np.argmax(np.abs(gp[k])), gp[k].shape)) for k in (0, 1)]
    picks.append(tuple(np.unravel_index(np.argsort(np.abs(gp[0]).ravel())[-2], gp[0].shape)))
    comps = [0, 1, 0]
    eps = 1e-5       # m/s
    for comp, (j, i) in zip(comps, picks):
        vals = []
        for sgn in (1.0, -1.0):
            c2 = coef.copy()
            c2[comp, j, i] += sgn * eps
            r, o = _forward(cfg, pert=c2)
            r.close()
            vals.append(_loss_terms(o, w).sum())
        fd = (vals[0] - vals[1]) / (2 * eps)
        rel = abs(gp[comp, j, i] - fd) / max(abs(fd), 1e-12)
        print(f"coef[{comp},{j},{i}]: adjoint {gp[comp, j, i]:.10e} fd {fd:.10e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-5


# ---------------------------------------------------------------------------
# Tangent-linear checks.  The Jacobian of one step at the current state is
# the step recorded alone and its adjoint with unit lambdas (row r of J is
# J^T e_r); that is the tangent-linear model of the step, evaluated in
# Dual<3> numbers.
# ---------------------------------------------------------------------------
from oracle import oracle as O                      # noqa: E402
from romspath_amd.runner import injected_deviates, ndev_of  # noqa: E402
from romspath_amd.schedule import schedule          # noqa: E402


def _step(run, st):
    prm = run.cfg.params
    n = run.cfg.particles.n
    run.load(st)
    run.ctx.set_deviates(injected_deviates(prm.seed, st.step_index, n, ndev_of(prm)))
    run.ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index)


def _step_jacobian(run, st):
    n = run.cfg.particles.n
    run.ctx.adjoint_record(1)
    _step(run, st)
    J = np.empty((n, 3, 3))
    for r in range(3):
        lam = [np.zeros(n), np.zeros(n), np.zeros(n)]
        lam[r] = np.ones(n)
        (gx, gy, gz), _ = run.ctx.adjoint(lam)
        J[:, r, 0], J[:, r, 1], J[:, r, 2] = gx, gy, gz
    run.ctx.adjoint_record(0)
    return J


def test_step_jacobian_matches_reference_step_differences():
    # the tangent-linear of one step against central differences of the
    # oracle's step (the FP64 restatement of the reference) at the same state
    K = 20
    cfg = _case(n=256, seed=12)
    n = cfg.particles.n
    prm = cfg.params
    sched = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, K + 1))
    print(f"case tangent-linear of step {K} vs oracle differences, {n} particles", flush=True)
    run = GpuRun(cfg, precision="fp64")
    for st in sched[:K]:
        _step(run, st)
    J = _step_jacobian(run, sched[K])
    out = run.ctx.get_particles()
    run.close()
    base = O.run_oracle(cfg, max_steps=K + 1)
    assert np.array_equal(out["x"], base.x) and np.array_equal(out["z"], base.z)

    def shifted(dim, eps):
        def on_step(st, state):
            if st.step_index == K - 1:
                (state.x, state.y, state.z)[dim][:] += eps
        return O.run_oracle(cfg, max_steps=K + 1, on_step=on_step)

    fd = {}
    ok = np.ones(n, dtype=bool)
    for eps in (1e-4, 1e-5):
        F = np.empty((n, 3, 3))
        for dim in range(3):
            op, om = shifted(dim, eps), shifted(dim, -eps)
            ok &= (op.status == base.status) & (om.status == base.status)
            F[:, 0, dim] = (op.x - om.x) / (2 * eps)
            F[:, 1, dim] = (op.y - om.y) / (2 * eps)
            F[:, 2, dim] = (op.z - om.z) / (2 * eps)
        fd[eps] = F
    f1, f2 = fd[1e-4].reshape(n, 9), fd[1e-5].reshape(n, 9)
    scale = np.maximum(np.abs(f1).max(axis=1, keepdims=True), 1e-12)
    trusted = ok & (np.abs(f1 - f2) <= 1e-5 * scale).all(axis=1)
    d = (J.reshape(n, 9) - f2)[trusted]
    nrm = np.linalg.norm(d) / np.linalg.norm(f2[trusted])
    print(f"step Jacobian: {trusted.sum()}/{ok.sum()} particles trusted, norm-wise rel {nrm:.2e}",
          flush=True)
    assert trusted.sum() >= 0.9 * ok.sum()
    assert nrm <= 1e-5


def test_adjoint_is_the_transpose_of_the_tangent_linear():
    # lambda^T (J_N ... J_1) from the per-step tangent-linear models equals the
    # checkpointed reverse sweep of the same run (segments of 8, recomputed)
    N = 24
    cfg = _case(n=256, seed=13)
    n = cfg.particles.n
    prm = cfg.params
    sched = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, N))
    print(f"case adjoint vs tangent-linear product over {N} steps", flush=True)
    run = GpuRun(cfg, precision="fp64")
    P = np.tile(np.eye(3), (n, 1, 1))
    for st in sched:
        P = _step_jacobian(run, st) @ P
    a = run.ctx.get_particles()
    run.close()
    run = GpuRun(cfg, precision="fp64")
    run.ctx.adjoint_record(8)
    for st in sched:
        _step(run, st)
    b = run.ctx.get_particles()
    lam = np.random.default_rng(3).standard_normal((n, 3))
    (gx, gy, gz), _ = run.ctx.adjoint((lam[:, 0], lam[:, 1], lam[:, 2]))
    run.close()
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["z"], b["z"])
    g = np.stack([gx, gy, gz], axis=1)
    want = np.einsum("ni,nij->nj", lam, P)
    err = np.abs(g - want).max() / np.abs(want).max()
    print(f"adjoint vs tangent-linear: max rel {err:.2e}", flush=True)
    assert err <= 1e-10
