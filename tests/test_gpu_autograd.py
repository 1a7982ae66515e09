"""The adjoint through torch.autograd (verification harness only).

ParticleRun is a torch.autograd.Function around a recorded run: forward
steps the particles through the C-ABI with rp_adjoint_record (checkpoints
every CKPT steps) and returns the final x, y, z; backward hands the
incoming gradients to rp_adjoint as lambda and returns dL/d(release x, y, z)
and dL/d(velocity-perturbation coefficients).  The loss is the kind the
adjoint workload uses: squared distances from the final positions to
target points plus a Gaussian soft-box count of particles in a target
region, both written in torch, so torch's chain rule meets the adjoint.
Checked against central finite differences of the same forward."""
import numpy as np
import pytest
import torch

from romspath_amd import synth
from romspath_amd.runner import GpuRun

pytestmark = pytest.mark.gpu

STEPS = 60
CKPT = 16


class ParticleRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X0, Y0, Z0, coef, cfg, precision):
        p = cfg.particles
        parts = synth.Particles(X0.detach().numpy().copy(), Y0.detach().numpy().copy(),
                                Z0.detach().numpy().copy(), p.dob.copy())
        run = GpuRun(synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, parts, cfg.refine),
                     precision=precision)
        run.ctx.set_perturbation(coef.detach().numpy())
        run.ctx.adjoint_record(CKPT)
        out = run.run(max_steps=STEPS)
        ctx.run = run
        ctx.coef_shape = tuple(coef.shape)
        return (torch.from_numpy(out["x"].copy()), torch.from_numpy(out["y"].copy()),
                torch.from_numpy(out["z"].copy()))

    @staticmethod
    def backward(ctx, gx, gy, gz):
        lam = tuple(g.detach().contiguous().numpy().astype(np.float64) for g in (gx, gy, gz))
        (ax, ay, az), gp = ctx.run.ctx.adjoint(lam, want_pert=True, pert_shape=ctx.coef_shape)
        ctx.run.close()
        ctx.run = None
        return (torch.from_numpy(ax), torch.from_numpy(ay), torch.from_numpy(az),
                torch.from_numpy(gp), None, None)


def final_positions(X0, Y0, Z0, coef, cfg, precision):
    """The same forward, for the differences.  It records too: a recorded
    run takes the step's FP64-arithmetic path (the map the adjoint
    differentiates), which in the FP32-field form differs from the plain
    run's faster one."""
    p = cfg.particles
    parts = synth.Particles(X0.copy(), Y0.copy(), Z0.copy(), p.dob.copy())
    run = GpuRun(synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, parts, cfg.refine),
                 precision=precision)
    run.ctx.set_perturbation(coef)
    run.ctx.adjoint_record(CKPT)
    out = run.run(max_steps=STEPS)
    run.close()
    return out


def loss_terms(x, y, z, tx, ty, box, lib):
    """Per-particle loss: squared distance (cells) to the target plus a
    Gaussian soft count of the particle inside the target box."""
    (bx0, bx1, by0, by1, sig) = box
    d2 = (x - tx) ** 2 + (y - ty) ** 2
    inside = (lib.exp(-((x - 0.5 * (bx0 + bx1)) / sig) ** 2) * lib.exp(-((y - 0.5 * (by0 + by1)) / sig) ** 2))
    return d2 - 3.0 * inside + 0.1 * z


def _check(g, fd1, fd2, ok, name, tol=1e-5):
    # difference quotients are trusted where two step sizes agree (the
    # trajectory map is piecewise smooth); the gradient must match them
    floor = 1e-3 * np.median(np.abs(fd1[ok]))
    trusted = ok & (np.abs(fd1 - fd2) <= tol * np.maximum(np.abs(fd1), floor))
    d = (g - fd2)[trusted]
    nrm = np.linalg.norm(d) / np.linalg.norm(fd2[trusted])
    print(f"{name}: {trusted.sum()}/{ok.sum()} trusted, norm-wise rel {nrm:.2e}", flush=True)
    assert trusted.sum() >= 0.9 * ok.sum()
    assert nrm <= tol


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_autograd_release_and_perturbation_gradients(precision):
    cfg = synth.config_mini(n_particles=256, days=STEPS * 120 / 86400.0, dob_spread=0.0, seed=13)
    n = cfg.particles.n
    rng = np.random.default_rng(21)
    P = cfg.particles
    tx = P.x + rng.normal(0.0, 2.0, n)
    ty = P.y + rng.normal(0.0, 2.0, n)
    box = (float(P.x.min()) + 3, float(P.x.min()) + 9, float(P.y.min()) + 3, float(P.y.min()) + 9, 4.0)
    coef = 0.02 * rng.standard_normal((2, 16, 16))
    print(f"case autograd {precision}: {n} particles, {STEPS} steps, checkpoints every {CKPT}", flush=True)
    X0 = torch.tensor(P.x, requires_grad=True)
    Y0 = torch.tensor(P.y, requires_grad=True)
    Z0 = torch.tensor(P.z, requires_grad=True)
    C = torch.tensor(coef, requires_grad=True)
    x, y, z = ParticleRun.apply(X0, Y0, Z0, C, cfg, precision)
    L = loss_terms(x, y, z, torch.from_numpy(tx), torch.from_numpy(ty), box, torch).sum()
    L.backward()
    gX, gZ, gC = X0.grad.numpy(), Z0.grad.numpy(), C.grad.numpy()
    assert np.all(np.isfinite(gX)) and np.all(np.isfinite(gC))

    def loss_np(out):
        return loss_terms(out["x"], out["y"], out["z"], tx, ty, box, np)

    base = final_positions(P.x, P.y, P.z, coef, cfg, precision)
    assert np.array_equal(base["x"], x.detach().numpy())
    # release x and z: every particle's final state depends only on its own
    # release, so one shifted pair gives every particle's difference quotient
    for name, g, shift in (("d/dx0", gX, 0), ("d/dz0", gZ, 2)):
        fds, ok = [], np.ones(n, dtype=bool)
        for eps in (1e-4, 1e-5):
            vals = []
            for sgn in (1.0, -1.0):
                r = [P.x.copy(), P.y.copy(), P.z.copy()]
                r[shift] += sgn * eps
                o = final_positions(*r, coef, cfg, precision)
                ok &= o["status"] == base["status"]
                vals.append(loss_np(o))
            fds.append((vals[0] - vals[1]) / (2 * eps))
        _check(g, fds[0], fds[1], ok, name)
    # the two most influential perturbation coefficients
    for comp in (0, 1):
        j, i = np.unravel_index(np.argmax(np.abs(gC[comp])), gC[comp].shape)
        eps = 1e-5
        vals = []
        for sgn in (1.0, -1.0):
            c2 = coef.copy()
            c2[comp, j, i] += sgn * eps
            vals.append(loss_np(final_positions(P.x, P.y, P.z, c2, cfg, precision)).sum())
        fd = (vals[0] - vals[1]) / (2 * eps)
        rel = abs(gC[comp, j, i] - fd) / max(abs(fd), 1e-12)
        print(f"dL/dcoef[{comp},{j},{i}]: autograd {gC[comp, j, i]:.10e} fd {fd:.10e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-5
