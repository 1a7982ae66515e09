"""Seeded synthetic ROMS grids, hydrodynamic records and particle releases.

The reference reads grids and hydrodynamic snapshots from ROMS netCDF files
(grid_module.f90:41-422, hydrodynamic_module.f90:112-760).  No such files ship
with the reference, so every workload here is generated from closed forms:

* Cartesian grids (spherical=0, angle=0, pm=pn=1/dx), rho point (i,j) at
  x=(i-1)*dx+x0, y=(j-1)*dy+y0, with the reference's 1-based index
  convention: a position X is a fractional rho index (rho i sits at X=i).
* s-coordinates with Vtransform=2 and the published ROMS Vstretching=4
  stretching function.
* Velocities are depth-uniform, divergence-free streamfunction fields (double
  gyre, drifting Gaussian eddies) plus an optional M2 tide; w follows from the
  kinematic bottom condition w=-u.grad(h); AKs is a parabolic profile.

Array conventions follow the reference's Fortran column-major storage: a
Fortran array F(xi, eta[, s]) is held as a C-ordered numpy array of shape
([s,] eta, xi), so the bytes are identical to the Fortran array's.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

M2_PERIOD = 44714.0


def vstretching4(n: int, theta_s: float, theta_b: float):
    """ROMS Vstretching=4 (Shchepetkin 2010) s and C(s) at rho and w levels.

    Returns (s_rho[N], cs_r[N], s_w[N+1], cs_w[N+1]) ordered bottom to top,
    matching the reference's 1-based s_rho(1..N), s_w(1..N+1).
    """
    def stretch(s):
        if theta_s > 0.0:
            csur = (1.0 - np.cosh(theta_s * s)) / (np.cosh(theta_s) - 1.0)
        else:
            csur = -s * s
        if theta_b > 0.0:
            return (np.exp(theta_b * csur) - 1.0) / (1.0 - np.exp(-theta_b))
        return csur

    k_r = np.arange(1, n + 1, dtype=np.float64)
    s_rho = (k_r - n - 0.5) / n
    k_w = np.arange(0, n + 1, dtype=np.float64)
    s_w = (k_w - n) / n
    cs_r = stretch(s_rho)
    cs_w = stretch(s_w)
    cs_w[0] = -1.0
    cs_w[-1] = 0.0
    return s_rho, cs_r, s_w, cs_w


@dataclass
class Grid:
    """One ROMS grid (the fields of GRIDDATA, grid_module.f90:14-36)."""
    L: int                 # xi_rho
    M: int                 # eta_rho
    N: int                 # s_rho
    dx: float
    dy: float
    x0: float
    y0: float
    h: np.ndarray          # (M, L)
    mask_rho: np.ndarray   # (M, L) int32
    vtransform: int = 2
    hc: float = 10.0
    theta_s: float = 5.0
    theta_b: float = 0.4
    s_rho: np.ndarray = field(default=None)
    cs_r: np.ndarray = field(default=None)
    s_w: np.ndarray = field(default=None)
    cs_w: np.ndarray = field(default=None)
    # coordinates x = x0 + (i*dx*den)/den: a refined child computes its
    # points as x0 + (i*1000)/3 so that every third one equals a parent
    # point bit for bit (the nested maps search for exact matches)
    den: int = 1

    def __post_init__(self):
        if self.s_rho is None:
            self.s_rho, self.cs_r, self.s_w, self.cs_w = vstretching4(
                self.N, self.theta_s, self.theta_b)
        self.mask_rho = np.ascontiguousarray(self.mask_rho, dtype=np.int32)
        self.h = np.ascontiguousarray(self.h, dtype=np.float64)

    @property
    def pm(self):
        return np.full((self.M, self.L), 1.0 / self.dx)

    @property
    def pn(self):
        return np.full((self.M, self.L), 1.0 / self.dy)

    @property
    def mask_u(self):
        m = self.mask_rho
        return np.ascontiguousarray(m[:, :-1] * m[:, 1:], dtype=np.int32)

    @property
    def mask_v(self):
        m = self.mask_rho
        return np.ascontiguousarray(m[:-1, :] * m[1:, :], dtype=np.int32)

    def _axis(self, n, d0, d):
        i = np.arange(n, dtype=np.float64)
        if self.den == 1:
            return d0 + i * d
        return d0 + (i * (d * self.den)) / self.den

    @property
    def x_rho(self):
        return np.broadcast_to(self._axis(self.L, self.x0, self.dx), (self.M, self.L)).copy()

    @property
    def y_rho(self):
        return np.broadcast_to(self._axis(self.M, self.y0, self.dy)[:, None], (self.M, self.L)).copy()

    # coordinates of the staggered points (metres)
    def xy_rho(self):
        return self.x_rho, self.y_rho

    def xy_u(self):
        xr = self._axis(self.L, self.x0, self.dx)
        x = 0.5 * (xr[:-1] + xr[1:])
        y = self._axis(self.M, self.y0, self.dy)
        return np.meshgrid(x, y)

    def xy_v(self):
        x = self._axis(self.L, self.x0, self.dx)
        yr = self._axis(self.M, self.y0, self.dy)
        y = 0.5 * (yr[:-1] + yr[1:])
        return np.meshgrid(x, y)

    def z_w(self, zeta):
        """z of the w levels (Vtransform 2), shape (N+1, M, L)."""
        h = self.h
        S = (self.hc * self.s_w[:, None, None] + h[None] * self.cs_w[:, None, None]) / (self.hc + h[None])
        return zeta[None] + (zeta[None] + h[None]) * S


# --------------------------------------------------------------------------
# analytic flow fields
# --------------------------------------------------------------------------

class DoubleGyre:
    """Steady depth-uniform double gyre psi=(U0*Ly/pi) sin(pi*x') sin(pi*y')."""

    def __init__(self, grid: Grid, U0=0.3):
        self.U0 = U0
        self.Lx = (grid.L - 1) * grid.dx
        self.Ly = (grid.M - 1) * grid.dy
        self.x0, self.y0 = grid.x0, grid.y0

    def uv(self, x, y, t):
        xp = 2.0 * (x - self.x0) / self.Lx
        yp = (y - self.y0) / self.Ly
        u = -self.U0 * np.sin(np.pi * xp) * np.cos(np.pi * yp)
        v = self.U0 * (2.0 * self.Ly / self.Lx) * np.cos(np.pi * xp) * np.sin(np.pi * yp)
        return u, v


class EddyField:
    """Sum of drifting Gaussian-streamfunction eddies plus an M2 tide in x."""

    def __init__(self, rng, n_eddies, xlim, ylim, rmin, rmax, vmax=0.5,
                 drift=0.05, tide=0.2, scale_grid=None):
        self.xc = rng.uniform(xlim[0], xlim[1], n_eddies)
        self.yc = rng.uniform(ylim[0], ylim[1], n_eddies)
        self.R = rng.uniform(rmin, rmax, n_eddies)
        peak = rng.uniform(0.3, 1.0, n_eddies) * vmax
        sign = np.where(rng.uniform(size=n_eddies) < 0.5, -1.0, 1.0)
        self.A = sign * peak * self.R * math.exp(0.5)
        ang = rng.uniform(0.0, 2.0 * np.pi, n_eddies)
        self.cx = drift * np.cos(ang)
        self.cy = drift * np.sin(ang)
        self.tide = tide
        self.scale = 1.0
        if scale_grid is not None:
            x, y = scale_grid
            u, v = self._eddies(x, y, 0.0)
            sp = float(np.sqrt(u * u + v * v).max())
            if sp > vmax:
                self.scale = vmax / sp

    def _eddies(self, x, y, t):
        u = np.zeros(np.broadcast(x, y).shape)
        v = np.zeros_like(u)
        for e in range(self.A.size):
            dx = x - (self.xc[e] + self.cx[e] * t)
            dy = y - (self.yc[e] + self.cy[e] * t)
            R2 = self.R[e] * self.R[e]
            g = self.A[e] * np.exp(-(dx * dx + dy * dy) / (2.0 * R2)) / R2
            u += g * dy          # u = -dpsi/dy
            v -= g * dx          # v =  dpsi/dx
        return u, v

    def uv(self, x, y, t):
        u, v = self._eddies(x, y, t)
        u = u * self.scale + self.tide * math.cos(2.0 * np.pi * t / M2_PERIOD)
        return u, v * self.scale


class SumField:
    def __init__(self, *parts):
        self.parts = parts

    def uv(self, x, y, t):
        u = 0.0
        v = 0.0
        for p in self.parts:
            a, b = p.uv(x, y, t)
            u = u + a
            v = v + b
        return u, v


@dataclass
class HydroModel:
    """Closed-form hydrodynamics on one grid: records at t_k=(k-1)*dt."""
    grid: Grid
    flow: object
    hx: object = None            # callable (x, y) -> dh/dx   (None: flat)
    hy: object = None
    zeta_amp: float = 0.0        # tidal elevation amplitude (m)
    aks_peak: float = 0.0        # parabolic AKs peak (m2/s)
    stokes: tuple = None         # (u_s0, k, direction_rad): u_s0*exp(2kz)

    def zeta(self, t):
        g = self.grid
        if self.zeta_amp == 0.0:
            return np.zeros((g.M, g.L))
        x, y = g.xy_rho()
        Lx = (g.L - 1) * g.dx
        return (self.zeta_amp * math.cos(2.0 * np.pi * t / M2_PERIOD)
                * np.cos(np.pi * (x - g.x0) / Lx))

    def record(self, t: float) -> dict:
        g = self.grid
        N = g.N
        zeta = self.zeta(t) * g.mask_rho
        xu, yu = g.xy_u()
        u2, _ = self.flow.uv(xu, yu, t)
        xv, yv = g.xy_v()
        _, v2 = self.flow.uv(xv, yv, t)
        u2 = u2 * g.mask_u
        v2 = v2 * g.mask_v
        u = np.broadcast_to(u2, (N, g.M, g.L - 1)).copy()
        v = np.broadcast_to(v2, (N, g.M - 1, g.L)).copy()
        if self.stokes is not None:
            us0, k, ang = self.stokes
            # Stokes drift profile at rho-level depths of the u / v points
            zr = self._z_rho(zeta)
            zu = 0.5 * (zr[:, :, :-1] + zr[:, :, 1:])
            zv = 0.5 * (zr[:, :-1, :] + zr[:, 1:, :])
            su = us0 * math.cos(ang) * np.exp(2.0 * k * zu) * g.mask_u
            sv = us0 * math.sin(ang) * np.exp(2.0 * k * zv) * g.mask_v
        else:
            su = sv = None
        xr, yr = g.xy_rho()
        if self.hx is not None:
            ur, vr = self.flow.uv(xr, yr, t)
            w2 = -(ur * self.hx(xr, yr) + (vr * self.hy(xr, yr) if self.hy else 0.0))
            w2 = w2 * g.mask_rho
        else:
            w2 = np.zeros((g.M, g.L))
        w = np.broadcast_to(w2, (N + 1, g.M, g.L)).copy()
        if self.aks_peak > 0.0:
            zw = g.z_w(zeta)
            sig = (zw + g.h[None]) / (zeta[None] + g.h[None])
            aks = 4.0 * self.aks_peak * sig * (1.0 - sig) * g.mask_rho[None]
        else:
            aks = np.zeros((N + 1, g.M, g.L))
        rec = dict(zeta=zeta, u=u, v=v, w=w, aks=aks)
        if su is not None:
            rec["su"] = su
            rec["sv"] = sv
        return rec

    def _z_rho(self, zeta):
        g = self.grid
        h = g.h
        S = (g.hc * g.s_rho[:, None, None] + h[None] * g.cs_r[:, None, None]) / (g.hc + h[None])
        return zeta[None] + (zeta[None] + h[None]) * S


# --------------------------------------------------------------------------
# run configuration (the namelist scalars the hot path reads)
# --------------------------------------------------------------------------

@dataclass
class RunParams:
    """Subset of the ROMSPath.data namelist on the hot path (ROMSPath.h:77-258)."""
    dt: int = 3600               # timeparam: external step (s)
    idt: int = 60                # timeparam: internal step (s)
    days: float = 1.0            # timeparam
    iprint: int = 3600           # timeparam
    VTurbOn: bool = False        # turbparam
    HTurbOn: bool = False
    serr: float = 1.0e-4
    smth: float = 0.6
    sub: float = 4.0
    deltat: float = 1.0
    AKSback: float = 1.0e-8
    ConstantHTurb: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    scheme: int = 1              # advectparam
    nsb: int = 0
    vertdist: float = 0.0
    OpenOceanBoundary: bool = True   # behavparam
    swimfast: float = 0.0            # behavparam (written to the file only)
    seed: int = 1                    # other
    SaltTempMean: bool = False       # other: tracer means between outputs
    TempOffset: float = 0.0          # other: added to the temperature records

    @property
    def n_internal(self):
        return int(self.dt // self.idt)

    @property
    def n_external(self):
        return int(int(self.days * 86400.0) // self.dt)

    @property
    def loop(self):
        return int(self.idt // int(self.deltat))


@dataclass
class Particles:
    """Release state, the par(:,k) columns of ROMSPath.f90:104-129."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    dob: np.ndarray
    gid: np.ndarray = None       # 1-based grid id
    status: np.ndarray = None
    age: np.ndarray = None
    isin: np.ndarray = None

    def __post_init__(self):
        n = self.x.size
        if self.gid is None:
            self.gid = np.ones(n, dtype=np.int32)
        if self.status is None:
            self.status = np.zeros(n)
        if self.age is None:
            self.age = np.zeros(n)
        if self.isin is None:
            self.isin = np.ones(n, dtype=np.uint8)

    @property
    def n(self):
        return self.x.size

    def copy(self):
        return Particles(self.x.copy(), self.y.copy(), self.z.copy(), self.dob.copy(),
                         self.gid.copy(), self.status.copy(), self.age.copy(), self.isin.copy())


@dataclass
class Config:
    name: str
    grids: list                  # child first, as the reference requires
    hydro: list                  # one HydroModel per grid
    params: RunParams
    particles: Particles
    refine: tuple = (1.0,)

    @property
    def ngrid(self):
        return len(self.grids)


def bilinear_depth(g: Grid, X, Y):
    I = np.floor(X).astype(np.int64)
    J = np.floor(Y).astype(np.int64)
    fx = X - I
    fy = Y - J
    h = g.h
    v1 = h[J - 1, I - 1]
    v2 = h[J - 1, I]
    v3 = h[J, I]
    v4 = h[J, I - 1]
    return (1 - fx) * (1 - fy) * v1 + fx * (1 - fy) * v2 + fx * fy * v3 + (1 - fx) * fy * v4


def _release_uniform(rng, g: Grid, n, margin=2.0, smin=-0.98, smax=-0.02, require_wet=True):
    xs = []
    ys = []
    need = n
    m = g.mask_rho
    while need > 0:
        X = rng.uniform(margin, g.L - margin, 2 * need + 16)
        Y = rng.uniform(margin, g.M - margin, 2 * need + 16)
        if require_wet:
            I = np.floor(X).astype(np.int64)
            J = np.floor(Y).astype(np.int64)
            wet = (m[J - 1, I - 1] + m[J - 1, I] + m[J, I] + m[J, I - 1]) == 4
            X = X[wet]
            Y = Y[wet]
        xs.append(X[:need])
        ys.append(Y[:need])
        need -= xs[-1].size
    X = np.concatenate(xs)
    Y = np.concatenate(ys)
    s = rng.uniform(smin, smax, n)
    Z = s * bilinear_depth(g, X, Y)
    return X, Y, Z


def _tanh_shelf(L, M, dx, hmin=20.0, dh=90.0, xmid=200e3, width=25e3, x0=0.0):
    x = x0 + np.arange(L) * dx
    h1 = hmin + dh * (1.0 + np.tanh((x - xmid) / width))
    return np.broadcast_to(h1, (M, L)).copy()


def _island_mask(L, M, dx, dy, xc, yc, r, x0=0.0, y0=0.0):
    x = x0 + np.arange(L) * dx
    y = y0 + np.arange(M) * dy
    X, Y = np.meshgrid(x, y)
    return ((X - xc) ** 2 + (Y - yc) ** 2 >= r * r).astype(np.int32)


def config1(n_particles=10_000, seed=1, days=1.0):
    """Plain method, CPU-runnable: 100x80x20 flat 50 m, double gyre, RK4 only."""
    L, M, N = 100, 80, 20
    g = Grid(L, M, N, 1000.0, 1000.0, 0.0, 0.0, h=np.full((M, L), 50.0),
             mask_rho=np.ones((M, L), np.int32))
    hyd = HydroModel(g, DoubleGyre(g, 0.3))
    rng = np.random.default_rng(seed)
    X = rng.uniform(2.0, 98.0, n_particles)
    Y = rng.uniform(2.0, 78.0, n_particles)
    s = rng.uniform(-0.98, -0.02, n_particles)
    Z = s * 50.0
    p = Particles(X, Y, Z, np.zeros(n_particles))
    prm = RunParams(dt=3600, idt=60, days=days, seed=seed)
    return Config("C1", [g], [hyd], prm, p)


def _shelf_hx(dh=90.0, xmid=200e3, width=25e3):
    def hx(x, y):
        c = np.cosh((x - xmid) / width)
        return dh / (width * c * c)
    return hx


def config2(n_particles=1_000_000, seed=2, days=3.0, L=400, M=300, N=30,
            n_eddies=16, zeta_amp=0.0):
    """Plain method at GPU scale: 400x300x30 tanh shelf + island, eddies + M2, RK4+RW."""
    dx = 1000.0
    xmid = 0.5 * (L - 1) * dx if L != 400 else 200e3
    width = 25e3 * (L / 400.0)
    h = _tanh_shelf(L, M, dx, xmid=xmid, width=width)
    xc, yc = 0.5 * (L - 1) * dx, 0.5 * (M - 1) * dx
    rad = 15e3 * (L / 400.0)
    mask = _island_mask(L, M, dx, dx, xc, yc, rad)
    g = Grid(L, M, N, dx, dx, 0.0, 0.0, h=h, mask_rho=mask)
    rng = np.random.default_rng(seed)
    scale = L / 400.0
    xr, yr = g.xy_rho()
    flow = EddyField(rng, n_eddies, (0.0, (L - 1) * dx), (0.0, (M - 1) * dx),
                     5e3 * scale, 20e3 * scale, vmax=0.5, drift=0.05, tide=0.2,
                     scale_grid=(xr, yr))
    hyd = HydroModel(g, flow, hx=_shelf_hx(xmid=xmid, width=width), hy=None,
                     zeta_amp=zeta_amp, aks_peak=1e-2)
    X, Y, Z = _release_uniform(rng, g, n_particles)
    p = Particles(X, Y, Z, np.zeros(n_particles))
    prm = RunParams(dt=3600, idt=120, days=days, VTurbOn=True, HTurbOn=True,
                    deltat=12.0, ConstantHTurb=(1.0, 1.0, 1.0, 1.0, 1.0), seed=seed)
    return Config("C2", [g], [hyd], prm, p)


def config_mini(n_particles=600, seed=7, days=0.25, vturb=True, hturb=True,
                L=48, M=40, N=10, zeta_amp=0.3, dob_spread=3 * 3600.0, idt=120,
                deltat=12.0):
    """Small C2-family case that exercises every branch: island (partial-wet
    cells, land rejection), open-boundary exits, sloping bottom, tidal zeta,
    staggered releases and both random walks."""
    dx = 1000.0
    xmid = 0.5 * (L - 1) * dx
    h = _tanh_shelf(L, M, dx, hmin=8.0, dh=40.0, xmid=xmid, width=6e3)
    xc, yc = 0.5 * (L - 1) * dx + 0.37 * dx, 0.5 * (M - 1) * dx - 0.21 * dx
    mask = _island_mask(L, M, dx, dx, xc, yc, 6.3e3)
    g = Grid(L, M, N, dx, dx, 0.0, 0.0, h=h, mask_rho=mask)
    rng = np.random.default_rng(seed)
    xr, yr = g.xy_rho()
    flow = EddyField(rng, 5, (0.0, (L - 1) * dx), (0.0, (M - 1) * dx), 3e3, 8e3,
                     vmax=0.6, drift=0.05, tide=0.25, scale_grid=(xr, yr))
    hyd = HydroModel(g, flow, hx=_shelf_hx(dh=40.0, xmid=xmid, width=6e3), hy=None,
                     zeta_amp=zeta_amp, aks_peak=1e-2)
    X, Y, Z = _release_uniform(rng, g, n_particles, margin=2.0)
    dob = np.floor(rng.uniform(0.0, dob_spread, n_particles))
    p = Particles(X, Y, Z, dob)
    prm = RunParams(dt=3600, idt=idt, days=days, VTurbOn=vturb, HTurbOn=hturb,
                    deltat=deltat, ConstantHTurb=(1.0,) * 5, seed=seed)
    return Config("mini", [g], [hyd], prm, p)


def _nested(L, M, N, Lc, Mc, i0, j0, refine, n_particles, seed, days, n_eddies, n_child_eddies,
            rad_scale, zeta_amp, hturb, stokes, name, xmid=200e3, width=25e3, island_r=15e3):
    """Parent + centred refined child (grid 1 = child, 2 = parent, as the
    reference orders them).  Child point (1,1) sits on parent point (i0,j0)."""
    dx = 1000.0
    hp = _tanh_shelf(L, M, dx, xmid=xmid, width=width)
    xc, yc = 0.5 * (L - 1) * dx, 0.5 * (M - 1) * dx
    parent = Grid(L, M, N, dx, dx, 0.0, 0.0, h=hp,
                  mask_rho=_island_mask(L, M, dx, dx, xc, yc, island_r))
    x0, y0 = (i0 - 1) * dx, (j0 - 1) * dx
    dxc = dx / refine
    child = Grid(Lc, Mc, N, dxc, dxc, x0, y0, h=np.zeros((Mc, Lc)),
                 mask_rho=np.ones((Mc, Lc), np.int32), den=int(refine))
    xr, yr = child.xy_rho()
    child.h = np.ascontiguousarray(20.0 + 90.0 * (1.0 + np.tanh((xr - xmid) / width)))
    child.mask_rho = np.ascontiguousarray(((xr - xc) ** 2 + (yr - yc) ** 2 >= island_r ** 2).astype(np.int32))
    rng = np.random.default_rng(seed)
    pxr, pyr = parent.xy_rho()
    flow = EddyField(rng, n_eddies, (0.0, (L - 1) * dx), (0.0, (M - 1) * dx),
                     5e3 * rad_scale, 20e3 * rad_scale, vmax=0.5, drift=0.05, tide=0.2,
                     scale_grid=(pxr, pyr))
    small = EddyField(rng, n_child_eddies, (x0, x0 + (Lc - 1) * dxc), (y0, y0 + (Mc - 1) * dxc),
                      1e3 * rad_scale, 3e3 * rad_scale, vmax=0.2, drift=0.05, tide=0.0,
                      scale_grid=(xr, yr))
    hx = _shelf_hx(xmid=xmid, width=width)
    hyd_p = HydroModel(parent, flow, hx=hx, zeta_amp=zeta_amp, aks_peak=1e-2, stokes=stokes)
    hyd_c = HydroModel(child, SumField(flow, small), hx=hx, zeta_amp=zeta_amp, aks_peak=1e-2,
                       stokes=stokes)
    nh = n_particles // 2
    Xc, Yc, Zc = _release_uniform(rng, child, nh)
    # parent releases outside the child's footprint
    xs, ys = [], []
    need = n_particles - nh
    x1c, y1c = x0 + (Lc - 1) * dxc, y0 + (Mc - 1) * dxc
    while need > 0:
        X, Y, _ = _release_uniform(rng, parent, 2 * need + 16)
        xm, ym = (X - 1.0) * dx, (Y - 1.0) * dx
        out = ~((xm >= x0 - dx) & (xm <= x1c + dx) & (ym >= y0 - dx) & (ym <= y1c + dx))
        xs.append(X[out][:need])
        ys.append(Y[out][:need])
        need -= xs[-1].size
    Xp, Yp = np.concatenate(xs), np.concatenate(ys)
    s_ = rng.uniform(-0.98, -0.02, Xp.size)
    Zp = s_ * bilinear_depth(parent, Xp, Yp)
    gid = np.r_[np.ones(nh, np.int32), np.full(Xp.size, 2, np.int32)]
    p = Particles(np.r_[Xc, Xp], np.r_[Yc, Yp], np.r_[Zc, Zp], np.zeros(n_particles), gid=gid)
    prm = RunParams(dt=3600, idt=120, days=days, VTurbOn=True, HTurbOn=hturb, deltat=12.0,
                    ConstantHTurb=(1.0,) * 5, seed=seed)
    return Config(name, [child, parent], [hyd_c, hyd_p], prm, p, refine=(float(refine),))


def config3(n_particles=4_000_000, seed=3, days=5.0):
    """Nested grids + Stokes drift: parent 600x500x30, 3:1 child 300x300x30 at the
    centre, 32 child-only eddies, u_s = 0.1 exp(2kz) from 30 deg, k = 0.1."""
    return _nested(600, 500, 30, 300, 300, 251, 201, 3, n_particles, seed, days, 16, 32, 1.0,
                   0.0, False, (0.1, 0.1, np.deg2rad(30.0)), "C3")


def config_mini_nested(n_particles=600, seed=8, days=2 / 24.0, zeta_amp=0.2, hturb=True):
    """Small C3-family case: parent 48x40x8 and a 3:1 child 30x30x8."""
    return _nested(48, 40, 8, 30, 30, 20, 15, 3, n_particles, seed, days, 5, 6, 0.25, zeta_amp,
                   hturb, (0.1, 0.1, np.deg2rad(30.0)), "mini-nested", xmid=24e3, width=6e3,
                   island_r=3.3e3)


def config4(n_scen=256, seed=4, days=3.0, nmin=1_000, nmax=200_000, sigma_km=2.0,
            start_spread=86400.0, field_seed=2):
    """Batched release ensembles on the C2 grid and fields: n_scen scenarios of
    round(exp(U[ln nmin, ln nmax])) particles, each a Gaussian cloud (sigma 2 km)
    around a random fully wet cell at random depths, released together at a
    scenario start time U[0, 24 h) (whole seconds).  Particles are stored
    scenario after scenario; returns (Config, counts)."""
    base = config2(n_particles=1, seed=field_seed, days=days)
    g = base.grids[0]
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_scen))).astype(np.int64)
    m = g.mask_rho
    sig = sigma_km * 1000.0 / g.dx
    xs, ys, zs, dobs = [], [], [], []
    for s_ in range(n_scen):
        # centre: a random cell whose four corners are wet, away from the edges
        while True:
            cx = rng.uniform(4.0, g.L - 4.0)
            cy = rng.uniform(4.0, g.M - 4.0)
            I, J = int(cx), int(cy)
            if m[J - 1, I - 1] + m[J - 1, I] + m[J, I] + m[J, I - 1] == 4:
                break
        n = int(counts[s_])
        X = np.empty(0)
        Y = np.empty(0)
        while X.size < n:
            k = 2 * (n - X.size) + 16
            x = rng.normal(cx, sig, k)
            y = rng.normal(cy, sig, k)
            ok = (x >= 2.0) & (x < g.L - 2.0) & (y >= 2.0) & (y < g.M - 2.0)
            x, y = x[ok], y[ok]
            I = x.astype(np.int64)
            J = y.astype(np.int64)
            wet = (m[J - 1, I - 1] + m[J - 1, I] + m[J, I] + m[J, I - 1]) == 4
            X = np.r_[X, x[wet]]
            Y = np.r_[Y, y[wet]]
        X, Y = X[:n], Y[:n]
        Z = rng.uniform(-0.98, -0.02, n) * bilinear_depth(g, X, Y)
        xs.append(X)
        ys.append(Y)
        zs.append(Z)
        dobs.append(np.full(n, float(np.floor(rng.uniform(0.0, start_spread)))))
    p = Particles(np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(dobs))
    cfg = Config("C4", base.grids, base.hydro, base.params, p)
    return cfg, counts
