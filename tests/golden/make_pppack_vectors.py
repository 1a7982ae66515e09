"""Writes tests/golden/pppack_smooth.npz: inputs and outputs of the
reference's own PPPACK (pppack.f90 compiled as is into oracle/_ref by
oracle/Makefile) for smooth (the smoothing spline of VTurb) and ppvalu, so
that the oracle's restatement stays pinned where the reference sources are
absent.  Run from the repository root after `make -C oracle`."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402


def profile(n, rng):
    """A VTurb-like AKs profile on a uniform column (same family as the unit tests)."""
    z0 = -rng.uniform(10.0, 200.0)
    x = z0 + (-z0) * np.arange(n) / (n - 1)
    sig = (x - z0) / (-z0)
    y = 4e-2 * sig * (1 - sig) + rng.normal(0, 1e-4, n)
    return x, y


def main():
    R = O.ref_pppack()
    if R is None:
        raise SystemExit("oracle/_ref/libpppack.so is not built (make -C oracle)")
    rng = np.random.default_rng(2024)
    cases = {"n": [], "s": [], "x": [], "y": [], "dy": [], "sf": [], "a": []}
    for _ in range(24):
        n = int(rng.integers(6, 65))
        x, y = profile(n, rng)
        dy = np.full(n, 1e-4)
        s = float(rng.choice([0.6, 0.0, 1e3, 1e-3, 0.05]))
        v = np.zeros(n * 7)
        a = np.zeros(n * 4)
        nn, ss = C.c_int(n), C.c_double(s)
        sf = R.smooth_(O.dptr(x), O.dptr(y), O.dptr(dy), C.byref(nn), C.byref(ss), O.dptr(v), O.dptr(a))
        for k, val in (("n", n), ("s", s), ("x", x), ("y", y), ("dy", dy), ("sf", sf), ("a", a)):
            cases[k].append(val)
    # ppvalu on the last case's spline: value and first derivative at random points
    n = cases["n"][-1]
    x = cases["x"][-1]
    coef = np.ascontiguousarray(cases["a"][-1])
    px = rng.uniform(x[0] - 1.0, x[-1] + 1.0, 100)
    pv = np.empty((100, 2))
    for i, xx in enumerate(px):
        for jd in (0, 1):
            l, k, xv, j = C.c_int(n - 1), C.c_int(4), C.c_double(xx), C.c_int(jd)
            pv[i, jd] = R.ppvalu_(O.dptr(x), O.dptr(coef), C.byref(l), C.byref(k), C.byref(xv), C.byref(j))
    offs = np.cumsum([0] + cases["n"])
    np.savez_compressed(
        Path(__file__).with_name("pppack_smooth.npz"),
        n=np.array(cases["n"]), s=np.array(cases["s"]), sf=np.array(cases["sf"]), offs=offs,
        x=np.concatenate(cases["x"]), y=np.concatenate(cases["y"]), dy=np.concatenate(cases["dy"]),
        a=np.concatenate(cases["a"]), ppvalu_x=px, ppvalu=pv)


if __name__ == "__main__":
    main()
