"""Nested-grid affine maps scl/off (grid_module.f90:437-481).

For every ordered pair of grids the reference searches for the first rho
point of one grid (smallest i, then j) whose coordinates equal a rho point of
the other exactly, and derives X_other = scl*X + off from it and the
refinement ratio.  The loops run ng2 outer, ng inner, and later pairs
overwrite the entries earlier pairs wrote; the final values are reproduced
here in that order.
"""
from __future__ import annotations

import numpy as np

MAXGRID = 5


def refine_matrix(ngrid, refine):
    g = np.ones((ngrid, ngrid))
    for a in range(ngrid):
        for b in range(ngrid):
            d = abs(a - b)
            for _ in range(d):
                g[a, b] = refine[d - 1] * g[a, b]
    return g


def _first_match(ga, gb):
    """(i, j, ii, jj), 1-based: first point of ga (i outer, then j) that
    coincides with a point of gb, and that point of gb (smallest ii, jj)."""
    xa, ya = ga.x_rho, ga.y_rho
    xb, yb = gb.x_rho, gb.y_rho
    lut = {}
    # gb points in (ii outer, jj inner) order: keep the first per coordinate
    for ii in range(gb.L):
        for jj in range(gb.M):
            key = (xb[jj, ii], yb[jj, ii])
            if key not in lut:
                lut[key] = (ii + 1, jj + 1)
    for i in range(ga.L):
        for j in range(ga.M):
            hit = lut.get((xa[j, i], ya[j, i]))
            if hit is not None:
                return i + 1, j + 1, hit[0], hit[1]
    # no match: Fortran DO variables end one past their last value
    return ga.L + 1, ga.M + 1, gb.L + 1, gb.M + 1


def scale_offsets(grids, refine):
    """Return scl[ng][i][2], off[ng][i][2] (0-based ng and i) for all grids."""
    n = len(grids)
    grefine = refine_matrix(n, refine)
    scl = np.ones((n, MAXGRID, 2))
    off = np.zeros((n, MAXGRID, 2))
    for ng2 in range(n):
        for ng in range(n):
            i, j, ii, jj = _first_match(grids[ng], grids[ng2])
            scl[ng, ng2, :] = 1.0 / grefine[ng, ng2]
            scl[ng2, ng, :] = grefine[ng2, ng]
            off[ng, ng2, 0] = float(ii) - scl[ng, ng2, 0] * float(i)
            off[ng, ng2, 1] = float(jj) - scl[ng, ng2, 1] * float(j)
            off[ng2, ng, 0] = float(i) - scl[ng2, ng, 0] * float(ii)
            off[ng2, ng, 1] = float(j) - scl[ng2, ng, 1] * float(jj)
    return scl, off
