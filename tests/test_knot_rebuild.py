"""The identities behind the FP32 vertical walk's knot rebuild at small
smoothing weight (knots_from_curvature): smooth's knot values
f = y - 6(1-p) serr^2 Q u (pppack.f90:5305-5314) have second differences
Q'f = p h tridiag(1,4,1) u, and f's least-squares line is y's.  The kernel's
two passes (y's line and, in closed form over T = tridiag(1,4,1) u, the line
of g with g(1) = g(2) = 0 and g's second differences p h^2 T; then the
recurrence plus the difference of the lines) are restated here in FP64 and
must give the oracle's knot values for every smoothing weight, p = 0 (the
least-squares line) included.  No GPU."""
import numpy as np
import pytest

from oracle import oracle as O


def two_pass_knots(y, u, h, p):
    """The kernel's two passes; u(1) = u(np) = 0."""
    n = len(y)
    c = 0.5 * (n + 1)
    sy = y.sum()
    sjy = ((np.arange(1, n + 1) - c) * y).sum()
    a0 = a1 = 0.0
    for j in range(2, n):
        T = u[j - 2] + 4.0 * u[j - 1] + u[j]
        m = n - j
        m2 = 0.5 * m * (m + 1.0)
        a0 += T * m2
        a1 += T * ((j - c) * m2 + m2 * (2.0 * m + 1.0) / 3.0)
    w = p * h * h
    sjj = n * (n * n - 1.0) / 12.0
    da, db = (sy - w * a0) / n, (sjy - w * a1) / sjj
    f = np.empty(n)
    f[0], f[1] = da + db * (1.0 - c), da + db * (2.0 - c)
    gm = g0 = 0.0
    for j in range(2, n):
        g1 = 2.0 * g0 - gm + w * (u[j - 2] + 4.0 * u[j - 1] + u[j])
        f[j] = g1 + da + db * (j + 1.0 - c)
        gm, g0 = g0, g1
    return f


def smooth(y, h, serr, s):
    """The oracle's smooth on uniform points: knot values, 6 p u, and p
    recovered from the second differences of f (Q'f = p h Tri u)."""
    n = len(y)
    x = np.arange(n) * h
    dy = np.full(n, serr)
    v, a = np.zeros(n * 7), np.zeros(n * 4)
    O.lib().og_smooth(O.dptr(x), O.dptr(y), O.dptr(dy), n, s, O.dptr(v), O.dptr(a))
    A = a.reshape(4, n)
    return A[0].copy(), A[2].copy()


@pytest.mark.parametrize("n,s", [(124, 0.6), (124, 30.0), (124, 1e6), (67, 5.0), (9, 2.0), (3, 1.0)])
def test_knots_from_curvature_equal_smooth(n, s):
    rng = np.random.default_rng(n + int(s))
    h, serr = 0.43, 1e-4
    z = np.arange(n) * h
    y = 1e-3 * (1.0 + np.sin(z / 7.0) + 0.3 * rng.standard_normal(n))
    f_ref, six_pu = smooth(y, h, serr, s)
    # the kernel takes p and u separately; only their product p u enters
    # (w u = p h^2 u), so p = 1 with u = p u reproduces it
    f = two_pass_knots(y, six_pu / 6.0, h, 1.0)
    # 1e-9 relative: the oracle's own FP64 difference y - 6(1-p) serr^2 Q u
    # cancels as p -> 0 (1e-10 relative at s = 1e6), 1e-15 elsewhere
    assert np.abs(f - f_ref).max() <= 1e-9 * np.abs(f_ref).max()


def test_p0_knots_are_the_least_squares_line():
    rng = np.random.default_rng(5)
    n, h = 50, 0.5
    y = rng.standard_normal(n)
    f = two_pass_knots(y, np.zeros(n), h, 0.0)
    j = np.arange(1, n + 1)
    b, a = np.polyfit(j, y, 1)
    assert np.allclose(f, a + b * j, rtol=0, atol=1e-12)
    f_ref, six_pu = smooth(y, h, 1e-4, 1e12)
    assert not six_pu.any()                          # smooth's p = 0 branch
    assert np.allclose(f, f_ref, rtol=0, atol=1e-9)   # the oracle's FP64 cancellation
