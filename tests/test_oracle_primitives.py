"""Oracle primitives pinned by mathematics (CPU).

Boys function (restating mdr.c:6-34) against mpmath's incomplete gamma and
its special cases; Rys roots/weights (the role of rys.c:2217-2243) against
the Gauss-rule moment conditions, the Hermite large-T limit (the constants
of rys.c:433-434) and the Gauss-Legendre T=0 limit; the Hermite expansion
coefficients (mdd.c:3-43) against the Gaussian product theorem.
"""
import math

import mpmath
import numpy as np
import pytest

import oracle

mpmath.mp.dps = 40


def boys_mp(m, T):
    if T == 0.0:
        return mpmath.mpf(1) / (2 * m + 1)
    T = mpmath.mpf(T)
    return mpmath.gammainc(m + mpmath.mpf(1) / 2, 0, T) / (2 * T ** (m + mpmath.mpf(1) / 2))


@pytest.mark.parametrize("T", [0.0, 1e-12, 1e-6, 0.03, 0.7, 2.0, 5.5, 11.0, 19.9, 30.0, 36.5, 44.0, 80.0, 400.0])
def test_boys_vs_incomplete_gamma(T):
    F = oracle.boys(12, T)
    for m in range(13):
        ref = float(boys_mp(m, T))
        assert abs(F[m] - ref) <= 2e-14 * max(ref, 1e-300) + 1e-300, (m, T, F[m], ref)


def test_boys_special_cases():
    F0 = oracle.boys(8, 0.0)
    assert np.allclose(F0, [1.0 / (2 * m + 1) for m in range(9)], rtol=0, atol=1e-15)
    for T in (0.01, 0.5, 3.0, 17.0, 60.0):
        erf_form = math.sqrt(math.pi) * math.erf(math.sqrt(T)) / (2 * math.sqrt(T))
        assert abs(oracle.boys(0, T)[0] - erf_form) < 1e-14
        # downward recursion identity F_m = (2T F_{m+1} + e^-T) / (2m+1)
        F = oracle.boys(10, T)
        for m in range(10):
            assert abs(F[m] - (2 * T * F[m + 1] + math.exp(-T)) / (2 * m + 1)) < 1e-14 * F[m] + 1e-300


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("T", [0.0, 1e-9, 0.2, 1.0, 3.7, 9.0, 15.0, 24.0, 33.0, 41.0, 55.0, 90.0, 250.0])
def test_rys_moment_exactness(n, T):
    """An n-point Gauss rule for the weight exp(-T t^2) on [0,1] (in x = t^2)
    integrates x^k exactly for k < 2n: sum_i w_i x_i^k = F_k(T)."""
    x, w = oracle.rys(n, T)
    assert np.all(x > 0) and np.all(x < 1) and np.all(w > 0)
    assert np.all(np.diff(x) > 0)
    for k in range(2 * n):
        ref = float(boys_mp(k, T))
        got = float(np.sum(w * x ** k))
        assert abs(got - ref) <= 5e-14 * ref, (n, T, k, got, ref)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_large_T_hermite_limit(n):
    """T -> inf: T*t_i^2 -> squares of the positive roots of H_2n; for n=2
    these are (3 -+ sqrt 6)/2 (the rys.c:433-434 asymptote constants)."""
    h, hw = np.polynomial.hermite.hermgauss(2 * n)
    order = np.argsort(h[h > 0])
    hpos, hwpos = h[h > 0][order], hw[h > 0][order]
    T = 4.0e4
    x, w = oracle.rys(n, T)
    assert np.allclose(T * x, hpos ** 2, rtol=2e-4, atol=0)
    if n == 2:
        assert np.allclose(np.sort(hpos ** 2), [(3 - math.sqrt(6)) / 2, (3 + math.sqrt(6)) / 2], rtol=1e-14)
    # weights: w_i -> hw_i / sqrt(T) (half-line Hermite rule), summing to F_0(T) ~ sqrt(pi/T)/2
    assert np.allclose(w * math.sqrt(T), hwpos, rtol=2e-4)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_small_T_legendre_limit(n):
    """T = 0: Gauss-Legendre on [-1,1] with 2n points folded onto [0,1]."""
    g, gw = np.polynomial.legendre.leggauss(2 * n)
    pos = g > 0
    order = np.argsort(g[pos])
    x, w = oracle.rys(n, 0.0)
    assert np.allclose(x, g[pos][order] ** 2, rtol=1e-13, atol=0)
    assert np.allclose(w, gw[pos][order], rtol=1e-13, atol=0)


def _E_gpt(l1, l2, a, b, Ax, Bx):
    """Hermite expansion coefficients E^{ij}_t (Helgaker-Jorgensen-Olsen
    9.5.6-9.5.7) for i<=l1, j<=l2, by the textbook recurrence in mpmath."""
    p = a + b
    mu = a * b / p
    X = Ax - Bx
    Px = (a * Ax + b * Bx) / p
    XPA, XPB = Px - Ax, Px - Bx
    E = {}
    E[(0, 0, 0)] = mpmath.e ** (-mu * X * X)

    def get(i, j, t):
        if t < 0 or t > i + j or i < 0 or j < 0:
            return mpmath.mpf(0)
        return E[(i, j, t)]

    for i in range(l1 + 1):
        for j in range(l2 + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = get(i - 1, j, t - 1) / (2 * p) + XPA * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1)
                else:
                    v = get(i, j - 1, t - 1) / (2 * p) + XPB * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1)
                E[(i, j, t)] = v
    return E


EMAX = 8      # or_hermite_e layout: d[(i*EMAX + j)*(2*EMAX+1) + n]
RMAX = 13     # or_hermite_r layout: r[(t*RMAX + u)*RMAX + v]


@pytest.mark.parametrize("l1,l2", [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 2)])
def test_hermite_e_vs_recurrence(l1, l2):
    """or_hermite_e (mdd.c:3-43) against the textbook recurrence.  The
    reference's E carries no Gaussian exponential factor (it is applied in
    the prefactor), so E^{00}_0 = 1 there."""
    a, b, Ax, Bx = 1.3, 0.45, 0.2, -0.9
    p = a + b
    d = np.zeros(EMAX * EMAX * (2 * EMAX + 1))
    oracle.lib().or_hermite_e(d, l1, l2, 1.0 / (2 * p), (a * Ax + b * Bx) / p - Ax, (a * Ax + b * Bx) / p - Bx)
    d = d.reshape(EMAX, EMAX, 2 * EMAX + 1)
    E = _E_gpt(l1, l2, mpmath.mpf(a), mpmath.mpf(b), mpmath.mpf(Ax), mpmath.mpf(Bx))
    scale = E[(0, 0, 0)]
    assert d[0, 0, 0] == 1.0
    for i in range(l1 + 1):
        for j in range(l2 + 1):
            for t in range(i + j + 1):
                assert abs(d[i, j, t] - float(E[(i, j, t)] / scale)) < 1e-14, (i, j, t)


@pytest.mark.parametrize("ltot", [0, 2, 4])
def test_hermite_r_vs_derivatives(ltot):
    """or_hermite_r (mdr.c:37-2163): R_tuv = sr * d^t/dX d^u/dY d^v/dZ
    F_0(w |PQ|^2), checked by mpmath differentiation of the closed form
    F_0(T) = sqrt(pi) erf(sqrt T) / (2 sqrt T)."""
    w, sr = 0.8, 1.7
    pq = np.array([0.3, -0.5, 0.8])
    T = w * float(pq @ pq)
    r = np.zeros(RMAX ** 3)
    oracle.lib().or_hermite_r(r, ltot, sr, T, w, pq)
    r = r.reshape(RMAX, RMAX, RMAX)

    def f(x, y, z):
        t = w * (x * x + y * y + z * z)
        return mpmath.sqrt(mpmath.pi) * mpmath.erf(mpmath.sqrt(t)) / (2 * mpmath.sqrt(t))

    for t in range(ltot + 1):
        for u in range(ltot + 1 - t):
            for v in range(ltot + 1 - t - u):
                ref = sr * float(mpmath.diff(f, tuple(mpmath.mpf(c) for c in pq), (t, u, v)))
                assert abs(r[t, u, v] - ref) < 1e-13 * max(1.0, abs(ref)), (t, u, v, r[t, u, v], ref)
