"""Independent brute-force ERIs for the CPU tests (not collected).

Textbook McMurchie-Davidson (Helgaker, Jorgensen, Olsen ch. 9): Hermite
expansion coefficients by the E recurrence, Hermite integrals R_tuv by the
R recurrence started from Boys values evaluated in mpmath, and the integral
(ab|cd) = 2 pi^{5/2}/(pq sqrt(p+q)) sum E^ab_tuv (-1)^{tau+nu+phi} E^cd R.
Normalisation is imposed from the definitions: contraction coefficients
multiply normalised primitives (primitive norm proportional to
alpha^((2l+3)/4)), and every Cartesian component of every contracted shell
has unit self-overlap (the convention that shell.c:47-77 and
basis.c:309-319 implement).  Nothing here calls the oracle.
"""
import math

import mpmath
import numpy as np


def cart_components(l):
    """Cartesian exponents, lx descending then ly descending (basis.c:294-308)."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def e_coefs(la, lb, a, b, Ax, Bx):
    """E[i, j, t] including the factor exp(-ab/(a+b) (Ax-Bx)^2)."""
    p = a + b
    Px = (a * Ax + b * Bx) / p
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-a * b / p * (Ax - Bx) ** 2)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    src, X = E[i - 1, j], Px - Ax
                else:
                    src, X = E[i, j - 1], Px - Bx
                v = X * src[t] + (t + 1) * src[t + 1]
                if t > 0:
                    v += src[t - 1] / (2 * p)
                E[i, j, t] = v
    return E


def boys_list(nmax, T):
    if T < 1e-300:
        return [1.0 / (2 * m + 1) for m in range(nmax + 1)]
    Tm = mpmath.mpf(T)
    return [float(mpmath.gammainc(m + mpmath.mpf(0.5), 0, Tm) / (2 * Tm ** (m + mpmath.mpf(0.5))))
            for m in range(nmax + 1)]


def r_tensor(L, alpha, PQ):
    """R[t, u, v] = R^0_tuv(alpha, PQ) for t+u+v <= L."""
    T = alpha * float(PQ @ PQ)
    F = boys_list(L, T)
    Rn = np.zeros((L + 1, L + 1, L + 1, L + 1))
    for n in range(L + 1):
        Rn[n, 0, 0, 0] = (-2 * alpha) ** n * F[n]
    for n in range(L - 1, -1, -1):
        for t in range(L - n + 1):
            for u in range(L - n + 1 - t):
                for v in range(L - n + 1 - t - u):
                    if t + u + v == 0:
                        continue
                    if t > 0:
                        val = PQ[0] * Rn[n + 1, t - 1, u, v] + ((t - 1) * Rn[n + 1, t - 2, u, v] if t > 1 else 0.0)
                    elif u > 0:
                        val = PQ[1] * Rn[n + 1, t, u - 1, v] + ((u - 1) * Rn[n + 1, t, u - 2, v] if u > 1 else 0.0)
                    else:
                        val = PQ[2] * Rn[n + 1, t, u, v - 1] + ((v - 1) * Rn[n + 1, t, u, v - 2] if v > 1 else 0.0)
                    Rn[n, t, u, v] = val
    return Rn[0]


def prim_coefs(shell):
    """Raw contraction coefficients times the primitive normalisation
    (up to an l-dependent constant, which the contraction norm removes)."""
    return [c * a ** ((2 * shell.l + 3) / 4.0) for a, c in zip(shell.al, shell.co)]


def _overlap_prim(la, lb, a, b, A, B):
    p = a + b
    out = np.zeros((len(cart_components(la)), len(cart_components(lb))))
    Es = [e_coefs(la, lb, a, b, A[x], B[x]) for x in range(3)]
    for i, ca in enumerate(cart_components(la)):
        for j, cb in enumerate(cart_components(lb)):
            out[i, j] = (math.pi / p) ** 1.5 * np.prod([Es[x][ca[x], cb[x], 0] for x in range(3)])
    return out


def component_norms(shell, center):
    """1/sqrt(<phi_i|phi_i>) for each Cartesian component of a contracted
    shell built from the raw coefficients."""
    S = np.zeros((shell.nst, shell.nst))
    co = prim_coefs(shell)
    for a, ca in zip(shell.al, co):
        for b, cb in zip(shell.al, co):
            S += ca * cb * _overlap_prim(shell.l, shell.l, a, b, center, center)
    return 1.0 / np.sqrt(np.diag(S))


def _hermite_block(la, lb, a, b, A, B):
    """(ncomp_a*ncomp_b, (L+1)^3) Hermite coefficients E^{ab}_{tuv}."""
    L = la + lb
    Es = [e_coefs(la, lb, a, b, A[x], B[x]) for x in range(3)]
    ca, cb = cart_components(la), cart_components(lb)
    out = np.zeros((len(ca) * len(cb), (L + 1) ** 3))
    for i, ea in enumerate(ca):
        for j, eb in enumerate(cb):
            ex = Es[0][ea[0], eb[0], :L + 1]
            ey = Es[1][ea[1], eb[1], :L + 1]
            ez = Es[2][ea[2], eb[2], :L + 1]
            out[i * len(cb) + j] = np.einsum("t,u,v->tuv", ex, ey, ez).ravel()
    return out


def eri_block(system, I, J, K, L):
    """Contracted, normalised (IJ|KL) block, shape (nst_I, nst_J, nst_K, nst_L)."""
    sh = [system.shells[x] for x in (I, J, K, L)]
    R = [np.asarray(system.coords[s.cen], dtype=float) for s in sh]
    l = [s.l for s in sh]
    L12, L34 = l[0] + l[1], l[2] + l[3]
    Lt = L12 + L34
    sign = np.array([(-1) ** (t + u + v) for t in range(L34 + 1) for u in range(L34 + 1) for v in range(L34 + 1)],
                    dtype=float)
    n = [s.nst for s in sh]
    out = np.zeros((n[0] * n[1], n[2] * n[3]))
    i12 = np.arange(L12 + 1)
    i34 = np.arange(L34 + 1)
    co = [prim_coefs(x) for x in sh]
    for a, ca in zip(sh[0].al, co[0]):
        for b, cb in zip(sh[1].al, co[1]):
            p = a + b
            P = (a * R[0] + b * R[1]) / p
            Eab = _hermite_block(l[0], l[1], a, b, R[0], R[1])
            for c, cc in zip(sh[2].al, co[2]):
                for d, cd in zip(sh[3].al, co[3]):
                    q = c + d
                    Q = (c * R[2] + d * R[3]) / q
                    Ecd = _hermite_block(l[2], l[3], c, d, R[2], R[3]) * sign
                    Rt = r_tensor(Lt, p * q / (p + q), P - Q)
                    # H[tuv, tau nu phi] = R[t+tau, u+nu, v+phi]
                    H = Rt[i12[:, None, None, None, None, None] + i34[None, None, None, :, None, None],
                           i12[None, :, None, None, None, None] + i34[None, None, None, None, :, None],
                           i12[None, None, :, None, None, None] + i34[None, None, None, None, None, :]]
                    H = H.reshape((L12 + 1) ** 3, (L34 + 1) ** 3)
                    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * ca * cb * cc * cd
                    out += pref * (Eab @ H @ Ecd.T)
    out = out.reshape(n)
    for x in range(4):
        nv = component_norms(sh[x], R[x])
        shape = [1, 1, 1, 1]
        shape[x] = n[x]
        out = out * nv.reshape(shape)
    return out


def overlap_block(system, I, J):
    sI, sJ = system.shells[I], system.shells[J]
    A, B = np.asarray(system.coords[sI.cen], float), np.asarray(system.coords[sJ.cen], float)
    S = np.zeros((sI.nst, sJ.nst))
    for a, ca in zip(sI.al, prim_coefs(sI)):
        for b, cb in zip(sJ.al, prim_coefs(sJ)):
            S += ca * cb * _overlap_prim(sI.l, sJ.l, a, b, A, B)
    return S * component_norms(sI, A)[:, None] * component_norms(sJ, B)[None, :]
