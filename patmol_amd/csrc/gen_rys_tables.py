"""Generate rys_table.inc: piecewise-Chebyshev fits of the Rys quadrature.

For n = 1..NMAX roots and the weight e^{-T t^2} on t in [0,1], the nodes
x_i = t_i^2 and weights w_i (sum w_i = F_0(T)) are tabulated on unit
intervals T in [k, k+1), k < TMAX(n), as degree-DEG Chebyshev series in
u = 2(T-k)-1.  For T >= TMAX(n) the device uses the large-T limit
x_i = r_i/T, w_i = h_i/sqrt(T), where +/-sqrt(r_i) are the roots of the
Hermite polynomial H_{2n} and h_i the matching Gauss-Hermite weights.

Exact values come from the Golub-Welsch eigenproblem of the Jacobi matrix
built (Chebyshev algorithm, Gautschi 2004 sec. 2.1.7) from the moments
F_k(T) = gamma(k+1/2, T) / (2 T^(k+1/2)), all at 40 decimal digits with
mpmath, so the tables are limited only by the final rounding to double.

Usage: python gen_rys_tables.py [out.inc]   (takes ~1 minute on 8 cores)
"""
import multiprocessing as mp_
import os
import sys

import mpmath as mp
import numpy as np

NMAX = 7
DEG = 11
NC = DEG + 1


def tmax(n):
    return 36 + 8 * (n - 1)


def moments(T, K):
    T = mp.mpf(T)
    if T == 0:
        return [mp.mpf(1) / (2 * k + 1) for k in range(K)]
    return [mp.gammainc(k + mp.mpf(1) / 2, 0, T) / (2 * T ** (k + mp.mpf(1) / 2)) for k in range(K)]


def rule_from_moments(mom, n):
    a = [mp.mpf(0)] * n
    b = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mom[: 2 * n])
    a[0] = mom[1] / mom[0]
    b[0] = mom[0]
    for k in range(1, n):
        sn = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sn[l] = sig[l + 1] - a[k - 1] * sig[l] - b[k - 1] * sig_prev[l]
        a[k] = sn[k + 1] / sn[k] - sig[k] / sig[k - 1]
        b[k] = sn[k] / sig[k - 1]
        sig_prev, sig = sig, sn
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = a[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(b[i + 1])
    E, Q = mp.eigsy(J)
    o = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in o], [b[0] * Q[0, i] ** 2 for i in o]


def exact_all(T):
    mp.mp.dps = 40
    mom = moments(T, 2 * NMAX)
    out = {}
    for n in range(1, NMAX + 1):
        x, w = rule_from_moments(mom, n)
        out[n] = ([float(v) for v in x], [float(v) for v in w])
    return out


def hermite_limit(n):
    """r_i (ascending) and h_i for the large-T rule."""
    mp.mp.dps = 40
    m = 2 * n
    # Gauss-Hermite via Golub-Welsch: Jacobi matrix with off-diagonal sqrt(k/2)
    J = mp.matrix(m, m)
    for k in range(1, m):
        J[k, k - 1] = J[k - 1, k] = mp.sqrt(mp.mpf(k) / 2)
    E, Q = mp.eigsy(J)
    pts = [(E[i], mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0]
    pts.sort(key=lambda p: p[0])
    return [float(p[0] ** 2) for p in pts], [float(p[1]) for p in pts]


def cheb_nodes():
    k = np.arange(NC)
    return np.cos(np.pi * (k + 0.5) / NC)


def main(out_path):
    nodes = cheb_nodes()
    kmax = tmax(NMAX)
    Ts = [k + (u + 1.0) / 2.0 for k in range(kmax) for u in nodes]
    with mp_.Pool(min(8, os.cpu_count() or 1)) as pool:
        vals = pool.map(exact_all, Ts, chunksize=8)
    # coefficients: table[n][k][c][v], v = (x_0..x_{n-1}, w_0..w_{n-1})
    tables = {}
    for n in range(1, NMAX + 1):
        tab = np.zeros((tmax(n), NC, 2 * n))
        for k in range(tmax(n)):
            f = np.array([vals[k * NC + j][n][0] + vals[k * NC + j][n][1] for j in range(NC)])
            for c in range(NC):
                tab[k, c] = 2.0 / NC * np.sum(f * np.cos(c * np.arccos(nodes))[:, None], axis=0)
            tab[k, 0] *= 0.5
        tables[n] = tab
    herm = {n: hermite_limit(n) for n in range(1, NMAX + 1)}

    # verification against fresh exact values, Clenshaw in double as on device
    rng = np.random.default_rng(7)
    worst = 0.0
    for n in range(1, NMAX + 1):
        for T in list(rng.uniform(0, tmax(n), 40)) + [0.0, 1e-12, tmax(n) - 1e-9, tmax(n), tmax(n) + 3.0]:
            ex = exact_all(T)[n]
            exv = np.array(ex[0] + ex[1])
            got = evaluate(tables, herm, n, T)
            worst = max(worst, float(np.max(np.abs(got - exv) / np.abs(exv))))
    print("max relative error over checks: %.3e" % worst)
    if worst > 1e-14:
        raise SystemExit("table accuracy check failed")

    offs, flat = [], []
    for n in range(1, NMAX + 1):
        offs.append(len(flat))
        flat.extend(tables[n].reshape(-1).tolist())
    with open(out_path, "w") as f:
        f.write("// Generated by gen_rys_tables.py -- do not edit.\n")
        f.write("// Includers define RYS_STORAGE (e.g. static const, or __device__ const).\n")
        f.write("// Rys nodes x=t^2 and weights: Chebyshev fits (degree %d) on unit T intervals.\n" % DEG)
        f.write("// max relative error vs 40-digit Golub-Welsch: %.2e\n" % worst)
        f.write("#define RYS_NMAX %d\n#define RYS_NC %d\n" % (NMAX, NC))
        f.write("RYS_STORAGE int rys_tmax[%d] = {0, %s};\n" % (NMAX + 1, ", ".join(str(tmax(n)) for n in range(1, NMAX + 1))))
        f.write("RYS_STORAGE int rys_off[%d] = {0, %s};\n" % (NMAX + 1, ", ".join(str(o) for o in offs)))
        f.write("#define RYS_TABLE_SIZE %d\n" % len(flat))
        f.write("RYS_STORAGE double rys_cheb[RYS_TABLE_SIZE] = {\n")
        for i in range(0, len(flat), 4):
            f.write("  " + ", ".join(float.hex(v) for v in flat[i:i + 4]) + ",\n")
        f.write("};\n")
        rs = [v for n in range(1, NMAX + 1) for v in herm[n][0]]
        hs = [v for n in range(1, NMAX + 1) for v in herm[n][1]]
        f.write("// large-T limit: r_i and h_i for n = 1..%d, offset n(n-1)/2\n" % NMAX)
        f.write("RYS_STORAGE double rys_herm_r[%d] = {%s};\n" % (len(rs), ", ".join(float.hex(v) for v in rs)))
        f.write("RYS_STORAGE double rys_herm_h[%d] = {%s};\n" % (len(hs), ", ".join(float.hex(v) for v in hs)))
    print("wrote", out_path, len(flat), "doubles")


def evaluate(tables, herm, n, T):
    if T >= tmax(n):
        r, h = herm[n]
        return np.array([ri / T for ri in r] + [hi / np.sqrt(T) for hi in h])
    k = int(T)
    u = 2.0 * (T - k) - 1.0
    c = tables[n][k]
    b1 = np.zeros(2 * n)
    b2 = np.zeros(2 * n)
    for j in range(NC - 1, 0, -1):
        b1, b2 = 2.0 * u * b1 - b2 + c[j], b1
    return u * b1 - b2 + c[0]


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table.inc"))
