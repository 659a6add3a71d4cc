"""Oracle ERIs, canonical lists and the G scatter, pinned on the CPU.

or_eri_block (twoints.c:23-186 with rys.c:2269-2335) against an
independent McMurchie-Davidson brute force over every class to (dd|dd),
the contracted (ss|ss) closed form, permutational symmetry, translation
and rotation invariance and the Schwarz inequality; the canonical filter
of twoints.c:316-337 against closed-form counts; the G scatter of
scfaux.c:428-714 against einsum over the full four-index tensor.
"""
import math
import os

import numpy as np
import pytest

import oracle
from patmol_amd import systems
from tests import md_brute


def _by_l(s):
    out = {}
    for i, sh in enumerate(s.shells):
        out.setdefault(sh.l, []).append(i)
    return out


@pytest.fixture(scope="module")
def water_d():
    s = systems.water(d_shells=True)
    return s, oracle.Basis(s)


def test_eri_blocks_vs_md_brute_force_all_classes(water_d):
    s, b = water_d
    byl = _by_l(s)
    rng = np.random.default_rng(5)
    worst = 0.0
    for la in range(3):
        for lb in range(3):
            for lc in range(3):
                for ld in range(3):
                    for _ in range(2):
                        I, J, K, L = (int(rng.choice(byl[x])) for x in (la, lb, lc, ld))
                        ref = md_brute.eri_block(s, I, J, K, L)
                        got = b.eri_block(I, J, K, L)
                        assert got.shape == ref.shape
                        worst = max(worst, float(np.abs(got - ref).max()))
    assert worst < 1e-13


def test_eri_blocks_contracted_sf6_vs_md_brute_force(golden_dir):
    """Contracted shells (SF6 input, contraction up to 10 primitives)."""
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    b = oracle.Basis(s)
    byl = _by_l(s)
    rng = np.random.default_rng(9)
    classes = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 1, 1, 0), (2, 0, 1, 0), (2, 1, 1, 1), (2, 2, 0, 0)]
    for cls in classes:
        I, J, K, L = (int(rng.choice(byl[x])) for x in cls)
        npq = np.prod([s.shells[x].npr for x in (I, J, K, L)])
        if npq > 3000:
            continue
        ref = md_brute.eri_block(s, I, J, K, L)
        got = b.eri_block(I, J, K, L)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), cls


def test_ssss_closed_form():
    """Contracted (ss|ss) = sum c c c c 2 pi^(5/2)/(pq sqrt(p+q)) K_ab K_cd F0(T)."""
    ang = systems.ANGSTROM
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.1 * ang], [0.7, -0.4, 0.3]])
    al = [[3.1, 0.6, 0.15], [1.2, 0.3], [0.8]]
    co = [[0.2, 0.5, 0.4], [0.6, 0.5], [1.0]]
    shells = [systems.Shell(0, c, al[c], co[c]) for c in range(3)]
    s = systems.System(np.array([1.0, 1.0, 1.0]), xyz, shells, nelec=2)
    b = oracle.Basis(s)
    cn = []
    for sh in shells:       # normalised contraction coefficients
        a = np.array(sh.al)
        c = np.array(sh.co) * (2 * a / math.pi) ** 0.75
        nrm = sum(c[i] * c[j] * (math.pi / (a[i] + a[j])) ** 1.5 for i in range(len(a)) for j in range(len(a)))
        cn.append(c / math.sqrt(nrm))

    def F0(T):
        return 1.0 if T < 1e-15 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))

    def ssss(I, J, K, L):
        v = 0.0
        R = [xyz[shells[x].cen] for x in (I, J, K, L)]
        for a, ca in zip(shells[I].al, cn[I]):
            for bb, cb in zip(shells[J].al, cn[J]):
                p = a + bb
                P = (a * R[0] + bb * R[1]) / p
                Kab = math.exp(-a * bb / p * np.sum((R[0] - R[1]) ** 2))
                for c, cc in zip(shells[K].al, cn[K]):
                    for d, cd in zip(shells[L].al, cn[L]):
                        q = c + d
                        Q = (c * R[2] + d * R[3]) / q
                        Kcd = math.exp(-c * d / q * np.sum((R[2] - R[3]) ** 2))
                        T = p * q / (p + q) * np.sum((P - Q) ** 2)
                        v += ca * cb * cc * cd * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0(T)
        return v

    for q4 in [(0, 0, 0, 0), (1, 0, 0, 0), (2, 1, 1, 0), (2, 0, 2, 1), (1, 1, 2, 2)]:
        assert abs(b.eri_block(*q4).item() - ssss(*q4)) < 1e-14


def test_permutation_symmetry_and_schwarz(water_d):
    s, b = water_d
    rng = np.random.default_rng(7)
    for _ in range(60):
        I, J, K, L = (int(x) for x in rng.integers(0, s.nshell, 4))
        v = b.eri_block(I, J, K, L)
        assert np.allclose(v, b.eri_block(J, I, K, L).transpose(1, 0, 2, 3), atol=1e-14, rtol=0)
        assert np.allclose(v, b.eri_block(I, J, L, K).transpose(0, 1, 3, 2), atol=1e-14, rtol=0)
        assert np.allclose(v, b.eri_block(K, L, I, J).transpose(2, 3, 0, 1), atol=1e-14, rtol=0)
        # the primitive screen (sr < 1e-15, twoints.c:110) can zero a whole
        # diagonal block; positivity holds for the unscreened integrals
        dIJ = np.einsum("ijij->ij", b.eri_block_unscreened(I, J, I, J))
        dKL = np.einsum("klkl->kl", b.eri_block_unscreened(K, L, K, L))
        assert np.all(dIJ > 0) and np.all(dKL > 0)
        assert np.all(v ** 2 <= dIJ[:, :, None, None] * dKL[None, None, :, :] * (1 + 1e-12) + 1e-28)


def test_translation_and_rotation_invariance():
    """Translation leaves every block unchanged; a rotation acts on p
    components as an orthogonal matrix (s/p basis)."""
    s = systems.water()
    b = oracle.Basis(s)
    s2 = systems.water()
    s2.coords = s.coords + np.array([0.7, -1.3, 2.2])
    b2 = oracle.Basis(s2)
    rot = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))[0]
    s3 = systems.water()
    s3.coords = s.coords @ rot.T
    b3 = oracle.Basis(s3)
    rng = np.random.default_rng(2)
    byl = _by_l(s)
    for _ in range(20):
        I, K = int(rng.choice(byl[1])), int(rng.choice(byl[1]))
        J, L = (int(x) for x in rng.integers(0, s.nshell, 2))
        v = b.eri_block(I, J, K, L)
        assert np.allclose(v, b2.eri_block(I, J, K, L), atol=1e-13, rtol=0)
        w = b3.eri_block(I, J, K, L)
        # rotate back: components of I and K are vectors (x, y, z order)
        vr = np.einsum("ia,kc,ajcl->ijkl", rot.T, rot.T, w)
        if s.shells[J].l == 1:
            vr = np.einsum("jb,ibkl->ijkl", rot.T, vr)
        if s.shells[L].l == 1:
            vr = np.einsum("ld,ijkd->ijkl", rot.T, vr)
        assert np.allclose(v, vr, atol=1e-12, rtol=0)


def _loop_tuples(ns):
    return sum((i + 1) * (i + 1) * (i + 2) // 2 for i in range(ns))


def test_canonical_counts_water():
    """twoints.c:228-337 on the 26-shell water: 64,701 loop tuples, and the
    canonical filter keeps each unique integral once: M(M+1)/2 with
    M = N(N+1)/2 = 741 -> 274,911."""
    s = systems.water()
    b = oracle.Basis(s)
    loops, shellq, ints = b.eri_count()
    N = s.norb
    M = N * (N + 1) // 2
    assert (loops, ints) == (64701, 274911)
    assert loops == _loop_tuples(s.nshell)
    assert ints == M * (M + 1) // 2
    # shell quartets that own at least one canonical integral, by brute force
    off = s.offsets()
    nst = [sh.nst for sh in s.shells]
    n = 0
    for I in range(s.nshell):
        for J in range(I + 1):
            for K in range(I + 1):
                for L in range(K + 1):
                    ok = False
                    for i in range(off[I], off[I] + nst[I]):
                        for j in range(off[J], off[J] + nst[J]):
                            if j > i:
                                continue
                            for k in range(off[K], off[K] + nst[K]):
                                if k > i:
                                    continue
                                for l in range(off[L], off[L] + nst[L]):
                                    if l <= k and not (i == k and l > j):
                                        ok = True
                                        break
                                if ok:
                                    break
                            if ok:
                                break
                        if ok:
                            break
                    n += ok
    assert shellq == n


@pytest.mark.slow
def test_canonical_counts_water8():
    """(H2O)8, 232 shells, N = 448: 367,351,062 loop tuples and M(M+1)/2
    unique integrals.  The loop visits more shell quartets with a canonical
    component than there are unique shell quartets (27028*27029/2 =
    365,269,906): for I == K and J < L <= I the tuple (IJ|IL) still owns the
    components with k < i."""
    s = systems.water8()
    b = oracle.Basis(s)
    loops, shellq, ints = b.eri_count()
    M = s.norb * (s.norb + 1) // 2
    npair = s.nshell * (s.nshell + 1) // 2
    assert loops == _loop_tuples(s.nshell) == 367351062
    assert ints == M * (M + 1) // 2
    assert shellq > npair * (npair + 1) // 2 == 365269906


def test_two_electron_list_start_rows():
    """ERINT list restricted to ish >= start (twoints.c:228) is the tail of
    the full list, and every record is canonical."""
    s = systems.water()
    b = oracle.Basis(s)
    full = b.two_electron_ints().array()
    off = s.offsets()
    start = 14
    tail = b.two_electron_ints(start=start).array()
    keep = full[full["i"] >= off[start]]
    assert len(tail) == len(keep)
    assert np.array_equal(np.sort(tail["val"]), np.sort(keep["val"]))
    i, j, k, l = (full[x].astype(np.int64) for x in "ijkl")
    assert np.all(j <= i) and np.all(l <= k) and np.all(k <= i)
    assert np.all(~((i == k) & (l > j)))
    assert np.all(np.abs(full["val"]) > 1e-14)      # write filter twoints.c:375


def _full(p, n):
    return oracle.packed_to_full(p, n)


def _rand_density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.3
    return A @ A.T / n


def _tensor_from_list(arr, n, b, s):
    """Full (ij|kl) tensor by 8-fold symmetric fill of the canonical list;
    spot-checked against shell blocks."""
    T = np.zeros((n, n, n, n))
    i, j, k, l = (arr[x].astype(np.int64) for x in "ijkl")
    v = arr["val"]
    for a, bb, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                        (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
        T[a, bb, c, d] = v
    off = s.offsets()
    rng = np.random.default_rng(12)
    for _ in range(40):
        I, J, K, L = (int(x) for x in rng.integers(0, s.nshell, 4))
        blk = b.eri_block(I, J, K, L)
        sub = T[off[I]:off[I] + blk.shape[0], off[J]:off[J] + blk.shape[1],
                off[K]:off[K] + blk.shape[2], off[L]:off[L] + blk.shape[3]]
        # equal up to rounding and the |v| > 1e-14 write filter (twoints.c:375)
        assert np.abs(sub - blk).max() < 2e-14
    return T


def test_rhf_and_uhf_scatter_vs_einsum(water_d):
    s, b = water_d
    n = s.norb
    ints = b.two_electron_ints()
    T = _tensor_from_list(ints.array(), n, b, s)
    P = _rand_density(n, 3)
    J = np.einsum("ijkl,kl->ij", T, P)
    K = np.einsum("ikjl,kl->ij", T, P)
    G = ints.rhf_gmatrix(oracle.full_to_packed(P))
    assert np.abs(_full(G, n) - (2 * J - K)).max() < 1e-12
    # multi-threaded scatter is the same sum
    assert np.abs(ints.rhf_gmatrix(oracle.full_to_packed(P), nthreads=4) - G).max() < 1e-12
    PA, PB = _rand_density(n, 4), _rand_density(n, 5)
    GA, GB = ints.uhf_gmatrix(oracle.full_to_packed(PA), oracle.full_to_packed(PB))
    JT = np.einsum("ijkl,kl->ij", T, PA + PB)
    assert np.abs(_full(GA, n) - (JT - np.einsum("ikjl,kl->ij", T, PA))).max() < 1e-12
    assert np.abs(_full(GB, n) - (JT - np.einsum("ikjl,kl->ij", T, PB))).max() < 1e-12


def test_direct_gmatrix_and_rows_match_scatter(water_d):
    s, b = water_d
    n = s.norb
    P = oracle.full_to_packed(_rand_density(n, 8))
    G = b.two_electron_ints().rhf_gmatrix(P)
    Gd, done = b.direct_gmatrix(P, nthreads=4)
    assert np.abs(Gd - G).max() < 1e-12
    rows = [0, 5, 17, s.nshell - 1]
    Gr = b.g_rows(_full(P, n), rows).reshape(-1, n)
    Gf = _full(G, n)
    off = s.offsets()
    sel = np.concatenate([np.arange(off[r], off[r] + s.shells[r].nst) for r in rows])
    assert np.abs(Gr - Gf[sel]).max() < 1e-12


def test_sampled_direct_build_is_a_subset(water_d):
    """The timing form of the direct build (Schwarz screen + tuple sampling)
    reduces to the full build with tau = 0 and no sampling, and its bounds
    match the unscreened diagonal blocks."""
    s, b = water_d
    n = s.norb
    P = oracle.full_to_packed(_rand_density(n, 9))
    q = b.pair_schwarz(nthreads=2)
    I, J = 7, 3
    d = np.einsum("abab->ab", b.eri_block_unscreened(I, J, I, J))
    assert abs(q[I, J] - np.sqrt(np.abs(d).max())) < 1e-15 and q[I, J] == q[J, I]
    G0, n0 = b.direct_gmatrix(P, nthreads=2)
    npair = s.nshell * (s.nshell + 1) // 2
    assert n0 == npair * (npair + 1) // 2      # every canonical shell quartet, counted once
    G1, n1 = b.direct_gmatrix_sampled(P, q, 0.0, 1, nthreads=2)
    assert n0 == n1 and np.abs(G0 - G1).max() < 1e-12
    _, n3 = b.direct_gmatrix_sampled(P, q, 1e-12, 3, nthreads=2)
    assert 0 < n3 < n0 / 2
