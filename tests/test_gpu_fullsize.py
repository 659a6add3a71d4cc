"""The G build at the benchmark's full size, (H2O)8 on a 2x2x2 lattice
(232 shells, 448 functions, 138M Schwarz-kept shell quartets), on the GPU:

* against the oracle's direct build of the same screened quartet set (the
  reference's Rys algorithm and canonical loop, twoints.c:188-418 +
  scfaux.c:428-552, restated in C and run on the host's cores: about 10 s
  on 16 threads), element by element;
* linearity in the density, G(P1 + 2 P2) = G(P1) + 2 G(P2);
* the symmetry of the two-electron form, tr(P1 G(P2)) = tr(P2 G(P1)), which
  a wrong index or component order in any one class kernel would break;
* the mixed-precision build (work items whose bound x max|P| is under the
  threshold run the FP32 kernels, FP64 accumulation) against the FP64 build,
  at the benchmark's threshold 1e-6 and at 1e-3 (most items in FP32).

Tolerances (measured: 5e-14, 5e-15, 2e-16): 1e-11 absolute on G elements
against the oracle (|G| up to ~3, sums of up to 1e7 integrals in another
order; both screen with τ = 1e-12 on Schwarz bounds computed separately),
1e-12 absolute on linearity (FP64 atomics make every build's summation order
different), 1e-13 relative on the bilinear traces.
"""
import os

import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note

pytestmark = pytest.mark.gpu

TAU = 1e-12


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.1
    return (A @ A.T / n)[np.tril_indices(n)].copy()


def _trace(Pa, Gb, n):
    # tr(Pa Gb) over packed lower triangles: off-diagonal terms twice
    w = np.full(Pa.shape, 2.0)
    w[np.arange(n) * (np.arange(n) + 3) // 2] = 1.0
    return float(np.sum(w * Pa * Gb))


@pytest.fixture(scope="module")
def water8_store():
    s = systems.water8()
    pm.load_system(s)
    pm.set_options(tau=TAU, precision=0, strict=False)
    st = pm.EriStore()
    yield s, st
    st.close()


def test_water8_gmatrix_vs_oracle_direct(water8_store):
    s, st = water8_store
    P = _density(s.norb, 11)
    G = st.rhf_gmatrix(P)
    threads = min(16, os.cpu_count() or 1)
    b = oracle.Basis(s)
    q = b.pair_schwarz(nthreads=threads)
    Gref, done = b.direct_gmatrix_sampled(P, q, TAU, 1, nthreads=threads)
    d = float(np.abs(G - Gref).max())
    note(f"(H2O)8 G: {done} oracle shell quartets, max|G| {np.abs(Gref).max():.2f}, max|dG| {d:.2e}")
    assert d < 1e-11


def test_water8_gmatrix_linear_and_symmetric(water8_store):
    s, st = water8_store
    n = s.norb
    P1, P2 = _density(n, 12), _density(n, 13)
    G1, G2 = st.rhf_gmatrix(P1), st.rhf_gmatrix(P2)
    G12 = st.rhf_gmatrix(P1 + 2.0 * P2)
    lin = float(np.abs(G12 - (G1 + 2.0 * G2)).max())
    t12, t21 = _trace(P1, G2, n), _trace(P2, G1, n)
    rel = abs(t12 - t21) / abs(t12)
    note(f"(H2O)8 linearity max|dG| {lin:.2e}; tr(P1 G(P2)) {t12:.12f} vs tr(P2 G(P1)) {t21:.12f} (rel {rel:.1e})")
    assert lin < 1e-12
    assert rel < 1e-13


@pytest.mark.parametrize("thr,tol", [(1e-6, 1e-9), (1e-3, 1e-5)])
def test_water8_mixed_precision_vs_fp64(water8_store, thr, tol):
    """Tolerances: an FP32 item's integrals carry ~1e-7 relative error on
    contributions bounded by thr x |P| each (measured: 9.6e-11 with 23,873
    FP32 items at 1e-6, 1.4e-7 with 86,880 at 1e-3)."""
    s, st = water8_store
    P = _density(s.norb, 14)
    G64 = st.rhf_gmatrix(P)
    pm.set_options(precision=1, mixed_threshold=thr)
    try:
        Gmx = st.rhf_gmatrix(P)
        n32 = st.stats()["work_items_fp32"]
    finally:
        pm.set_options(precision=0)
    d = float(np.abs(Gmx - G64).max())
    note(f"(H2O)8 mixed precision, threshold {thr:g}: {n32} FP32 work items, max|G - G_fp64| {d:.2e}")
    assert n32 > 0
    assert d < tol
