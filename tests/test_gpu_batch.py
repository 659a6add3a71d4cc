"""Batched G build and RHF of many molecules on the GPU.

Molecules of the seeded H/C/N/O batch (2-20 atoms, uneven sizes) go into one
batch: one work queue per angular-momentum class over all molecules, the G
build captured as a HIP graph.  Checked: each molecule's G against its own
single-molecule build (bitwise: the J/K sums are fixed point) and against
the oracle's scatter; graph replays, with the program's own work between
them, against the eager launches; the batched SCF's per-molecule energies
against the oracle's SCF (1e-8 Eh), with the molecules leaving the replays
as they converge and every molecule built again afterwards; the
mixed-precision batch against FP64 (|dE| <= 1e-6 Eh).
"""
import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu


def _small_batch(k=8):
    mols = systems.molecule_batch(64, 1002)
    return sorted(mols, key=lambda m: m.norb)[:k]


def _densities(mols, seed):
    return np.concatenate([packed(random_density(m.norb, seed + i, 0.2)) for i, m in enumerate(mols)])


def test_batch_gmatrix_vs_single_and_oracle_and_graph():
    mols = _small_batch(8)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    b = pm.Batch()
    for m in mols:
        b.add(m)
    b.finalize()
    note(f"batch of {len(mols)}: N = {[m.norb for m in mols]}, {b.info()}")
    P = _densities(mols, 100)
    G_eager = b.gmatrix(P, use_graph=False)
    off = b.offsets()
    for i, m in enumerate(mols):
        pm.load_system(m)
        st = pm.EriStore()
        Pi = P[off[i]:off[i + 1]]
        Gs = st.rhf_gmatrix(Pi)
        st.close()
        assert np.array_equal(G_eager[off[i]:off[i + 1]], Gs), f"molecule {i}"
        ref = oracle.Basis(m).two_electron_ints().rhf_gmatrix(Pi)
        d = float(np.abs(Gs - ref).max())
        assert d < 1e-10, (i, d)
    # graph replays with eager work in between: every replay equals the eager build
    for k in range(3):
        G = b.gmatrix(P, use_graph=True)
        assert np.array_equal(G, G_eager), f"replay {k}"
        b.gmatrix(_densities(mols, 200 + k), use_graph=False)
    b.close()


def test_batch_rhf_scf_vs_oracle_and_mixed():
    mols = _small_batch(8)
    pm.set_options(tau=1e-12, precision=0, strict=False, scf_density=1)
    b = pm.Batch()
    for m in mols:
        b.add(m)
    b.finalize()
    Pr = _densities(mols, 300)
    G_before = b.gmatrix(Pr, use_graph=True)
    e, it, cv, P, tot, gms = b.rhf_scf(accel=3)
    # the SCF drops converged molecules from its replays and restores the
    # full build when it returns
    assert np.array_equal(b.gmatrix(Pr, use_graph=True), G_before)
    b.close()
    note(f"batched SCF of {len(mols)} molecules: {tot:.0f} ms ({gms:.0f} ms in G graphs); iterations {list(it)}")
    assert cv.all() and len(set(int(x) for x in it)) > 1
    for i, m in enumerate(mols):
        ob = oracle.Basis(m)
        ref, _, _, _ = oracle.rhf_scf_ex(ob, ob.two_electron_ints(), m.nelec, maxits=300, eps=1e-10, scf_a=2)
        note(f"  mol {i} (N={m.norb}): batch {e[i]:.10f} oracle {ref.energy:.10f} d {abs(e[i] - ref.energy):.2e}")
        assert ref.converged and abs(e[i] - ref.energy) < 1e-8
    pm.set_options(precision=1, mixed_threshold=1e-6)
    bm = pm.Batch()
    for m in mols:
        bm.add(m)
    bm.finalize(precision=1, threshold=1e-6, dmax=1.0)
    em, itm, cvm, _, _, _ = bm.rhf_scf(accel=3)
    info = bm.info()
    bm.close()
    pm.set_options(precision=0)
    note(f"mixed: FP32 items {info['work_items_fp32']}/{info['work_items']}, max |dE| {np.abs(em - e).max():.2e}")
    assert cvm.all() and np.abs(em - e).max() < 1e-6
