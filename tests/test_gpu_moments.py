"""GPU parity of the multipole path (momints.c, scfaux.c:140-340,
rhfscf.c:176-215): device moment integrals and the MOMINTS.DAT writer
against the oracle, moment analysis and finite-field polarisability through
the device SCF against the oracle's restatement.

Tolerances: moment integrals 1e-12 relative to their largest element;
moments of a converged density 1e-7 (the densities themselves agree to
~1e-8); polarisabilities 1e-4 relative (they are second differences of
energies converged to ~1e-11 Eh at F = 1e-2).
"""
import os

import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,factory", [
    ("h2o_d", lambda: systems.water(d_shells=True)),
    ("benzene_ring", systems.benzene_ring),
])
def test_moment_ints_vs_oracle(name, factory):
    s = factory()
    pm.load_system(s)
    d, q = pm.moment_ints()
    do, qo = oracle.Basis(s).moment_ints()
    ed = float(np.abs(d - do).max() / np.abs(do).max())
    eq = float(np.abs(q - qo).max() / np.abs(qo).max())
    note(f"moment ints {name}: rel d(dip) {ed:.2e} rel d(quad) {eq:.2e}")
    assert ed < 1e-12 and eq < 1e-12


def test_momints_file_vs_oracle(golden_dir, tmp_path):
    """The MOMINTS.DAT records (momints.c:113-180) of contracted SF6: order
    (ish, jsh <= ish, components with jr <= ir) and values."""
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    pm.load_system(s)
    path = str(tmp_path / "MOMINTS.DAT")
    pm.write_momints(path)
    rec = np.fromfile(path, dtype=pm.MOMINTS_DTYPE)
    order = []
    for i, si in enumerate(s.shells):
        for j in range(i + 1):
            sj = s.shells[j]
            oi, oj = s.offsets()[i], s.offsets()[j]
            for a in range(si.nst):
                for b_ in range(sj.nst):
                    if oj + b_ > oi + a:
                        break
                    order.append((oi + a, oj + b_))
    order = np.array(order)
    assert np.array_equal(rec["ir"], order[:, 0]) and np.array_equal(rec["jr"], order[:, 1])
    do, qo = oracle.Basis(s).moment_ints()
    ij = rec["ir"] * (rec["ir"] + 1) // 2 + rec["jr"]
    assert np.abs(rec["dip"] - do[:, ij].T).max() < 1e-12 * np.abs(do).max()
    assert np.abs(rec["quad"] - qo[:, ij].T).max() < 1e-12 * np.abs(qo).max()


def test_rhf_moments_vs_oracle():
    s = systems.water(d_shells=True)
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0)
    res, P, ev = pm.rhf_scf_device()
    e, n, t = pm.rhf_moments(P)
    b = oracle.Basis(s)
    do, qo = b.moment_ints()
    eo, no, to = oracle.rhf_moments(P, do, qo, s.charges, s.coords)
    note(f"water dipole (total) {t[:3]}, quadrupole diag {t[[3, 6, 8]]}")
    assert np.abs(n - no).max() < 1e-12
    assert np.abs(e - eo).max() < 1e-10 * np.abs(eo).max()
    # the electron count: trace of 2 P S is N, the same weights as the moments
    S, _, _ = pm.one_electron_ints()
    f = np.full(len(P), 4.0)
    f[[i * (i + 1) // 2 + i for i in range(s.norb)]] = 2.0
    assert abs((f * P * S).sum() - s.nelec) < 1e-8


def test_uhf_moments_closed_shell_equals_rhf():
    s = systems.water()
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0)
    r, P, _ = pm.rhf_scf_device()
    e1, _, t1 = pm.rhf_moments(P)
    e2, _, t2 = pm.uhf_moments(P, P)
    assert np.abs(e1 - e2).max() < 1e-12 * np.abs(e1).max()


def test_finite_field_vs_oracle():
    s = systems.water()
    pm.load_system(s)
    pm.set_options(tau=0.0, precision=0)
    F = 1e-2
    alfa, gs = pm.rhf_finite_field(F)
    b = oracle.Basis(s)
    ints = b.two_electron_ints()
    ao, a0o, a2o = oracle.rhf_finite_field(b, ints, s.nelec, F, maxits=s.maxits, eps=s.scf_eps, scf_a=s.scf_a)
    note(f"finite field F={F}: alpha {alfa[:3]} (oracle {ao}); alpha0 {alfa[3]:.6f} alpha2 {alfa[4]:.6f}")
    assert np.abs(alfa[:3] - ao).max() < 1e-4 * np.abs(ao).max()
    assert abs(alfa[3] - a0o) < 1e-4 * abs(a0o)
    assert abs(alfa[4] - a2o) < 1e-4 * abs(a0o)
