"""ERINT record interop (erifs.h:19-22) on the GPU: the writer against the
oracle's restatement of two_electron_ints (twoints.c:188-418) record by
record, and the record-driven G builds (scfaux.c:428-714) against the
oracle's scatter and the direct device build.

Tolerances: integrals 1e-11 absolute (the parity bar of the ERI path), G
elements 1e-10 absolute.  The record set is compared exactly: the write
threshold 1e-14 sees values that differ from the oracle's by ~1e-25 there.
"""
import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,factory,start", [
    ("h2o_d", lambda: systems.water(d_shells=True), 0),
    ("h2o_d_start", lambda: systems.water(d_shells=True), 7),
    ("water_dimer", lambda: systems.water_cluster(2, 1, 1, seed=7), 0),
])
def test_write_erints_vs_oracle(name, factory, start, tmp_path):
    s = factory()
    pm.load_system(s)
    pm.set_options(tau=0.0, precision=0)
    st = pm.EriStore(start=0)
    path = str(tmp_path / "MINTS.DAT")
    n = st.write_erints(path, start=start)
    got = np.fromfile(path, dtype=pm.ERINT_DTYPE)
    ref = oracle.Basis(s).two_electron_ints(start).array()
    note(f"write_erints {name}: {n} records (oracle {len(ref)})")
    assert n == len(got) == len(ref)
    for f in "ijkl":
        assert np.array_equal(got[f], ref[f]), f
    assert np.abs(got["val"] - ref["val"]).max() < 1e-11
    st.close()


def test_rhf_gmatrix_from_written_file(tmp_path):
    s = systems.water(d_shells=True)
    pm.load_system(s)
    pm.set_options(tau=0.0, precision=0)
    st = pm.EriStore()
    path = str(tmp_path / "MINTS.DAT")
    st.write_erints(path)
    P = packed(random_density(s.norb, 12))
    G_file = pm.rhf_gmatrix_erints(None, P, path=path)
    G_dir = st.rhf_gmatrix(P)
    G_ref = oracle.Basis(s).two_electron_ints().rhf_gmatrix(P)
    d1 = float(np.abs(G_file - G_ref).max())
    d2 = float(np.abs(G_file - G_dir).max())
    note(f"G from records: vs oracle scatter {d1:.2e}, vs direct device build {d2:.2e}")
    assert d1 < 1e-10 and d2 < 1e-10
    # accumulate semantics, as rhf_form_gmatrix
    G2 = pm.rhf_gmatrix_erints(None, P, G=G_file.copy(), path=path)
    assert np.allclose(G2, 2 * G_file, rtol=0, atol=1e-12)
    st.close()


def test_gmatrix_from_oracle_records():
    """Records produced by the oracle (the reference's format and order) fed
    to the device scatter, RHF and UHF, against the oracle's scatter."""
    s = systems.water_cluster(2, 1, 1, seed=7)
    pm.load_system(s)
    ints = oracle.Basis(s).two_electron_ints()
    rec = ints.array().view(pm.ERINT_DTYPE)
    P = packed(random_density(s.norb, 21))
    d = float(np.abs(pm.rhf_gmatrix_erints(rec, P) - ints.rhf_gmatrix(P)).max())
    PA, PB = packed(random_density(s.norb, 22)), packed(random_density(s.norb, 23))
    GA, GB = pm.uhf_gmatrix_erints(rec, PA, PB)
    GAr, GBr = ints.uhf_gmatrix(PA, PB)
    du = max(float(np.abs(GA - GAr).max()), float(np.abs(GB - GBr).max()))
    note(f"{len(rec)} oracle records: RHF |dG| {d:.2e}, UHF |dG| {du:.2e}")
    assert d < 1e-10 and du < 1e-10


def test_records_out_of_range_rejected():
    s = systems.water()
    pm.load_system(s)
    rec = np.zeros(3, dtype=pm.ERINT_DTYPE)
    rec["val"] = 1.0
    rec["i"] = [0, 1, s.norb]          # the last index lies outside the basis
    P = np.zeros(s.norb * (s.norb + 1) // 2)
    pm.set_options(strict=False)
    with pytest.raises(pm.PatmolError):
        pm.rhf_gmatrix_erints(rec, P)
