"""C ABI and host-side checks that need no GPU.

libpatmol_hip.so must export every function include/patmol_hip.h declares
(the drop-in names of erifs.h:27-28, twoints.h:14, scfaux.h:22-26 and
dpmints.h:12 among them); libpatmol_host.so provides the basis.h accessors
the library reads and the one-electron integrals.  Without a GPU the device
entry points must report an error (and raise in Python), never fall back to
a CPU path.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "patmol_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(names))


def _exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if len(ln.split()) >= 3 and ln.split()[1] in "TtWw"}


def test_libraries_built():
    for p in (pm.LIB_HIP, pm.LIB_HOST):
        assert os.path.exists(p), f"{p} missing: run __graft_entry__.build()"


def test_every_declared_symbol_is_exported():
    decl = _declared()
    assert {"init_ERIFS", "clean_ERIFS", "two_electron_ints", "rhf_form_gmatrix", "uhf_form_gmatrix",
            "dpm_ints"} <= set(decl)
    exp = _exported(pm.LIB_HIP)
    missing = [n for n in decl if n not in exp]
    assert not missing, missing
    H, G = pm.lib()
    for n in decl:
        getattr(G, n)
    assert sorted(pm.exported_symbols()) == decl


def test_drop_in_signatures_match_reference_headers():
    """The drop-in prototypes keep the reference's parameter lists."""
    text = open(HEADER).read()
    expect = [
        r"ERIFS \*init_ERIFS\(int nst, int nend, const char \*sname\);",
        r"void clean_ERIFS\(ERIFS \*e\);",
        r"void two_electron_ints\(ERIFS \*tfiles, int start\);",
        r"void rhf_form_gmatrix\(const ERIFS \*tfiles, const double \*Pmat, double \*Gmat\);",
        r"void uhf_form_gmatrix\(const ERIFS \*tfiles, const double \*PmatA, const double \*PmatB,\s+"
        r"double \*GmatA, double \*GmatB\);",
        r"void dpm_ints\(double \*hmat\);",
        r"void one_electron_ints\(double \*Smat, double \*Tmat, double \*Hmat\);",
    ]
    for e in expect:
        assert re.search(e, text), e


def test_host_basis_accessors_exported():
    exp = _exported(pm.LIB_HOST)
    for n in ("shells", "centers", "number_of_shells", "number_of_orbitals", "number_of_centers",
              "skip_center", "set_center_position", "nuclear_repulsion", "patmol_host_one_electron_ints",
              "init_Basis", "dpm_augment", "patmol_basis_load", "patmol_rhf_scf", "patmol_uhf_scf"):
        assert n in exp, n


@pytest.mark.parametrize("factory", [lambda: systems.water(d_shells=True), systems.benzene_ring])
def test_host_one_electron_vs_oracle(factory):
    s = factory()
    pm.load_system(s)
    H, _ = pm.lib()
    assert H.number_of_orbitals() == s.norb and H.number_of_shells() == s.nshell
    S, T, Hm = pm.host_one_electron_ints()
    So, To, Ho = oracle.Basis(s).one_electron_ints()
    assert np.abs(S - So).max() < 1e-13
    assert np.abs(T - To).max() < 1e-11 * np.abs(To).max()
    assert np.abs(Hm - Ho).max() < 1e-11 * np.abs(Ho).max()
    assert abs(H.nuclear_repulsion() - s.nuclear_repulsion()) < 1e-10


def test_host_sf6_patin_reader(golden_dir):
    """init_Basis reads patin.dat from the working directory (basis.c:251)."""
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    pm.load_system(s)
    H, _ = pm.lib()
    assert H.number_of_orbitals() == 214 and H.number_of_shells() == 86
    S, T, Hm = pm.host_one_electron_ints()
    So, To, Ho = oracle.Basis(s).one_electron_ints()
    assert np.abs(Hm - Ho).max() < 1e-10 * np.abs(Ho).max()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_device_path_fails_loudly_without_gpu():
    pm.load_system(systems.water())
    pm.set_options(strict=False)
    with pytest.raises(pm.PatmolError):
        pm.EriStore()
