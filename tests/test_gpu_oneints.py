"""GPU parity of the device one_electron_ints (oneints.c:137-234) against
the oracle's restatement and the host library's CPU version.

Tolerances (FP64): S to 1e-13 absolute (overlaps are <= 1); T and H to
1e-12 relative to their largest element (the Rys nodes and weights of V carry
~1e-14 relative error).
"""
import os

import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note

pytestmark = pytest.mark.gpu

S_TOL = 1e-13
REL_TOL = 1e-12


def _compare(tag, got, ref):
    S, T, H = got
    So, To, Ho = ref
    dS = float(np.abs(S - So).max())
    dT = float(np.abs(T - To).max()) / float(np.abs(To).max())
    dH = float(np.abs(H - Ho).max()) / float(np.abs(Ho).max())
    note(f"{tag}: max|dS| {dS:.2e}  rel dT {dT:.2e}  rel dH {dH:.2e}")
    assert dS < S_TOL and dT < REL_TOL and dH < REL_TOL


@pytest.mark.parametrize("name,factory", [
    ("h2o", systems.water),
    ("h2o_d", lambda: systems.water(d_shells=True)),
    ("benzene_ring", systems.benzene_ring),
    ("water_cluster_2x2x2", lambda: systems.water_cluster(2, 2, 2, seed=1001)),
])
def test_oneints_vs_oracle(name, factory):
    s = factory()
    note(f"one_electron_ints {name} nshell={s.nshell} norb={s.norb}")
    pm.load_system(s)
    _compare(name, pm.one_electron_ints(), oracle.Basis(s).one_electron_ints())


def test_oneints_sf6_contracted(golden_dir):
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    note("one_electron_ints SF6 (contracted shells)")
    pm.load_system(s)
    got = pm.one_electron_ints()
    _compare("sf6 vs oracle", got, oracle.Basis(s).one_electron_ints())
    _compare("sf6 vs host CPU", got, pm.host_one_electron_ints())


def test_oneints_skip_center_dpm_basis():
    """With the positron centre added and visible, V leaves that centre out
    (oneints.c:51-53): the electron block equals the plain basis's H, and the
    whole matrix equals the CPU version over the augmented basis."""
    s = systems.water(d_shells=True)
    pm.load_system(s)
    H, _ = pm.lib()
    H.patmol_basis_add_dpm_center()
    H.dpm_augment()
    H.set_center_position(s.ncen, np.array([0.4, -0.3, 1.7]))
    assert H.skip_center() == s.ncen
    got = pm.one_electron_ints()
    _compare("dpm-augmented vs host CPU", got, pm.host_one_electron_ints())
    n0 = s.norb * (s.norb + 1) // 2
    ref = oracle.Basis(s).one_electron_ints()
    _compare("electron block vs plain oracle", [g[:n0] for g in got], ref)


def test_oneints_repeat_identical():
    """One writer per element: two calls give bit-identical results."""
    pm.load_system(systems.water(d_shells=True))
    a = pm.one_electron_ints()
    b = pm.one_electron_ints()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
