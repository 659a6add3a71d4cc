"""Bitwise reproducibility of the device G build and of the device SCF.

J' and K' are summed in fixed point across workgroups and in a fixed copy
order inside a workgroup, so the G of one density must come out identical
bit for bit on every build, and an SCF on identical inputs must take the
same path: the same energies at every iteration and the same iteration
count.  Run to run identical G is what lets the reference's stop rule
(|dE| < 10 eps or pdiff < eps, rhfscf.c:52-59) and its damping on the sign
of dE (rhfscf.c:60-69) give one iteration count for one input.
"""
import numpy as np
import pytest

import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,factory,tau", [
    ("h2o_d", lambda: systems.water(d_shells=True), 0.0),
    ("dimer", lambda: systems.water_cluster(2, 1, 1, seed=7), 1e-12),
    ("water8", systems.water8, 1e-12),
])
def test_rhf_gmatrix_bitwise_repeatable(name, factory, tau):
    s = factory()
    note(f"repeat G {name} N={s.norb}")
    pm.load_system(s)
    pm.set_options(tau=tau, precision=0, strict=False)
    st = pm.EriStore()
    P = packed(random_density(s.norb, 21))
    G0 = st.rhf_gmatrix(P)
    for k in range(4):
        G = st.rhf_gmatrix(P)
        assert np.array_equal(G, G0), f"build {k + 1} differs by {np.abs(G - G0).max():.3e}"
    st.close()


def test_uhf_gmatrix_bitwise_repeatable():
    s = systems.water_cluster(2, 1, 1, seed=7)
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    st = pm.EriStore()
    PA, PB = packed(random_density(s.norb, 22)), packed(random_density(s.norb, 23))
    A0, B0 = st.uhf_gmatrix(PA, PB)
    for _ in range(3):
        A, B = st.uhf_gmatrix(PA, PB)
        assert np.array_equal(A, A0) and np.array_equal(B, B0)
    st.close()


@pytest.mark.parametrize("name,factory,mode", [
    ("h2o", systems.water, 1),
    ("ring", systems.benzene_ring, 1),
    ("ring_diag", systems.benzene_ring, 0),
])
def test_device_scf_repeatable(name, factory, mode):
    s = factory()
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False, scf_density=mode)
    st = pm.EriStore()
    r0, P0, _ = pm.rhf_scf_device(store=st)
    r1, P1, _ = pm.rhf_scf_device(store=st)
    st.close()
    pm.set_options(scf_density=1)
    note(f"{name}: E {r0['energy']:.12f} it {r0['iterations']} / E {r1['energy']:.12f} it {r1['iterations']}")
    assert r0["converged"] == 1 and r1["converged"] == 1
    assert r0["iterations"] == r1["iterations"]
    assert r0["energy"] == r1["energy"] and r0["energy0"] == r1["energy0"]
    assert np.array_equal(P0, P1)
