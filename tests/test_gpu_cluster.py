"""GPU parity on multi-tile systems (several tiles per angular momentum)."""
import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tau", [0.0, 1e-12])
def test_rhf_gmatrix_two_waters(tau):
    s = systems.water_cluster(2, 1, 1, seed=7)
    pm.load_system(s)
    pm.set_options(tau=tau, precision=0)
    st = pm.EriStore()
    P = packed(random_density(s.norb, 9))
    note(f"two waters N={s.norb} tau={tau}: GPU G")
    G = st.rhf_gmatrix(P)
    b = oracle.Basis(s)
    note("  oracle ERI list")
    ints = b.two_electron_ints()
    Gr = ints.rhf_gmatrix(P, nthreads=8)
    d = np.abs(G - Gr)
    i = int(np.argmax(d))
    note(f"  max |dG| = {d.max():.3e} at packed {i}; |G|max {np.abs(Gr).max():.2f}; stats {st.stats()}")
    assert d.max() < (1e-10 if tau == 0 else 1e-8)
