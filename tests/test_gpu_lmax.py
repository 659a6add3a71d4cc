"""Shells above d are refused with an error, not computed wrongly.

The reference runs to l = 5 (pconfig.h MAXL, rootN of rys.c:245-323 for six
or more roots); the device kernels are instantiated for s, p and d only, and
every device entry point that takes the basis refuses an f shell with
PATMOL_HIP_EUNSUPPORTED before it launches anything.  A later call on an
s/p/d basis in the same process works as before.
"""
import dataclasses

import numpy as np
import pytest

import patmol_amd as pm
from patmol_amd import systems
from patmol_amd.systems import Shell

pytestmark = pytest.mark.gpu


def _water_with_f():
    s = systems.water()
    return dataclasses.replace(s, shells=list(s.shells) + [Shell(3, 0, [0.8], [1.0])])


def test_f_shell_refused_then_spd_works():
    s = _water_with_f()
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    with pytest.raises(pm.PatmolError, match="l > 2"):
        pm.one_electron_ints()
    with pytest.raises(pm.PatmolError, match="l=3"):
        pm.EriStore()
    # the next basis without f shells builds as usual
    w = systems.water()
    pm.load_system(w)
    S, T, H = pm.one_electron_ints()
    st = pm.EriStore()
    n = w.norb
    P = np.zeros(n * (n + 1) // 2)
    P[np.arange(n) * (np.arange(n) + 3) // 2] = 0.1
    G = st.rhf_gmatrix(P)
    st.close()
    assert np.isfinite(G).all() and np.abs(G).max() > 0
    assert np.allclose(S[np.arange(n) * (np.arange(n) + 3) // 2], 1.0)
