"""Allocation-free device G build against the allocating one, on the GPU.

patmol_hip_rhf_gmatrix_device_noalloc keeps its work buffers in a fixed
device array instead of allocated memory; it must give the same G as
patmol_hip_rhf_gmatrix_device (FP64 fixed-point J/K sums: bitwise equal),
through the C ABI, on two calls in a row whose second basis is smaller, and
reject a basis above its size limit before any device work.
"""
import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import DeviceArray, note, packed, random_density

pytestmark = pytest.mark.gpu


def _both(sys_, seed):
    pm.load_system(sys_)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    st = pm.EriStore()
    rng = np.random.default_rng(seed)
    P = packed(random_density(sys_.norb, seed))
    G0 = rng.standard_normal(len(P))          # G accumulates: start from a nonzero G
    dP, dG_ref, dG_na = DeviceArray(P), DeviceArray(G0), DeviceArray(G0)
    st.rhf_gmatrix_device(dP.addr, dG_ref.addr, 0)
    st.rhf_gmatrix_device_noalloc(dP.addr, dG_na.addr, 0)
    g_ref, g_na = dG_ref.get(), dG_na.get()
    for d in (dP, dG_ref, dG_na):
        d.free()
    st.close()
    return P, G0, g_ref, g_na


def test_noalloc_matches_allocating_entry_twice():
    big, small = systems.water(d_shells=True), systems.water()
    assert big.norb > small.norb
    for k, s in enumerate((big, small)):
        note(f"noalloc call {k + 1}: {s.name} N={s.norb}")
        P, G0, g_ref, g_na = _both(s, 31 + k)
        d = float(np.abs(g_na - g_ref).max())
        note(f"  max |G_noalloc - G_device| = {d:.3e}")
        assert np.array_equal(g_na, g_ref)
        # and both are the oracle's G (scfaux.c:428-552) on top of G0
        ref = oracle.Basis(s).two_electron_ints().rhf_gmatrix(P)
        assert float(np.abs((g_na - G0) - ref).max()) < 1e-10


def test_noalloc_rejects_large_basis():
    s = systems.benzene_ring()
    assert s.norb > pm.NOALLOC_MAX_NORB
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    st = pm.EriStore()
    no2 = s.norb * (s.norb + 1) // 2
    dP, dG = DeviceArray(np.zeros(no2)), DeviceArray(np.zeros(no2))
    with pytest.raises(pm.PatmolError, match="allocation-free"):
        st.rhf_gmatrix_device_noalloc(dP.addr, dG.addr, 0)
    assert not np.any(dG.get())
    dP.free()
    dG.free()
    st.close()
