"""Analytic gradient of the DPM-augmented RHF energy (one grid point of the
DPM loop, rhfscf.c:217-263: the basis after dpm_augment(), the positron at
skip_center(), H carrying -V of dpm_ints, dpmints.c:120-199) on the GPU.

* The DPM potential's derivative kernel (patmol_hip_dpm_gradient) against
  central differences of -2 sum_ab P_ab V_ab (V from dpm_ints) at a fixed
  seeded density, for every coordinate of the three nuclei and the positron
  centre (h = 1e-4 bohr).
* The whole gradient (patmol_rhf_dpm_gradient: the RHF terms over the
  augmented basis with H - V, plus the DPM term) against central
  differences of the converged device SCF energy of the same point, every
  coordinate, within 1e-6 Eh/bohr; translational invariance (all centres
  moved together leave E unchanged: the rows sum to zero).
"""
import dataclasses

import numpy as np
import pytest

import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu

Q = [0.5, -0.4, 2.5]
STEP = 1e-4


def _setup():
    s = dataclasses.replace(systems.water(d_shells=True), scf_a=3)
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False, scf_incremental=0)
    pm.add_dpm_center()
    H, G = pm.lib()
    H.dpm_augment()
    nc = H.number_of_centers()
    xyz = [np.array(r, dtype=np.float64) for r in s.coords] + [np.array(Q, dtype=np.float64)]
    assert len(xyz) == nc
    H.set_center_position(nc - 1, xyz[-1])
    return H, G, nc, xyz


def _displaced(H, xyz, c, x, d, fn):
    r = xyz[c].copy()
    r[x] += d
    H.set_center_position(c, r)
    try:
        return fn()
    finally:
        H.set_center_position(c, xyz[c])


def test_dpm_term_vs_differences():
    H, G, nc, xyz = _setup()
    n = H.number_of_orbitals()
    P = packed(random_density(n, 31, 0.3))
    w = np.full(len(P), 2.0)
    w[np.arange(n) * (np.arange(n) + 3) // 2] = 1.0

    def e_dpm():
        V = np.zeros(len(P))
        G.dpm_ints(V)           # V = -V_dpm (packed)
        return 2.0 * float(np.sum(w * P * V))

    g = pm.dpm_gradient(P)
    fd = np.zeros((nc, 3))
    for c in range(nc):
        for x in range(3):
            fd[c, x] = (_displaced(H, xyz, c, x, STEP, e_dpm) - _displaced(H, xyz, c, x, -STEP, e_dpm)) / (2 * STEP)
    d = float(np.abs(g - fd).max())
    note(f"DPM term: max |analytic - difference| {d:.2e} (max |g| {np.abs(g).max():.3e}); rows sum "
         f"{np.abs(g.sum(axis=0)).max():.1e}")
    assert d < 1e-7
    assert np.abs(g.sum(axis=0)).max() < 1e-11


def test_dpm_scf_gradient_vs_differences():
    H, G, nc, xyz = _setup()

    def energy():
        st = pm.EriStore()
        try:
            r, P, _ = pm.rhf_scf_device(store=st, dpm=True)
        finally:
            st.close()
        assert r["converged"]
        return r["energy"]

    st = pm.EriStore()
    r, P, _ = pm.rhf_scf_device(store=st, dpm=True)
    assert r["converged"]
    g = pm.rhf_dpm_gradient(P, store=st)
    st.close()
    fd = np.zeros((nc, 3))
    for c in range(nc):
        for x in range(3):
            fd[c, x] = (_displaced(H, xyz, c, x, STEP, energy) - _displaced(H, xyz, c, x, -STEP, energy)) / (2 * STEP)
    d = float(np.abs(g - fd).max())
    note(f"DPM point E {r['energy']:.10f}: max |analytic - difference| {d:.2e} Eh/bohr; dE/dQ {g[-1]}; "
         f"rows sum {np.abs(g.sum(axis=0)).max():.1e}")
    assert d < 1e-6
    assert np.abs(g.sum(axis=0)).max() < 1e-8
