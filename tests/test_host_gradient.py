"""Host parts of the RHF gradient, pinned on the CPU by finite differences.

patmol_host_onee_gradient gives d/dR of E1 = 2 sum P H - 2 sum W S for fixed
packed P and W, with S and H the one-electron integrals of the host library
(patmol_host_one_electron_ints, itself pinned against the oracle and the
reference's one_electron_ints); central differences of E1 over moved centres
check every centre and axis, the attraction's Hellmann-Feynman term
included.  The nuclear repulsion gradient is checked the same way.
"""
import numpy as np
import pytest

import patmol_amd as pm
from patmol_amd import systems

H_STEP = 1e-4


def _e1(P, W):
    S, T, H = pm.host_one_electron_ints()
    n = int(round((np.sqrt(8 * len(P) + 1) - 1) / 2))
    w = np.full(len(P), 2.0)
    w[np.arange(n) * (np.arange(n) + 3) // 2] = 1.0
    return float(np.sum(w * (2.0 * P * H - 2.0 * W * S)))


@pytest.mark.parametrize("name,factory", [("h2o", systems.water), ("h2o_d", lambda: systems.water(d_shells=True))])
def test_onee_gradient_vs_finite_differences(name, factory):
    s = factory()
    pm.load_system(s)
    Hl, _ = pm.lib()
    n = s.norb
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n))
    P = (A @ A.T / n)[np.tril_indices(n)].copy()
    B = rng.standard_normal((n, n))
    W = ((B + B.T) / n)[np.tril_indices(n)].copy()
    g = pm.host_onee_gradient(P, W)
    fd = np.zeros_like(g)
    for c in range(s.ncen):
        for x in range(3):
            r = s.coords[c].copy()
            for sgn in (1, -1):
                rr = r.copy()
                rr[x] += sgn * H_STEP
                Hl.set_center_position(c, rr)
                fd[c, x] += sgn * _e1(P, W) / (2 * H_STEP)
            Hl.set_center_position(c, r)
    err = np.abs(g - fd).max()
    assert err < 1e-6 * max(1.0, np.abs(fd).max()), (err, g, fd)
    # translational invariance: the forces of a rigid shift cancel
    assert np.abs(g.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(g).max())


def test_nuclear_repulsion_gradient_vs_finite_differences():
    s = systems.water()
    pm.load_system(s)
    Hl, _ = pm.lib()
    g = pm.nuclear_repulsion_gradient()
    for c in range(s.ncen):
        for x in range(3):
            r = s.coords[c].copy()
            e = []
            for sgn in (1, -1):
                rr = r.copy()
                rr[x] += sgn * H_STEP
                Hl.set_center_position(c, rr)
                e.append(Hl.nuclear_repulsion())
            Hl.set_center_position(c, r)
            assert abs(g[c, x] - (e[0] - e[1]) / (2 * H_STEP)) < 1e-7
