"""Closed-form pins of the oracle's one-electron integrals (oneints.c:12-234)
and of the host library's CPU version, independent of the reference.

For normalised primitives a (exponent al, centre A) and b (be, B), with
p = al + be, mu = al be / p, R = |A - B|, K = exp(-mu R^2), P = (al A + be B)/p:
  <s_a|s_b>          = N_a N_b (pi/p)^(3/2) K
  <s_a|-lap/2|s_b>   = mu (3 - 2 mu R^2) <s_a|s_b>
  <s_a|-Z/|r-C||s_b> = -Z N_a N_b (2 pi/p) K F_0(p |P - C|^2)
  <p_z,a|s_b>        = N_pz N_b (pi/p)^(3/2) K (P_z - A_z)
with N_s = (2 al/pi)^(3/4) and N_pz = 2 sqrt(al) (2 al/pi)^(3/4).  A dropped
term, a wrong sign of P - A or a transposed pair each fail one of these.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

import oracle
import patmol_amd as pm
from patmol_amd import systems


def _boys0(t):
    return 1.0 if t < 1e-14 else 0.5 * math.sqrt(math.pi / t) * erf(math.sqrt(t))


A = np.array([0.0, 0.0, 0.0])
B = np.array([0.3, -0.2, 1.1])
C3 = np.array([-0.7, 0.5, 0.4])
ZS = [1.0, 2.0, 3.0]
AL, BE = 0.8, 1.7


def _sys():
    """s on A, s on B, p on A; nuclei at A, B and C3."""
    shells = [systems.Shell(0, 0, [AL], [1.0]), systems.Shell(0, 1, [BE], [1.0]), systems.Shell(1, 0, [AL], [1.0])]
    return systems.System(np.array(ZS), np.vstack([A, B, C3]), shells, nelec=2)


def _closed_forms():
    p, mu = AL + BE, AL * BE / (AL + BE)
    R2 = float(((A - B) ** 2).sum())
    K = math.exp(-mu * R2)
    P = (AL * A + BE * B) / p
    Na, Nb = (2 * AL / math.pi) ** 0.75, (2 * BE / math.pi) ** 0.75
    S = Na * Nb * (math.pi / p) ** 1.5 * K
    T = mu * (3 - 2 * mu * R2) * S
    V = sum(-z * Na * Nb * 2 * math.pi / p * K * _boys0(p * float(((P - c) ** 2).sum()))
            for z, c in zip(ZS, [A, B, C3]))
    Npz = 2 * math.sqrt(AL) * Na
    Spz = Npz * Nb * (math.pi / p) ** 1.5 * K * (P[2] - A[2])
    return S, T, V, Spz


def _idx(r, c):
    r, c = max(r, c), min(r, c)
    return r * (r + 1) // 2 + c


def _check(S, T, H):
    s_ab, t_ab, v_ab, s_pz_b = _closed_forms()
    # orbitals: 0 = s_A, 1 = s_B, 2..4 = p_A (x, y, z)
    assert abs(S[_idx(0, 0)] - 1.0) < 1e-14
    assert abs(S[_idx(1, 0)] - s_ab) < 1e-14
    assert abs(T[_idx(1, 0)] - t_ab) < 1e-13
    assert abs(H[_idx(1, 0)] - (t_ab + v_ab)) < 1e-12
    assert abs(S[_idx(4, 1)] - s_pz_b) < 1e-14
    # single s Gaussian: T = 3 al / 2
    assert abs(T[_idx(0, 0)] - 1.5 * AL) < 1e-13
    # p_A with s_A on the same centre: zero by parity
    assert abs(S[_idx(4, 0)]) < 1e-15


def test_oracle_one_electron_closed_forms():
    _check(*oracle.Basis(_sys()).one_electron_ints())


def test_host_one_electron_closed_forms():
    pm.load_system(_sys())
    _check(*pm.host_one_electron_ints())


@pytest.mark.gpu
def test_device_one_electron_closed_forms():
    pm.load_system(_sys())
    _check(*pm.one_electron_ints())


# ---------------------------------------------------------------------------
# multipole integrals (momints.c:14-92): closed forms for the same primitives
#   <s_a|x_k|s_b>      = P_k <s_a|s_b>
#   <s_a|x_k^2|s_b>    = (P_k^2 + 1/(2p)) <s_a|s_b>
#   <s_a|x_k x_l|s_b>  = P_k P_l <s_a|s_b>          (k != l)
#   <p_z,a|z|s_b>      = N_pz N_b (pi/p)^(3/2) K ((P_z - A_z) P_z + 1/(2p))
# ---------------------------------------------------------------------------
def _moment_closed_forms():
    p, mu = AL + BE, AL * BE / (AL + BE)
    R2 = float(((A - B) ** 2).sum())
    K = math.exp(-mu * R2)
    P = (AL * A + BE * B) / p
    Na, Nb = (2 * AL / math.pi) ** 0.75, (2 * BE / math.pi) ** 0.75
    S = Na * Nb * (math.pi / p) ** 1.5 * K
    dip = [P[k] * S for k in range(3)]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    quad = [(P[k] * P[l] + (0.5 / p if k == l else 0.0)) * S for k, l in pairs]
    Npz = 2 * math.sqrt(AL) * Na
    pz_z_s = Npz * Nb * (math.pi / p) ** 1.5 * K * ((P[2] - A[2]) * P[2] + 0.5 / p)
    return dip, quad, pz_z_s


def _check_moments(dip, quad):
    d, q, pzs = _moment_closed_forms()
    ij = _idx(1, 0)
    for k in range(3):
        assert abs(dip[k][ij] - d[k]) < 1e-14
    for k in range(6):
        assert abs(quad[k][ij] - q[k]) < 1e-14
    assert abs(dip[2][_idx(4, 1)] - pzs) < 1e-14
    # s_A at the origin: <x> = 0, <x^2> = 1/(4 al)
    assert abs(dip[0][_idx(0, 0)]) < 1e-15
    assert abs(quad[0][_idx(0, 0)] - 0.25 / AL) < 1e-14


def test_oracle_moment_closed_forms():
    _check_moments(*oracle.Basis(_sys()).moment_ints())


@pytest.mark.gpu
def test_device_moment_closed_forms():
    pm.load_system(_sys())
    _check_moments(*pm.moment_ints())


def test_oracle_dipole_is_energy_derivative():
    """Hellmann-Feynman: with H(F) = H - F d_z (rhf_getDipoleHam,
    rhfscf.c:474-489) the converged energy has dE/dF = -(electronic z moment
    of rhf_analyzeMoments); and the reference's finite-field polarisability
    -2 (E - E_first)/F^2 agrees with the central second difference to O(F)."""
    s = systems.water()
    b = oracle.Basis(s)
    ints = b.two_electron_ints()
    _, _, H = b.one_electron_ints()
    dip, quad = b.moment_ints()
    gs, P, _, _ = oracle.rhf_scf_ex(b, ints, s.nelec, eps=1e-12)
    elec, nuc, tot = oracle.rhf_moments(P, dip, quad, s.charges, s.coords)
    F = 2e-3
    ep, _, _, _ = oracle.rhf_scf_ex(b, ints, s.nelec, H=H - F * dip[2], guess=P, eps=1e-12)
    em, _, _, _ = oracle.rhf_scf_ex(b, ints, s.nelec, H=H + F * dip[2], guess=P, eps=1e-12)
    dEdF = (ep.energy - em.energy) / (2 * F)
    assert abs(dEdF + elec[2]) < 1e-5 * max(1.0, abs(elec[2]))
    a_cd = -(ep.energy + em.energy - 2 * gs.energy) / F ** 2
    alfa, a0, a2 = oracle.rhf_finite_field(b, ints, s.nelec, F, eps=1e-12)
    assert a_cd > 0 and abs(alfa[2] - a_cd) < 2e-2 * a_cd
