"""torch.autograd.Function over the device RHF energy and its analytic
nuclear gradient.

RhfEnergy.apply(coords, system) runs the device SCF (Pulay DIIS) of `system`
with its centres moved to `coords` (ncen x 3, bohr, float64) and returns the
total energy as a 0-d tensor; backward returns dE/dcoords from
patmol_rhf_gradient (two-electron derivative integrals on the GPU).  The
density of the last forward call is kept as the guess of the next one.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

import patmol_amd as pm


class RhfEnergy(torch.autograd.Function):
    _guess = {}

    @staticmethod
    def forward(ctx, coords, system, tau=1e-12, precision=0):
        xyz = coords.detach().cpu().numpy().astype(np.float64)
        s = dataclasses.replace(system, coords=xyz.copy(), scf_a=3)
        pm.load_system(s)
        pm.set_options(tau=tau, precision=precision, strict=False)
        st = pm.EriStore()
        key = (system.name, system.norb)
        guess = RhfEnergy._guess.get(key)
        res, P, _ = pm.rhf_scf_device(store=st, guess=guess)
        if not res["converged"]:
            st.close()
            raise pm.PatmolError("SCF did not converge")
        RhfEnergy._guess[key] = P
        g = pm.rhf_gradient(P, store=st)
        st.close()
        ctx.save_for_backward(torch.from_numpy(g).to(coords.dtype))
        return torch.tensor(res["energy"], dtype=coords.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None


def rhf_energy(coords, system, tau=1e-12, precision=0):
    return RhfEnergy.apply(coords, system, tau, precision)
