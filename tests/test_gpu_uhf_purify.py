"""The device UHF (patmol_hip_uhf_scf, uhfscf.c:22-168) on an open-shell
system large enough for its densities to come from trace-correcting
purification (128 functions and up): triplet (H2O)3, 168 functions, 30
electrons.  Purification and the eigensolver path must give the same
converged state (energy 1e-8 Eh; densities and orbital energies 1e-6: the
stop rule is on the energy change, |dE| < 1e-10, so the two paths stop at
different iterations and first-order quantities agree to about the square
root of that), with no purification falling back to the eigensolver, and
the energy must
match the oracle's UHF (tests/golden/uhf_triplet_water3.json, written by
make_uhf_triplet.py next to it) to 1e-8 Eh."""
import json
import os

import numpy as np
import pytest

import patmol_amd as pm
from tests.gpu_util import note
from tests.golden.make_uhf_triplet import triplet_water3

pytestmark = pytest.mark.gpu


def test_uhf_purification_matches_eigensolver_and_oracle(golden_dir):
    gold = json.load(open(os.path.join(golden_dir, "uhf_triplet_water3.json")))
    s = triplet_water3()
    assert s.norb >= 128
    pm.load_system(s)
    pm.set_options(tau=0.0, precision=0)
    out = {}
    try:
        for mode in (1, 0):
            pm.set_options(scf_density=mode)
            res, PA, PB, eA, eB = pm.uhf_scf_device()
            note(f"UHF triplet (H2O)3, scf_density {mode}: E={res['energy']:.12f} it={res['iterations']} "
                 f"converged={res['converged']} purification fallbacks {res['purify_fallbacks']}")
            out[mode] = (res, PA, PB, eA, eB)
    finally:
        pm.set_options(scf_density=1)
    rp, PAp, PBp, eAp, eBp = out[1]
    re, PAe, PBe, eAe, eBe = out[0]
    assert rp["converged"] == 1 and re["converged"] == 1
    assert rp["purify_fallbacks"] == 0
    assert abs(rp["energy"] - re["energy"]) < 1e-8
    assert np.abs(PAp - PAe).max() < 1e-6 and np.abs(PBp - PBe).max() < 1e-6
    assert np.abs(eAp - eAe).max() < 1e-6 and np.abs(eBp - eBe).max() < 1e-6
    note(f"oracle UHF: E={gold['energy']:.12f} in {gold['iterations']} iterations")
    assert abs(rp["energy"] - gold["energy"]) < 1e-8
