"""Incremental G builds in the device RHF (patmol_hip_set_scf_incremental)
and the per-item precision split, on the GPU.

Between full builds the SCF builds only G(P - Pref) and adds it to G(Pref)
(G is linear in P); quartets whose Schwarz bound times the largest density
difference of their six density blocks is below the Schwarz threshold are
skipped, and in mixed precision the same per-item product below 1e-6 selects
FP32.  Checked:

* water (38 functions), against the oracle's SCF (rhfscf.c:21-174): FP64
  incremental within 1e-8 Eh and the same iteration count as the full-build
  device SCF; mixed incremental within 1e-6 Eh;
* (H2O)8 (448 functions, DIIS): FP64 incremental within 1e-8 Eh of the full
  builds, mixed incremental within 1e-6 Eh, with work actually skipped and
  actually run in FP32;
* the incremental run is bitwise repeatable.
"""
import dataclasses

import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


def _run(store, precision, incr):
    pm.set_options(tau=1e-12, precision=precision, mixed_threshold=1e-6, strict=False, scf_incremental=incr)
    try:
        r, _, _ = pm.rhf_scf_device(store=store)
    finally:
        pm.set_options(precision=0, scf_incremental=0)
    return r


def _describe(tag, r):
    return (f"{tag}: E {r['energy']:.12f} it {r['iterations']} full {r['full_builds']} delta {r['delta_builds']} "
            f"items {r['items']:.0f} fp32 {r['items_fp32']:.0f} skipped {r['items_skipped']:.0f}")


def test_water_incremental_vs_oracle():
    s = systems.water()
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    st = pm.EriStore()
    full = _run(st, 0, 0)
    inc = _run(st, 0, 4)
    mix = _run(st, 1, 4)
    st.close()
    ob = oracle.Basis(s)
    ref, _, _ = oracle.rhf_scf(ob, ob.two_electron_ints(), s.nelec, maxits=s.maxits, eps=s.scf_eps, scf_a=s.scf_a)
    for tag, r in (("full FP64", full), ("incremental FP64", inc), ("incremental mixed", mix)):
        note(_describe(tag, r) + f"  |dE| vs oracle {abs(r['energy'] - ref.energy):.2e}")
    assert full["converged"] and inc["converged"] and mix["converged"]
    assert inc["delta_builds"] > 0 and inc["full_builds"] < full["full_builds"]
    assert abs(inc["energy"] - ref.energy) < 1e-8
    assert inc["iterations"] == full["iterations"]
    assert abs(mix["energy"] - ref.energy) <= 1e-6


def test_water8_incremental_fp64_and_mixed():
    s = dataclasses.replace(systems.water8(), scf_a=3)
    pm.load_system(s)
    pm.set_options(tau=1e-12, precision=0, strict=False)
    st = pm.EriStore()
    full = _run(st, 0, 0)
    inc = _run(st, 0, 8)
    inc2 = _run(st, 0, 8)
    mix = _run(st, 1, 8)
    st.close()
    for tag, r in (("full FP64", full), ("incremental FP64", inc), ("incremental mixed", mix)):
        note(_describe(tag, r) + f"  |dE| vs full {abs(r['energy'] - full['energy']):.2e}; "
             f"G {r['gbuild_ms'] / (r['iterations'] + 1):.2f} ms/iteration")
    assert full["converged"] and inc["converged"] and mix["converged"]
    assert abs(inc["energy"] - full["energy"]) < 1e-8
    assert abs(inc["iterations"] - full["iterations"]) <= 1
    assert inc["items_skipped"] > 0 and inc["items_fp32"] == 0
    # the incremental path is as reproducible as the full one
    assert inc2["energy"] == inc["energy"] and inc2["iterations"] == inc["iterations"]
    assert mix["items_fp32"] > 0
    assert abs(mix["energy"] - full["energy"]) <= 1e-6
