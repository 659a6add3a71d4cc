"""GPU parity of the McMurchie-Davidson J/K kernel (md_kernel.h), which runs
the FP64 work items of (dd|dd), (dd|dp), (dp|dp) and (dd|pp), against the
oracle's restatement of the reference's Rys path (twoints.c:23-186,
scfaux.c:428-714) and against the library's own Rys kernels for the same
classes (patmol_hip_set_jk_engine mask 0); the tests put all four classes
on the MD kernel (mask 15, with bit 4 so that the kept-share rule does not
send (dd|dp) back to its Rys kernel), not only the default (dd|dd) and
(dd|dp).

Tolerances as in test_gpu_parity.py: 1e-10 absolute on G elements against
the oracle; the two device engines agree to 1e-12 (different algorithms,
FP64 rounding only)."""
import os

import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note, packed, random_density

pytestmark = pytest.mark.gpu

G_TOL = 1e-10
ENGINE_TOL = 1e-12
ALL = 31   # (dd|dd), (dd|dp), (dp|dp), (dd|pp) on the MD kernel, whatever their kept share (bit 4)


def _store(sys_, tau=0.0, start=0):
    pm.load_system(sys_)
    pm.set_options(tau=tau, precision=0, strict=False, md_engine=ALL)
    return pm.EriStore(start=start)


def _both_engines(st, fn):
    """fn with every MD-capable class on the McMurchie-Davidson kernel, then
    with all of them on their Rys kernels."""
    pm.set_options(md_engine=ALL)
    try:
        a = fn(st)
        pm.set_options(md_engine=0)
        b = fn(st)
    finally:
        pm.set_options(md_engine=-1)
    return a, b


def test_md_rhf_gmatrix_water2_vs_oracle_and_rys():
    s = systems.water_cluster(2, 1, 1, seed=1001)
    note(f"MD rhf G (H2O)2 N={s.norb}")
    st = _store(s)
    P = packed(random_density(s.norb, 21))
    G_md, G_rys = _both_engines(st, lambda x: x.rhf_gmatrix(P))
    ints = oracle.Basis(s).two_electron_ints()
    G_ref = ints.rhf_gmatrix(P)
    d_ref = float(np.abs(G_md - G_ref).max())
    d_eng = float(np.abs(G_md - G_rys).max())
    note(f"  max |G_md - G_oracle| = {d_ref:.3e}, max |G_md - G_rys| = {d_eng:.3e} (|G|max {np.abs(G_ref).max():.2f})")
    assert d_ref < G_TOL
    assert d_eng < ENGINE_TOL
    # reproducible bit for bit
    pm.set_options(md_engine=ALL)
    try:
        assert np.array_equal(G_md, st.rhf_gmatrix(P))
    finally:
        pm.set_options(md_engine=-1)
    st.close()


def test_md_uhf_gmatrix_water2_vs_oracle():
    s = systems.water_cluster(2, 1, 1, seed=1001)
    st = _store(s)
    PA = packed(random_density(s.norb, 22))
    PB = packed(random_density(s.norb, 23))
    (GA, GB), (GAr, GBr) = _both_engines(st, lambda x: x.uhf_gmatrix(PA, PB))
    ints = oracle.Basis(s).two_electron_ints()
    GA_ref, GB_ref = ints.uhf_gmatrix(PA, PB)
    d = max(float(np.abs(GA - GA_ref).max()), float(np.abs(GB - GB_ref).max()))
    e = max(float(np.abs(GA - GAr).max()), float(np.abs(GB - GBr).max()))
    note(f"MD uhf (H2O)2: max |dG| vs oracle {d:.3e}, vs Rys engine {e:.3e}")
    assert d < G_TOL and e < ENGINE_TOL
    st.close()


def test_md_start_shell_vs_oracle():
    """two_electron_ints(start > 0): only quartets whose largest original
    shell index reaches start (the DPM xints filter, twoints.c:316-337)."""
    s = systems.water_cluster(2, 1, 1, seed=1001)
    start = 20   # inside the first water's d shells (original order)
    st = _store(s, start=start)
    P = packed(random_density(s.norb, 24))
    G_md, G_rys = _both_engines(st, lambda x: x.rhf_gmatrix(P))
    G_ref = oracle.Basis(s).two_electron_ints(start=start).rhf_gmatrix(P)
    d = float(np.abs(G_md - G_ref).max())
    e = float(np.abs(G_md - G_rys).max())
    note(f"MD start={start}: max |dG| vs oracle {d:.3e}, vs Rys engine {e:.3e}")
    assert d < G_TOL and e < ENGINE_TOL
    st.close()


def test_md_sf6_contracted_vs_rys(golden_dir):
    """SF6: contracted p shells (the MD kernel's primitive loops) against the
    Rys kernels, which the parity tests pin to the oracle."""
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    st = _store(s, tau=1e-14)
    P = packed(random_density(s.norb, 25))
    G_md, G_rys = _both_engines(st, lambda x: x.rhf_gmatrix(P))
    e = float(np.abs(G_md - G_rys).max())
    note(f"MD SF6 (N={s.norb}): max |G_md - G_rys| = {e:.3e} (|G|max {np.abs(G_rys).max():.2f})")
    assert e < ENGINE_TOL * max(1.0, float(np.abs(G_rys).max()))
    st.close()


def test_md_water8_vs_rys_screened():
    """(H2O)8 at the benchmark's Schwarz threshold: every MD class with many
    work items, against the Rys engine."""
    s = systems.water8()
    st = _store(s, tau=1e-12)
    P = packed(random_density(s.norb, 26))
    G_md, G_rys = _both_engines(st, lambda x: x.rhf_gmatrix(P))
    e = float(np.abs(G_md - G_rys).max())
    note(f"MD (H2O)8: max |G_md - G_rys| = {e:.3e}")
    assert e < ENGINE_TOL
    st.close()


def test_engine_mask_kept_share_rule():
    """(dd|dp) leaves the MD kernel by default where most of its work items'
    shell-quartet slots are kept (the water dimer at 2.9 A: 0.86 of them,
    counted on the oracle's work-list mirror), and stays on it when bit 4 is
    set, or when the screen empties most slots (the dimer at 8 A: 0.32);
    (dd|dd) stays in every case."""
    masks = {}
    try:
        for spacing in (2.9, 8.0):
            s = systems.water_cluster(2, 1, 1, seed=1001, spacing=spacing)
            pm.load_system(s)
            pm.set_options(tau=1e-12, precision=0, strict=False, md_engine=-1)
            st = pm.EriStore()
            masks[spacing] = st.engine_mask()
            if spacing == 2.9:
                pm.set_options(md_engine=3 | 16)
                masks["forced"] = st.engine_mask()
                pm.set_options(md_engine=-1)
            st.close()
    finally:
        pm.set_options(md_engine=-1)
    note(f"engine masks: {masks}")
    assert masks[2.9] == 1 and masks["forced"] == 3 and masks[8.0] == 3

