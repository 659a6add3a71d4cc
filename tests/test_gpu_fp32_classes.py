"""Error of the FP32 integrals of the mixed-precision path, per angular-
momentum class, against the oracle's FP64 integrals (twoints.c:23-186).

The mixed build sends a work item to FP32 when its Schwarz bound times the
largest density element it contracts with is below 1e-6, so what matters is
each FP32 integral's error relative to its Schwarz bound Q_IJ Q_KL: an FP32
item's contribution to G then errs by at most that ratio times 1e-6.  For
every class of the J/K kernels (21 with la >= lb, lc >= ld, pair class of
the bra >= the ket's), on (H2O)8 (s/p/d, 232 shells) and seeded shell
quartets of that class: the FP32 integrals through the C ABI
(patmol_hip_eri_blocks_prec, the arithmetic of jk_kernel3<..., float>)
against the oracle, as max |error| / (Q_IJ Q_KL) and max |error|; the FP64
integrals of the same quartets within 1e-11.  The table is printed
(profiles/fp32_class_errors_water8.md keeps one run of it).
"""
import numpy as np
import pytest

import oracle
import patmol_amd as pm
from patmol_amd import systems
from tests.gpu_util import note

pytestmark = pytest.mark.gpu

L = "spd"
PAIRS = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
PER_CLASS = 60
REL_BAR = 5e-5   # |error| / (Q_IJ Q_KL) of an FP32 integral: 5e-11 per contribution below 1e-6


def _classes():
    out = []
    for b in range(6):
        for k in range(b + 1):
            out.append(PAIRS[b] + PAIRS[k])
    return out


def test_fp32_integral_error_per_class():
    s = systems.water8()
    pm.load_system(s)
    pm.set_options(tau=0.0, precision=0, strict=False)
    st = pm.EriStore()
    ob = oracle.Basis(s)
    q = st.pair_schwarz()
    ls = np.array([sh.l for sh in s.shells])
    by_l = {l: np.flatnonzero(ls == l) for l in range(3)}
    rng = np.random.default_rng(2024)

    def qpair(i, j):
        i, j = max(i, j), min(i, j)
        return q[i * (i + 1) // 2 + j]

    rows = []
    worst_rel = 0.0
    for la, lb, lc, ld in _classes():
        quartets = np.stack([rng.choice(by_l[la], PER_CLASS), rng.choice(by_l[lb], PER_CLASS),
                             rng.choice(by_l[lc], PER_CLASS), rng.choice(by_l[ld], PER_CLASS)], axis=1)
        g32 = st.eri_blocks(quartets, ob.nst, fp32=True)
        g64 = st.eri_blocks(quartets, ob.nst)
        rel = absmax = err64 = 0.0
        for (I, J, K, Lq), a32, a64 in zip(quartets, g32, g64):
            ref = ob.eri_block(int(I), int(J), int(K), int(Lq))
            bound = qpair(int(I), int(J)) * qpair(int(K), int(Lq))
            e = float(np.abs(a32 - ref).max())
            absmax = max(absmax, e)
            if bound > 0:
                rel = max(rel, e / bound)
            err64 = max(err64, float(np.abs(a64 - ref).max()))
        name = f"({L[la]}{L[lb]}|{L[lc]}{L[ld]})"
        rows.append((name, rel, absmax, err64))
        worst_rel = max(worst_rel, rel)
        assert err64 < 1e-11, (name, err64)
    st.close()
    note("| class | FP32 max abs error / (Q_IJ Q_KL) | FP32 max abs error | FP64 max abs error |")
    note("|---|---|---|---|")
    for name, rel, absmax, err64 in rows:
        note(f"| {name} | {rel:.2e} | {absmax:.2e} | {err64:.2e} |")
    note(f"{PER_CLASS} seeded quartets per class on (H2O)8; worst FP32 error / bound {worst_rel:.2e} "
         f"(an FP32 item's G contributions then err by at most {worst_rel * 1e-6:.1e} each)")
    assert worst_rel < REL_BAR
