"""Writes dpm_ring.json: the oracle's DPM grid loop (rhf_dpm_calculation,
rhfscf.c:217-263, as restated in oracle/__init__.py) on the planar
C6H6-shaped ring (systems.benzene_ring: 300 electron functions, 42
electrons) with the positron at the ring centre and 2 bohr above and below
it on the six-fold axis.  The oracle stores every integral of the ring
(several GB of records) and takes several minutes, so its answer is kept as
a fixture.

    python tests/golden/make_dpm_ring.py [threads]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from patmol_amd import systems  # noqa: E402

POINTS = [[0.0, 0.0, 0.0], [0.0, 0.0, 2.0], [0.0, 0.0, -2.0]]


def main():
    nthreads = int(sys.argv[1]) if len(sys.argv) > 1 else 1
    s = systems.benzene_ring()
    t0 = time.time()
    gs, pts = oracle.rhf_dpm_calculation(s, POINTS, nthreads=nthreads)
    out = {"system": f"C6H6 ring, {s.nshell} shells, {s.norb} functions, 42 electrons; DPM centre with 7 shells",
           "points": POINTS,
           "ground_state": {"energy": gs.energy, "iterations": gs.iterations, "converged": gs.converged},
           "results": pts, "oracle_seconds": time.time() - t0}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dpm_ring.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
