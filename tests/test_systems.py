"""Synthetic workloads and the patin.dat reader/writer (CPU)."""
import os

import numpy as np

from patmol_amd import systems


def test_water_sizes():
    s = systems.water()
    assert (s.nshell, s.norb, s.nelec) == (26, 38, 10)
    sd = systems.water(d_shells=True)
    assert (sd.nshell, sd.norb) == (29, 56)


def test_water_clusters_sizes_and_seeds():
    s = systems.water8()
    assert (s.nshell, s.norb, s.nelec, s.ncen) == (232, 448, 80, 24)
    assert np.array_equal(s.coords, systems.water8().coords)   # seeded
    s16 = systems.water16()
    assert (s16.nshell, s16.norb, s16.nelec) == (464, 896, 160)
    # rigid rotations keep every O-H bond at its 0.9578 A length
    for w in range(8):
        o, h1, h2 = s.coords[3 * w: 3 * w + 3]
        for h in (h1, h2):
            assert abs(np.linalg.norm(h - o) / systems.ANGSTROM - np.hypot(0.757, 0.587)) < 1e-12


def test_molecule_batch_rules():
    mols = systems.molecule_batch(64, seed=1002)
    assert len(mols) == 64
    for m in mols:
        assert 1 <= m.ncen <= 21
        assert m.nelec % 2 == 0 and m.nelec == int(sum(m.charges))
        d = np.linalg.norm(m.coords[:, None] - m.coords[None], axis=-1) / systems.ANGSTROM
        assert np.all(d[np.triu_indices(m.ncen, 1)] >= 1.0 - 1e-12)
    again = systems.molecule_batch(64, seed=1002)
    assert all(np.array_equal(a.coords, b.coords) for a, b in zip(mols, again))


def test_benzene_ring():
    s = systems.benzene_ring()
    assert s.nelec == 42 and s.ncen == 12
    c = s.coords[:6] / systems.ANGSTROM
    assert abs(np.linalg.norm(c[0] - c[1]) - 1.39) < 1e-12


def test_patin_round_trip(tmp_path, golden_dir):
    s = systems.read_patin(os.path.join(golden_dir, "sf6_patin.dat"))
    assert (s.nshell, s.norb, s.ncen, s.nelec) == (86, 214, 7, 70)
    p = os.path.join(tmp_path, "patin.dat")
    systems.write_patin(s, p)
    t = systems.read_patin(p)
    assert np.array_equal(s.coords, t.coords) and np.array_equal(s.charges, t.charges)
    assert [(a.l, a.cen, a.al, a.co) for a in s.shells] == [(a.l, a.cen, a.al, a.co) for a in t.shells]
    assert (t.maxits, t.scf_eps, t.scf_a) == (s.maxits, s.scf_eps, s.scf_a)
