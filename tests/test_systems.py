"""Synthetic workloads and the patin.dat reader/writer (CPU)."""
import os

import numpy as np
import pytest

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


def test_posgrid_and_dpmvpol_files(tmp_path):
    """posgrid.dat as rhfscf.c:230-235 reads it (count, then x y z per point,
    any whitespace) and dpmvpol.dat in the layout of rhfscf.c:257-258."""
    p = tmp_path / "posgrid.dat"
    p.write_text("3\n0 0 0\n0.0 0.0\n2.0   0.5 -0.4 2.5\n")
    pts = systems.read_posgrid(p)
    assert pts.shape == (3, 3) and np.array_equal(pts[2], [0.5, -0.4, 2.5]) and pts[1, 2] == 2.0
    q = tmp_path / "again.dat"
    systems.write_posgrid(pts, q)
    assert np.array_equal(systems.read_posgrid(q), pts)
    bad = tmp_path / "short.dat"
    bad.write_text("2\n0 0 0\n")
    with pytest.raises(ValueError):
        systems.read_posgrid(bad)
    out = tmp_path / "dpmvpol.dat"
    res = [{"vpol": -0.125, "alfa": 0.0, "e_first": -10.5}, {"vpol": -0.0625, "alfa": 2.0, "e_first": -10.25},
           {"vpol": -0.03125, "alfa": 3.5, "e_first": -10.125}]
    systems.write_dpmvpol(out, pts, res, -11.0)
    lines = out.read_text().splitlines()
    assert len(lines) == 3
    assert lines[1] == "%15.6e %15.6e %15.6e %20.10e %15.6e %20.10e" % (0.0, 0.0, 2.0, -0.0625, 2.0, 0.75)
    cols = [float(x) for x in lines[2].split()]
    assert cols == [0.5, -0.4, 2.5, -0.03125, 3.5, 0.875]
