"""Molecular systems in the reference's input model.

A :class:`System` holds what the reference reads from ``patin.dat``
(basis.c:243-321): centres (charge, position in bohr), shells (primitive
count, angular momentum, centre, exponents, raw contraction coefficients) and
the SCF controls.  ``read_patin``/``write_patin`` use the reference's
positional format (basis.c:251-279, shell.c:25-43, center.c:131-136).

The builders below produce the synthetic workloads this package is measured
on: a water molecule with an even-tempered s/p basis, seeded water clusters
with O d shells, seeded H/C/N/O molecules, and a planar C6H6-shaped ring with
a diffuse positron basis.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence

import numpy as np

ANGSTROM = 1.8897261246257702  # bohr per Angstrom
ZNUC = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0, "F": 9.0, "S": 16.0}


@dataclasses.dataclass
class Shell:
    l: int
    cen: int
    al: List[float]
    co: List[float]

    @property
    def npr(self) -> int:
        return len(self.al)

    @property
    def nst(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2


@dataclasses.dataclass
class System:
    charges: np.ndarray          # (ncen,)
    coords: np.ndarray           # (ncen, 3) bohr
    shells: List[Shell]
    nelec: int
    mult: int = 1
    maxits: int = 200
    scf_eps: float = 1e-10
    scf_a: int = 2
    name: str = ""

    @property
    def ncen(self) -> int:
        return len(self.charges)

    @property
    def nshell(self) -> int:
        return len(self.shells)

    @property
    def norb(self) -> int:
        return sum(s.nst for s in self.shells)

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.nshell, dtype=np.int32)
        acc = 0
        for i, s in enumerate(self.shells):
            off[i] = acc
            acc += s.nst
        return off

    def flat(self):
        """Arrays in the order the C entry points take them."""
        npr = np.array([s.npr for s in self.shells], dtype=np.int32)
        lsh = np.array([s.l for s in self.shells], dtype=np.int32)
        cen = np.array([s.cen for s in self.shells], dtype=np.int32)
        al = np.array([a for s in self.shells for a in s.al], dtype=np.float64)
        co = np.array([c for s in self.shells for c in s.co], dtype=np.float64)
        q = np.ascontiguousarray(self.charges, dtype=np.float64)
        xyz = np.ascontiguousarray(self.coords, dtype=np.float64).reshape(-1)
        return q, xyz, npr, lsh, cen, al, co

    def nuclear_repulsion(self) -> float:
        e = 0.0
        for i in range(self.ncen):
            for j in range(i + 1, self.ncen):
                e += self.charges[i] * self.charges[j] / np.linalg.norm(self.coords[i] - self.coords[j])
        return float(e)

    def subset_shells(self, keep: Sequence[int]) -> "System":
        return dataclasses.replace(self, shells=[self.shells[i] for i in keep])


# ---------------------------------------------------------------------------
# patin.dat reader / writer (basis.c:251-279)
# ---------------------------------------------------------------------------
def read_patin(path) -> System:
    toks = open(path).read().split()
    it = iter(toks)
    nshell, norb, ncen = int(next(it)), int(next(it)), int(next(it))
    maxits, eps = int(next(it)), float(next(it))
    nelec, mult = int(next(it)), int(next(it))
    for _ in range(4):
        next(it)  # dpm dft dftx dftx
    _ffc, scf_a, _scf_g = int(next(it)), int(next(it)), int(next(it))
    for _ in range(3):
        next(it)
    q = np.zeros(ncen)
    xyz = np.zeros((ncen, 3))
    for c in range(ncen):
        q[c] = float(next(it))
        xyz[c] = [float(next(it)) for _ in range(3)]
    shells = []
    for _ in range(nshell):
        npr, l, cen = int(next(it)), int(next(it)), int(next(it))
        al, co = [], []
        for _ in range(npr):
            al.append(float(next(it)))
            co.append(float(next(it)))
        shells.append(Shell(l, cen, al, co))
    s = System(q, xyz, shells, nelec, mult, maxits, eps, scf_a, name=str(path))
    assert s.norb == norb, (s.norb, norb)
    return s


def write_patin(sys_: System, path, dpm: int = 0) -> None:
    with open(path, "w") as f:
        f.write(f"{sys_.nshell} {sys_.norb} {sys_.ncen}\n")
        f.write(f"{sys_.maxits} {sys_.scf_eps:.10e}\n")
        f.write(f"{sys_.nelec} {sys_.mult}\n")
        f.write(f"{dpm} 0 0 0\n0 {sys_.scf_a} 0\n0 0 0\n")
        for c in range(sys_.ncen):
            x, y, z = sys_.coords[c]
            f.write(f"{sys_.charges[c]:.17e} {x:.17e} {y:.17e} {z:.17e}\n")
        for s in sys_.shells:
            f.write(f"{s.npr} {s.l} {s.cen}\n")
            for a, c in zip(s.al, s.co):
                f.write(f"{a:.17e} {c:.17e}\n")


# ---------------------------------------------------------------------------
# even-tempered basis recipes
# ---------------------------------------------------------------------------
def _even(a0: float, ratio: float, n: int) -> List[float]:
    return [a0 * ratio ** k for k in range(n)]


def read_posgrid(path) -> np.ndarray:
    """Positron positions (bohr) from a posgrid.dat as the reference's DPM
    loop reads it (rhfscf.c:230-235): the point count, then x y z per
    point, whitespace separated.  Returns an (npts, 3) array."""
    with open(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    v = np.array([float(x) for x in tok[1:1 + 3 * n]], dtype=np.float64)
    if v.size != 3 * n:
        raise ValueError(f"{path}: {n} points announced, {v.size // 3} found")
    return v.reshape(n, 3)


def write_posgrid(points, path) -> None:
    """The inverse of read_posgrid."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(f"{len(pts)}\n")
        for x, y, z in pts:
            f.write(f"{x:.10f} {y:.10f} {z:.10f}\n")


def write_dpmvpol(path, points, results, ground_energy: float) -> None:
    """One line per point in the reference's dpmvpol.dat layout
    (rhfscf.c:257-258: "%15.6le %15.6le %15.6le %20.10le %15.6le %20.10le":
    x y z vpol alfa vstat).  results: the per-point dicts of
    patmol_amd.dpm_calculation.  The reference's vstat reads a variable it
    never sets (rhfscf.c:254); written here is E_first - E_ground, the
    energy of the ground-state density in the field of the positron less the
    ground-state energy."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as f:
        for (x, y, z), r in zip(pts, results):
            f.write("%15.6e %15.6e %15.6e %20.10e %15.6e %20.10e\n"
                    % (x, y, z, r["vpol"], r["alfa"], r["e_first"] - ground_energy))


def heavy_atom_shells(cen: int, zscale: float = 1.0, d: bool = True) -> List[Shell]:
    """O-type recipe: s 0.1*3^k (k<8), p 0.15*3^k (k<6), d 0.3*3^k (k<3),
    all exponents multiplied by zscale ((Z/8)^2 for C and N)."""
    sh = [Shell(0, cen, [a * zscale], [1.0]) for a in _even(0.1, 3.0, 8)]
    sh += [Shell(1, cen, [a * zscale], [1.0]) for a in _even(0.15, 3.0, 6)]
    if d:
        sh += [Shell(2, cen, [a * zscale], [1.0]) for a in _even(0.3, 3.0, 3)]
    return sh


def hydrogen_shells(cen: int) -> List[Shell]:
    return [Shell(0, cen, [a], [1.0]) for a in _even(0.05, 3.0, 6)]


WATER_GEOM = np.array([[0.0, 0.0, 0.0], [0.0, 0.757, 0.587], [0.0, -0.757, 0.587]])  # Angstrom


def water(d_shells: bool = False) -> System:
    """H2O, O at origin, H at (0, +/-0.757, 0.587) A; 10 electrons."""
    xyz = WATER_GEOM * ANGSTROM
    shells = heavy_atom_shells(0, d=d_shells) + hydrogen_shells(1) + hydrogen_shells(2)
    return System(np.array([8.0, 1.0, 1.0]), xyz, shells, nelec=10, name="h2o")


def _random_rotations(n: int, seed: int) -> np.ndarray:
    from scipy.spatial.transform import Rotation

    return Rotation.random(n, random_state=seed).as_matrix()


def water_cluster(nx: int, ny: int, nz: int, seed: int, spacing: float = 2.9,
                  d_shells: bool = True) -> System:
    """Waters on an nx*ny*nz cubic lattice (spacing in A), each rigidly
    rotated about its O by a seeded uniform random rotation."""
    nw = nx * ny * nz
    rots = _random_rotations(nw, seed)
    q, xyz, shells = [], [], []
    w = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                origin = np.array([i, j, k], dtype=float) * spacing
                geo = WATER_GEOM @ rots[w].T + origin
                base = len(q)
                q += [8.0, 1.0, 1.0]
                xyz += list(geo * ANGSTROM)
                shells += heavy_atom_shells(base, d=d_shells)
                shells += hydrogen_shells(base + 1) + hydrogen_shells(base + 2)
                w += 1
    return System(np.array(q), np.array(xyz), shells, nelec=10 * nw,
                  name=f"h2o_{nx}x{ny}x{nz}")


def water8() -> System:
    """(H2O)8 on a 2x2x2 lattice, seed 1001: 232 shells, 448 functions, 80 e."""
    return water_cluster(2, 2, 2, seed=1001)


def water16() -> System:
    """(H2O)16 on a 4x2x2 lattice, seed 1004: 464 shells, 896 functions."""
    return water_cluster(4, 2, 2, seed=1004)


def random_molecule(rng: np.random.Generator, name: str = "") -> System:
    """2-20 atoms from {H,C,N,O} uniform in a cube of volume 1.2 A^3 per atom,
    pairwise distance >= 1.0 A; electron count made even by removing the last
    H (or adding one H by the same sampler)."""
    nat = int(rng.integers(2, 21))
    elems = list(rng.choice(["H", "C", "N", "O"], size=nat))
    side = (1.2 * nat) ** (1.0 / 3.0)

    def place(existing):
        for _ in range(100000):
            p = rng.uniform(0.0, side, size=3)
            if all(np.linalg.norm(p - e) >= 1.0 for e in existing):
                return p
        raise RuntimeError("placement failed")

    pos = []
    for _ in elems:
        pos.append(place(pos))
    ne = int(sum(ZNUC[e] for e in elems))
    if ne % 2:
        if "H" in elems:
            idx = len(elems) - 1 - elems[::-1].index("H")
            del elems[idx]
            del pos[idx]
        else:
            elems.append("H")
            pos.append(place(pos))
    q, xyz, shells = [], [], []
    for c, (e, p) in enumerate(zip(elems, pos)):
        q.append(ZNUC[e])
        xyz.append(p * ANGSTROM)
        if e == "H":
            shells += hydrogen_shells(c)
        else:
            shells += heavy_atom_shells(c, zscale=(ZNUC[e] / 8.0) ** 2)
    ne = int(sum(q))
    return System(np.array(q), np.array(xyz), shells, nelec=ne, name=name)


def molecule_batch(n: int = 64, seed: int = 1002) -> List[System]:
    rng = np.random.default_rng(seed)
    return [random_molecule(rng, name=f"mol{i}") for i in range(n)]


def benzene_ring() -> System:
    """Planar D6h C6H6-shaped ring, C-C 1.39 A, C-H 1.09 A, 42 electrons,
    config-1 electron basis (O recipe scaled by (6/8)^2 on C)."""
    q, xyz, shells = [], [], []
    for k in range(6):
        ang = math.pi / 3 * k
        q.append(6.0)
        xyz.append(np.array([1.39 * math.cos(ang), 1.39 * math.sin(ang), 0.0]) * ANGSTROM)
    for k in range(6):
        ang = math.pi / 3 * k
        q.append(1.0)
        xyz.append(np.array([2.48 * math.cos(ang), 2.48 * math.sin(ang), 0.0]) * ANGSTROM)
    for c in range(6):
        shells += heavy_atom_shells(c, zscale=(6.0 / 8.0) ** 2)
    for c in range(6, 12):
        shells += hydrogen_shells(c)
    return System(np.array(q), np.array(xyz), shells, nelec=42, name="c6h6")
