"""Extract the SF6 known answer (the reference's only golden run) into
sf6_golden.json.  Reads the reference's committed output files
scfout.gs.dat and short.gs.dat; sf6_patin.dat is a verbatim copy of the
reference's patin.dat input (data, not source)."""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

lines = open(os.path.join(REF, "scfout.gs.dat")).read().splitlines()


def field(prefix):
    for ln in lines:
        if ln.startswith(prefix):
            return ln.split("=")[1].split()[0]
    raise KeyError(prefix)


orb = []
for ln in lines:
    t = ln.split()
    if len(t) == 3 and t[0].isdigit() and t[2] in ("0", "2"):
        orb.append(float(t[1]))
short = [float(x) for x in open(os.path.join(REF, "short.gs.dat")).read().split()]
out = {
    "source": "scfout.gs.dat / short.gs.dat of the reference (produced by its MPI binary)",
    "iterations": int(field("Final Iteration")),
    "energy": float(field("Hartree Fock Energy")),
    "electronic_energy": float(field("Electronic Energy")),
    "nuclear_repulsion": float(field("Nuclear Rep. Energy")),
    "kinetic_energy": float(field("Kinetic Energy")),
    "ediff": float(field("Difference in Final Energies")),
    "pdiff": float(field("Pmatrix diff norm")),
    "orbital_energies": orb,
    "short": short,
}
json.dump(out, open(os.path.join(HERE, "sf6_golden.json"), "w"), indent=1)
print(out["energy"], out["iterations"], len(orb))
