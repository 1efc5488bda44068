"""Writes tests/golden/reference_pins.json from the reference repository's own frozen outputs.

Reads (as data, nothing is imported or run):
  tools/oracle/INTERNAL_CONSISTENCY.json   pm3-rs total energies (eV) of the element hydrides
  tools/oracle/ALL_ELEMENTS_RESULTS.json   pm3-rs heats of formation (kcal/mol) of the same systems
  tools/oracle/all_element_validation.py   read as text for the geometry recipe only (restated below)
  tests/data/mopac_oracle.tsv              MOPAC v23.2.5 rows (HoF, dipole, IP, charges, geometry)

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import csv
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_pins.json")
SUPPORTED = {"H": 1, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Na": 11, "Mg": 12, "Al": 13,
             "Si": 14, "P": 15, "S": 16, "Cl": 17}
# Covalent radii the reference's validation script uses to pick the hydride bond length
# (tools/oracle/all_element_validation.py:29-41), for the supported elements.
COV = {1: 0.31, 3: 1.28, 4: 0.96, 5: 0.84, 6: 0.76, 7: 0.71, 8: 0.66, 9: 0.57, 11: 1.66, 12: 1.41, 13: 1.21,
       14: 1.11, 15: 1.07, 16: 1.05, 17: 1.02}


def hydride(sym, z, tore):
    """Closed-shell hydride geometry (Angstrom, 10 decimals as the reference writes its XYZ):
    tools/oracle/all_element_validation.py:70-110."""
    d = max(0.90, 1.05 * (COV[z] + COV[1]))
    if tore in (1, 7, 8):
        atoms, charge = [(sym, 0, 0, 0), ("H", d, 0, 0)], (1 if tore == 8 else 0)
    elif tore in (2, 6):
        atoms, charge = [(sym, 0, 0, 0), ("H", d, 0, 0), ("H", -0.25 * d, 0.9682458366 * d, 0)], 0
    elif tore == 3:
        atoms = [(sym, 0, 0, 0)] + [("H", x * d, y * d, 0) for x, y in ((1, 0), (-0.5, 0.8660254038), (-0.5, -0.8660254038))]
        charge = 0
    elif tore == 5:
        atoms = [(sym, 0, 0, 0)] + [("H", x * 0.9539392014 * d, y * 0.9539392014 * d, 0.3 * d)
                                    for x, y in ((1, 0), (-0.5, 0.8660254038), (-0.5, -0.8660254038))]
        charge = 0
    else:
        q = d / 3 ** 0.5
        atoms = [(sym, 0, 0, 0), ("H", q, q, q), ("H", -q, -q, q), ("H", -q, q, -q), ("H", q, -q, -q)]
        charge = 0
    return [[a[0]] + [float("%.10f" % c) for c in a[1:]] for a in atoms], charge


def main():
    ic = {r["name"]: r for r in json.load(open(os.path.join(REF, "tools/oracle/INTERNAL_CONSISTENCY.json")))["results"]}
    ae = {r["name"]: r for r in json.load(open(os.path.join(REF, "tools/oracle/ALL_ELEMENTS_RESULTS.json")))["results"]}
    rows = [r for r in csv.DictReader(l for l in open(os.path.join(REF, "src/data/element_data.csv")) if not l.startswith("#"))]
    tore = {int(r["z"]): int(float(r["tore"])) for r in rows}
    hyd = []
    for sym, z in SUPPORTED.items():
        name = f"{sym}H{ {1:1,3:1,4:2,5:3,6:4,7:3,8:2,9:1,11:1,12:2,13:3,14:4,15:3,16:2,17:1}[z]}_z{z}"
        if name not in ic:
            continue
        atoms, charge = hydride(sym, z, tore[z])
        hyd.append({"name": name, "atoms": atoms, "charge": charge, "energy_ev": ic[name]["energy_ev"],
                    "hof_kcal": ae[name]["energy_pm3_rs_kcal"] if name in ae else None})
    mopac, open_shell = [], []
    for line in open(os.path.join(REF, "tests/data/mopac_oracle.tsv")):
        if line.startswith("#") or line.startswith("name\t") or not line.strip():
            continue
        f = line.rstrip("\n").split("\t")
        atoms = [a.split() for a in f[9].split(";")]
        if any(a[0] not in SUPPORTED for a in atoms):
            continue
        row = {"name": f[0], "formula": f[1], "charge": float(f[2]), "multiplicity": int(f[3]),
               "hof_kcal": float(f[4]), "dipole_debye": float(f[5]) if f[5] else None,
               "charges": [float(c) for c in f[7].split(";")],
               "atoms": [[a[0]] + [float(v) for v in a[1:4]] for a in atoms]}
        if int(f[3]) != 1:
            open_shell.append(row)
            continue
        mopac.append({"name": f[0], "formula": f[1], "charge": float(f[2]), "hof_kcal": float(f[4]),
                      "dipole_debye": float(f[5]) if f[5] else None, "ip_ev": float(f[6]) if f[6] else None,
                      "charges": [float(c) for c in f[7].split(";")],
                      "atoms": [[a[0]] + [float(v) for v in a[1:4]] for a in atoms]})
    json.dump({"source": "reference tools/oracle/*.json and tests/data/mopac_oracle.tsv",
               "hydrides": hyd, "mopac_closed_shell": mopac, "mopac_open_shell": open_shell}, open(OUT, "w"), indent=1)
    print(f"wrote {OUT}: {len(hyd)} hydrides, {len(mopac)} MOPAC rows, {len(open_shell)} open-shell rows")


if __name__ == "__main__":
    main()
