import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pm3-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
SYM = {"H": 1, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Na": 11, "Mg": 12, "Al": 13, "Si": 14,
       "P": 15, "S": 16, "Cl": 17}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def pins():
    with open(os.path.join(ROOT, "tests", "golden", "reference_pins.json")) as fh:
        return json.load(fh)


def atoms_to_zx(atoms):
    import numpy as np
    z = np.array([SYM[a[0]] for a in atoms], dtype=np.int32)
    x = np.array([a[1:4] for a in atoms], dtype=float) * ANGSTROM_TO_BOHR
    return z, x


@pytest.fixture
def exact_eigensolves(monkeypatch):
    """Full-precision eigensolves in every SCF iteration (the engine's default loosens early
    iterations), for tests that follow the reference's iteration trajectory step for step."""
    monkeypatch.setenv("PM3_EIG_TOL_FACTOR", "0")


@pytest.fixture(scope="session")
def oracle():
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    import oracle as O
    O.lib()
    return O
