"""ASE calculator for PM3 over the GPU engine: the molecular part of the reference's calculator
(python/pm3_rs/ase.py:47-390) with the same constructor selectors, property names and units, plus
a batched form that evaluates many structures in one device launch.

ASE's units throughout: energies in eV, forces in eV/Angstrom, positions in Angstrom, Mulliken
charges in e, the dipole in e*Angstrom (Debye * 0.2081943, python/pm3_rs/ase.py:44), the external
field in V/Angstrom (src/python.rs:192-212).  Internally the engine works in Bohr and eV/Bohr;
the conversions happen here, at the boundary, as in the reference's bindings
(src/python.rs:359-427).

When ASE is importable, PM3 is an ase.calculators.calculator.Calculator and works with ASE's
Atoms, optimisers and MD drivers.  Without ASE it keeps the same protocol (results cached against
the structure, get_property / get_potential_energy / get_forces / get_charges /
get_dipole_moment) over any object with get_atomic_numbers() and get_positions(); Structure is a
minimal such object.

What the reference's calculator also offers and this one refuses: periodic structures (stress,
phonons, bands, Berry phase), the pm3-d3 / -d3h4 / -d3h4x corrections and the Hessian family --
none of them is part of the molecular PM3 path this engine implements.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

from . import (_ERRORS, Batch, InvalidInput, Molecule, Pm3Error, Pm3Options, Pm3Parameters, ScfNotConverged,
               Unsupported, closed_form_gradient)

BOHR_TO_ANGSTROM = 0.529177210903
ANGSTROM_TO_BOHR = 1.0 / BOHR_TO_ANGSTROM
# Debye -> e*Angstrom, ASE's dipole unit (python/pm3_rs/ase.py:44)
DEBYE_TO_E_ANGSTROM = 0.2081943

# IUPAC 2016 standard atomic weights (amu), the masses ASE assigns by default, for the elements
# this engine parameterises or refuses by name (Z 1-20)
ATOMIC_MASSES = {1: 1.008, 2: 4.002602, 3: 6.94, 4: 9.0121831, 5: 10.81, 6: 12.011, 7: 14.007, 8: 15.999,
                 9: 18.998403163, 10: 20.1797, 11: 22.98976928, 12: 24.305, 13: 26.9815385, 14: 28.085,
                 15: 30.973761998, 16: 32.06, 17: 35.45, 18: 39.948, 19: 39.0983, 20: 40.078}

try:
    from ase.calculators.calculator import Calculator, PropertyNotImplementedError, all_changes
    HAVE_ASE = True
except ImportError:
    HAVE_ASE = False
    all_changes = ["positions", "numbers", "cell", "pbc", "initial_charges", "initial_magmoms"]

    class PropertyNotImplementedError(NotImplementedError):
        pass

    class Calculator:
        """The part of ASE's Calculator protocol (ase/calculators/calculator.py) that PM3 uses:
        results cached against a snapshot of the structure, recomputed when it changes."""
        implemented_properties: List[str] = []

        def __init__(self, **kwargs):
            self.results = {}
            self.atoms = None
            self.parameters = dict(kwargs)

        def check_state(self, atoms, tol=1e-15):
            if self.atoms is None:
                return list(all_changes)
            changes = []
            if not np.array_equal(self.atoms.get_atomic_numbers(), np.asarray(atoms.get_atomic_numbers())):
                changes.append("numbers")
            a, b = self.atoms.get_positions(), np.asarray(atoms.get_positions(), dtype=float)
            if a.shape != b.shape or np.abs(a - b).max(initial=0.0) > tol:
                changes.append("positions")
            if not np.array_equal(self.atoms.get_pbc(), _pbc(atoms)):
                changes.append("pbc")
            return changes

        def reset(self):
            self.atoms = None
            self.results = {}

        def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
            if atoms is not None:
                self.atoms = Structure(atoms.get_atomic_numbers(), atoms.get_positions(), pbc=_pbc(atoms))

        def get_property(self, name, atoms=None, allow_calculation=True):
            if name not in self.implemented_properties:
                raise PropertyNotImplementedError(f"{name} property not implemented")
            if atoms is None:
                atoms = self.atoms
                system_changes = []
            else:
                system_changes = self.check_state(atoms)
                if system_changes:
                    self.results = {}
            if name not in self.results:
                if not allow_calculation:
                    return None
                if atoms is None:
                    raise RuntimeError("no structure to compute on: pass one, or evaluate a property first")
                self.calculate(atoms, [name], system_changes)
            value = self.results[name]
            return value.copy() if isinstance(value, np.ndarray) else value

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("free_energy" if force_consistent else "energy", atoms)

        def get_forces(self, atoms=None):
            return self.get_property("forces", atoms)

        def get_charges(self, atoms=None):
            return self.get_property("charges", atoms)

        def get_dipole_moment(self, atoms=None):
            return self.get_property("dipole", atoms)


def _pbc(atoms):
    get = getattr(atoms, "get_pbc", None)
    return np.zeros(3, dtype=bool) if get is None else np.asarray(get(), dtype=bool).reshape(3)


class Structure:
    """A minimal Atoms-like structure (Angstrom) for use without ASE: atomic numbers, positions,
    masses (IUPAC weights unless given) and momenta, with the accessors PM3 and the MD driver
    read."""

    def __init__(self, numbers, positions, masses=None, momenta=None, pbc=False, calc=None):
        self.numbers = np.asarray(numbers, dtype=np.int32).copy()
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3).copy()
        if len(self.numbers) != len(self.positions):
            raise InvalidInput("numbers and positions differ in length")
        self.masses = (np.array([ATOMIC_MASSES[int(z)] for z in self.numbers]) if masses is None
                       else np.asarray(masses, dtype=np.float64).copy())
        self.momenta = np.zeros_like(self.positions) if momenta is None else np.asarray(momenta, float).reshape(-1, 3).copy()
        self.pbc = np.zeros(3, dtype=bool) | np.asarray(pbc, dtype=bool)
        self.calc = calc

    def __len__(self):
        return len(self.numbers)

    def get_atomic_numbers(self):
        return self.numbers.copy()

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, positions):
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3).copy()

    def get_masses(self):
        return self.masses.copy()

    def get_momenta(self):
        return self.momenta.copy()

    def set_momenta(self, momenta):
        self.momenta = np.asarray(momenta, dtype=np.float64).reshape(-1, 3).copy()

    def get_velocities(self):
        return self.momenta / self.masses[:, None]

    def set_velocities(self, velocities):
        self.momenta = np.asarray(velocities, dtype=np.float64).reshape(-1, 3) * self.masses[:, None]

    def get_pbc(self):
        return self.pbc.copy()

    def get_kinetic_energy(self):
        return 0.5 * float(np.sum(self.momenta ** 2 / self.masses[:, None]))

    def copy(self):
        return Structure(self.numbers, self.positions, self.masses, self.momenta, self.pbc)

    def get_potential_energy(self):
        return self.calc.get_potential_energy(self)

    def get_forces(self, md=False):
        return self.calc.get_forces(self)


def field_to_engine(field_v_per_angstrom):
    """A uniform field in V/Angstrom to Pm3Options::field (eV/Bohr/e with the sign of
    E = E0 + mu.F), as src/python.rs:198-212 converts it."""
    if field_v_per_angstrom is None:
        return None
    f = np.asarray(field_v_per_angstrom, dtype=np.float64).reshape(-1)
    if f.size != 3:
        raise InvalidInput(f"field must have three components, got {f.size}")
    return tuple(float(-BOHR_TO_ANGSTROM * v) for v in f)


def _properties(energy_ev, gradient_ev_bohr, hof, charges, dipole_debye):
    return {"energy": float(energy_ev),
            # a molecule is filled by aufbau: no electronic entropy, free energy = energy
            "free_energy": float(energy_ev),
            "electronic_entropy_ts_ev": 0.0,
            "forces": -np.asarray(gradient_ev_bohr, dtype=np.float64) * ANGSTROM_TO_BOHR,
            "charges": np.asarray(charges, dtype=np.float64).copy(),
            "dipole": np.asarray(dipole_debye, dtype=np.float64) * DEBYE_TO_E_ANGSTROM,
            "heat_of_formation_kcal": float(hof)}


class PM3(Calculator):
    """PM3 calculator in ASE units (python/pm3_rs/ase.py:47-130, molecular part).

    charge, multiplicity, reference ("auto" | "rhf" | "uhf"), method ("pm3" only here) and field
    (V/Angstrom) are set at construction and used for every evaluation; changing one of them
    invalidates the cached results (python/pm3_rs/ase.py:145-163).  options takes further
    Pm3Options fields (tolerances, accelerator, guess) for the engine; params the parameter set.
    """

    implemented_properties = ["energy", "free_energy", "electronic_entropy_ts_ev", "forces", "charges", "dipole",
                              "heat_of_formation_kcal"]

    def __init__(self, charge=0, multiplicity=1, reference="auto", method="pm3", field=None,
                 options: Optional[Pm3Options] = None, params: Optional[Pm3Parameters] = None, **kwargs):
        super().__init__(**kwargs)
        self.charge = int(charge)
        self.multiplicity = int(multiplicity)
        self.reference = str(reference)
        self.method = str(method)
        self.field = None if field is None else [float(v) for v in np.asarray(field).reshape(-1)]
        self.options = options or Pm3Options()
        self.params = params or Pm3Parameters.standard()
        self._selectors_at_calculate = None
        self._check_method()

    def _check_method(self):
        if self.method != "pm3":
            raise Unsupported(f"method {self.method!r}: this engine computes plain PM3 only (the D3/H4 "
                              "corrections of the reference are not implemented)")

    def _selectors(self):
        return (self.charge, self.multiplicity, self.reference, self.method,
                None if self.field is None else tuple(self.field))

    def check_state(self, atoms, tol=1e-15):
        changes = list(super().check_state(atoms, tol))
        if self._selectors_at_calculate != self._selectors():
            changes.append("model")
        return changes

    def engine_options(self) -> Pm3Options:
        """The Pm3Options of every evaluation: the options template with this calculator's
        reference and field."""
        return dataclasses.replace(self.options, reference=self.reference, field=field_to_engine(self.field))

    def _molecule(self, atoms):
        if bool(np.any(_pbc(atoms))):
            raise Unsupported("periodic structures are not part of this engine; clear atoms.pbc")
        return Molecule(np.asarray(atoms.get_atomic_numbers(), np.int32),
                        np.asarray(atoms.get_positions(), np.float64) * ANGSTROM_TO_BOHR, float(self.charge),
                        self.multiplicity)

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        super().calculate(atoms, properties, system_changes)
        self._selectors_at_calculate = self._selectors()
        self._check_method()
        if "stress" in properties:
            raise PropertyNotImplementedError("stress is only defined for a periodic system")
        g = closed_form_gradient(self._molecule(self.atoms), self.params, self.engine_options())
        self.results.update(_properties(g.energy_ev, g.gradient, g.scf.heat_of_formation_kcal, g.scf.charges,
                                        g.scf.dipole_debye))

    def get_gradient(self, atoms=None):
        """Cartesian energy gradient in eV/Angstrom (the negative of the forces)."""
        return -self.get_forces(atoms)

    def calculate_many(self, structures) -> List[dict]:
        """Energies, forces, charges, dipoles and heats of formation of many structures with this
        calculator's selectors, in one batched SCF + gradient launch (no reference counterpart:
        the reference evaluates one structure per call).  Returns one property dict per structure;
        raises on the first structure that fails, naming it."""
        self._check_method()
        mols = [self._molecule(a) for a in structures]
        if not mols:
            return []
        b = Batch(mols, self.params, self.engine_options(), want_gradient=True)
        try:
            b.solve()
            status, res = b.results()
            grads = b.gradients()
        finally:
            b.close()
        out = []
        for k, (s, r) in enumerate(zip(status, res)):
            if s != 0 or not r.converged:
                cls = _ERRORS.get(int(s), Pm3Error) if s != 0 else ScfNotConverged
                raise cls(f"structure {k}: {cls.__name__} (status {int(s)}, {r.iterations} SCF iterations)")
            out.append(_properties(r.total_ev, grads[k], r.heat_of_formation_kcal, r.charges, r.dipole_debye))
        return out
