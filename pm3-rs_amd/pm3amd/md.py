"""Molecular dynamics on the PM3 surface for many independent structures at once.

The reference runs MD through its ASE calculator (python/pm3_rs/ase.py): ASE's VelocityVerlet
asks the calculator for forces once per step, one structure and one SCF at a time.  Here the same
integrator (ase/md/verlet.py: half kick, drift, new forces, half kick) advances a whole list of
structures in lockstep, and each step's forces for all of them come from one batched SCF +
gradient launch on the device-resident batch (new positions uploaded with
pm3_batch_set_positions, each SCF optionally started from the structure's previous density).

Units are ASE's: positions in Angstrom, masses in amu, momenta in amu*Angstrom/t with ASE's time
unit t = Angstrom*sqrt(amu/eV), energies in eV; time steps are given in femtoseconds.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import Batch, Molecule, ScfNotConverged, _ERRORS, Pm3Error
from .ase import ANGSTROM_TO_BOHR, PM3

# ASE's femtosecond in its own time unit, 1e-15 s * 1e10 Angstrom/m * sqrt(e / amu), and its
# Boltzmann constant in eV/K, both from the CODATA 2014 constants ASE uses by default
FS = 1.0e-5 * math.sqrt(1.6021766208e-19 / 1.660539040e-27)
KB_EV = 1.38064852e-23 / 1.6021766208e-19


def maxwell_boltzmann(structures, temperature_k, rng: np.random.Generator, stationary=True):
    """Momenta drawn from the Maxwell-Boltzmann distribution at temperature_k (ASE's
    MaxwellBoltzmannDistribution: p = N(0, 1) * sqrt(m kT) per component), with the centre-of-mass
    momentum removed when stationary (ASE's Stationary)."""
    for a in structures:
        m = np.asarray(a.get_masses(), dtype=np.float64)
        p = rng.standard_normal((len(m), 3)) * np.sqrt(m * KB_EV * temperature_k)[:, None]
        if stationary:
            p -= m[:, None] * (p.sum(axis=0) / m.sum())
        a.set_momenta(p)


@dataclasses.dataclass
class MDFrame:
    """One recorded step: per-structure potential and kinetic energies (eV)."""
    step: int
    time_fs: float
    potential_ev: np.ndarray
    kinetic_ev: np.ndarray

    @property
    def total_ev(self):
        return self.potential_ev + self.kinetic_ev


class BatchedVelocityVerlet:
    """Velocity Verlet (ase/md/verlet.py) for many structures at once.

    structures: Atoms-like objects (ASE Atoms or pm3amd.ase.Structure) whose positions and momenta
    this driver updates in place.  calc: the PM3 calculator whose selectors (charge, multiplicity,
    reference, field, options) apply to every structure.  warm_start starts every SCF after the
    first from the structure's density at the previous step.  evaluator, when given, replaces the
    engine: a callable taking a list of (n_k, 3) position arrays (Angstrom) and returning
    (energies eV, list of (n_k, 3) forces eV/Angstrom) -- for model potentials; without it every
    force comes from the GPU engine.
    """

    def __init__(self, structures: Sequence, timestep_fs: float, calc: Optional[PM3] = None, warm_start=True,
                 evaluator: Optional[Callable] = None):
        if not structures:
            raise ValueError("no structures")
        self.structures = list(structures)
        self.dt = float(timestep_fs) * FS
        self.timestep_fs = float(timestep_fs)
        self.calc = calc or PM3()
        self.evaluator = evaluator
        self.warm_start = bool(warm_start)
        self.masses = [np.asarray(a.get_masses(), dtype=np.float64)[:, None] for a in self.structures]
        self.nsteps = 0
        self.forces: Optional[List[np.ndarray]] = None
        self.energies: Optional[np.ndarray] = None
        self.scf_iterations = 0
        self._batch: Optional[Batch] = None

    def _engine(self, positions):
        mols = [Molecule(np.asarray(a.get_atomic_numbers(), np.int32), np.asarray(x) * ANGSTROM_TO_BOHR,
                         float(self.calc.charge), self.calc.multiplicity) for a, x in zip(self.structures, positions)]
        if self._batch is None:
            self.calc._check_method()
            self._batch = Batch(mols, self.calc.params, self.calc.engine_options(), want_gradient=True,
                                warm_start=self.warm_start)
        else:
            self._batch.set_positions([m.positions for m in mols])
        self._batch.solve()
        status, res = self._batch.results()
        for k, (s, r) in enumerate(zip(status, res)):
            if s != 0 or not r.converged:
                cls = _ERRORS.get(int(s), Pm3Error) if s != 0 else ScfNotConverged
                raise cls(f"structure {k} at MD step {self.nsteps}: {cls.__name__} (status {int(s)})")
        self.scf_iterations += sum(int(r.iterations) for r in res)
        grads = self._batch.gradients()
        return (np.array([r.total_ev for r in res]), [-g * ANGSTROM_TO_BOHR for g in grads])

    def _evaluate(self):
        positions = [np.asarray(a.get_positions(), dtype=np.float64) for a in self.structures]
        e, f = (self.evaluator or self._engine)(positions)
        self.energies = np.asarray(e, dtype=np.float64)
        self.forces = [np.asarray(v, dtype=np.float64).reshape(-1, 3) for v in f]

    def kinetic_energies(self):
        return np.array([0.5 * float(np.sum(np.asarray(a.get_momenta()) ** 2 / m))
                         for a, m in zip(self.structures, self.masses)])

    def potential_energies(self):
        if self.energies is None:
            self._evaluate()
        return self.energies.copy()

    def step(self):
        """One velocity-Verlet step of every structure (ase/md/verlet.py VelocityVerlet.step)."""
        if self.forces is None:
            self._evaluate()
        dt = self.dt
        for a, m, f in zip(self.structures, self.masses, self.forces):
            p = np.asarray(a.get_momenta(), dtype=np.float64) + 0.5 * dt * f
            a.set_positions(np.asarray(a.get_positions(), dtype=np.float64) + dt * p / m)
            a.set_momenta(p)
        self._evaluate()
        for a, f in zip(self.structures, self.forces):
            a.set_momenta(np.asarray(a.get_momenta(), dtype=np.float64) + 0.5 * dt * f)
        self.nsteps += 1

    def run(self, steps: int, record_every: int = 1) -> List[MDFrame]:
        """Advance every structure by steps; returns the frames at step 0 and every record_every
        steps after it."""
        frames = []
        if self.forces is None:
            self._evaluate()

        def frame():
            frames.append(MDFrame(self.nsteps, self.nsteps * self.timestep_fs, self.energies.copy(),
                                  self.kinetic_energies()))

        frame()
        for i in range(int(steps)):
            self.step()
            if record_every and (i + 1) % record_every == 0:
                frame()
        return frames

    def close(self):
        if self._batch is not None:
            self._batch.close()
            self._batch = None
