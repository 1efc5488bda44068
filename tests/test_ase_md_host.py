"""The ASE-style calculator and the batched MD driver on the host: unit conversions, the
calculator protocol (caching, invalidation, refusals) and the velocity-Verlet integrator against
closed forms, with no device call."""
import math

import numpy as np
import pytest


def test_units_match_ase_and_the_reference():
    from pm3amd import ase as PA
    from pm3amd import md
    # ase.units.fs = 1e-15 s in ASE's time unit Angstrom * sqrt(amu / eV)
    assert abs(md.FS - 0.09822694788464063) < 1e-15
    assert abs(PA.DEBYE_TO_E_ANGSTROM - 0.2081943) == 0.0  # python/pm3_rs/ase.py:44
    # src/python.rs:198-212: V/Angstrom -> eV/Bohr with the sign of E = E0 + mu.F
    f = PA.field_to_engine([1.0, -2.0, 0.5])
    assert f == (-0.529177210903, 2 * 0.529177210903, -0.5 * 0.529177210903)
    assert PA.field_to_engine(None) is None
    import pm3amd as G
    with pytest.raises(G.InvalidInput):
        PA.field_to_engine([1.0, 2.0])


def test_calculator_refusals_need_no_device():
    import pm3amd as G
    from pm3amd.ase import PM3, Structure
    with pytest.raises(G.Unsupported):
        PM3(method="pm3-d3")
    calc = PM3()
    s = Structure([8, 1, 1], [[0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0]], pbc=[True, False, False])
    with pytest.raises(G.Unsupported):
        calc.get_potential_energy(s)
    with pytest.raises(Exception):
        calc.get_property("hessian", s)


class _Counting:
    """PM3 with the engine call replaced by a harmonic model, to test the caching protocol."""

    def __init__(self):
        from pm3amd.ase import PM3

        class C(PM3):
            calls = 0

            def calculate(self, atoms=None, properties=("energy",), system_changes=None):
                super(PM3, self).calculate(atoms, properties, system_changes)
                self._selectors_at_calculate = self._selectors()
                C.calls += 1
                x = np.asarray(self.atoms.get_positions())
                self.results.update(energy=float(np.sum(x ** 2)) + self.charge, forces=-2 * x, free_energy=0.0,
                                    charges=np.zeros(len(x)), dipole=np.zeros(3), electronic_entropy_ts_ev=0.0,
                                    heat_of_formation_kcal=0.0)
        self.cls = C


def test_results_are_cached_per_structure_and_selectors():
    from pm3amd.ase import Structure
    C = _Counting().cls
    calc = C()
    s = Structure([1, 1], [[0, 0, 0], [0.74, 0, 0]])
    e1 = calc.get_potential_energy(s)
    calc.get_forces(s)
    calc.get_charges(s)
    assert C.calls == 1  # one evaluation serves every property
    s.set_positions([[0, 0, 0], [0.75, 0, 0]])
    assert calc.get_potential_energy(s) != e1 and C.calls == 2
    # a changed selector invalidates the results (python/pm3_rs/ase.py:145-163)
    calc.charge = 1
    assert calc.get_potential_energy(s) == pytest.approx(0.75 ** 2 + 1) and C.calls == 3
    f = calc.get_forces(s)
    f[0, 0] = 99.0  # results are handed out as copies
    assert calc.get_forces(s)[0, 0] != 99.0


def test_velocity_verlet_follows_the_discrete_harmonic_solution():
    """A single atom on a spring from rest: velocity Verlet gives x_n = A cos(n theta) exactly, with
    cos theta = 1 - (omega dt)^2 / 2 (a misplaced half kick or a wrong time unit breaks it)."""
    from pm3amd.ase import Structure
    from pm3amd.md import FS, BatchedVelocityVerlet
    k, mass, A, dt_fs = 3.0, 2.5, 0.1, 0.7
    s = Structure([1], [[A, 0.0, 0.0]], masses=[mass])

    def spring(positions):
        return [0.5 * k * float(np.sum(p ** 2)) for p in positions], [-k * p for p in positions]

    md = BatchedVelocityVerlet([s], dt_fs, evaluator=spring)
    omega_dt = math.sqrt(k / mass) * dt_fs * FS
    theta = math.acos(1.0 - 0.5 * omega_dt ** 2)
    frames = md.run(40)
    assert len(frames) == 41 and frames[-1].step == 40 and frames[-1].time_fs == pytest.approx(40 * dt_fs)
    assert abs(s.get_positions()[0, 0] - A * math.cos(40 * theta)) < 1e-12
    # the energy of the discrete flow is conserved to O(dt^2) and never drifts
    tot = np.array([f.total_ev[0] for f in frames])
    assert np.abs(tot - tot[0]).max() < 0.5 * omega_dt ** 2 * tot[0]


def test_velocity_verlet_moves_each_structure_independently():
    from pm3amd.ase import Structure
    from pm3amd.md import BatchedVelocityVerlet

    def spring(positions):
        return [float(np.sum(p ** 2)) for p in positions], [-2.0 * p for p in positions]

    a = [Structure([1, 1], [[0.3, 0, 0], [0, 0.2, 0]]), Structure([6], [[0, 0, -0.4]])]
    b = [Structure([1, 1], [[0.3, 0, 0], [0, 0.2, 0]])]
    c = [Structure([6], [[0, 0, -0.4]])]
    for group in (a, b, c):
        BatchedVelocityVerlet(group, 0.5, evaluator=spring).run(25)
    assert np.array_equal(a[0].get_positions(), b[0].get_positions())
    assert np.array_equal(a[1].get_positions(), c[0].get_positions())


def test_maxwell_boltzmann_temperature():
    from pm3amd.ase import Structure
    from pm3amd.md import KB_EV, maxwell_boltzmann
    rng = np.random.default_rng(3)
    ss = [Structure([8, 1, 1], np.zeros((3, 3))) for _ in range(4000)]
    maxwell_boltzmann(ss, 300.0, rng)
    ke = np.array([s.get_kinetic_energy() for s in ss])
    # 9 degrees of freedom less the 3 of the removed centre-of-mass motion
    assert abs(ke.mean() / (3.0 * KB_EV * 300.0) - 1.0) < 0.03
    p = ss[0].get_momenta()
    assert np.abs(p.sum(axis=0)).max() < 1e-12
