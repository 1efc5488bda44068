"""The ASE-style calculator (python/pm3_rs/ase.py, molecular part) and batched velocity-Verlet MD
on the GPU engine, against the oracle: every property in ASE units, the batched evaluation equal
to single ones, and MD trajectories that follow the same integrator driven by oracle forces."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = 1.0 / 0.529177210903
TIGHT = dict(p_tol=1e-10, e_tol=1e-12)


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def _structures():
    from pm3amd.ase import Structure
    return [Structure([8, 1, 1], [[0.0, 0.0, 0.0], [0.97, 0.02, 0.0], [-0.25, 0.93, 0.01]]),
            Structure([6, 8, 1, 1, 1, 1], [[0, 0, 0], [1.42, 0.02, 0], [1.75, 0.93, 0.05], [-0.37, 1.03, 0.02],
                                           [-0.38, -0.5, 0.9], [-0.36, -0.52, -0.88]]),
            Structure([6, 8, 1, 1], [[0, 0, 0], [0.05, 0, 1.22], [0.94, 0.1, -0.54], [-0.94, 0, -0.5]])]


def _oracle_properties(oracle, z, x_ang, charge=0.0, mult=1, field=None):
    opts = oracle.options(**TIGHT) if field is None else oracle.options(field=field, **TIGHT)
    st, g = oracle.gradient(z, x_ang * A, charge, opts, mult)
    st2, r = oracle.run_pm3(z, x_ang * A, charge, mult, opts)
    assert st == 0 and st2 == 0
    return g, r


def test_calculator_properties_against_the_oracle(G, oracle):
    from pm3amd.ase import DEBYE_TO_E_ANGSTROM, PM3, Structure, field_to_engine
    cases = [(s, 0, 1, None) for s in _structures()]
    cases.append((Structure([6, 6, 1, 1, 1, 1, 1], [[0, 0, 0], [1.5, 0, 0], [-0.4, 1.0, 0.1], [-0.4, -0.5, 0.9],
                                                     [-0.35, -0.5, -0.85], [2.0, 0.95, 0.1], [2.05, -0.9, -0.1]]),
                  0, 2, None))
    cases.append((_structures()[0], 0, 1, [0.0, 0.4, -0.2]))  # V/Angstrom
    for s, q, mult, field in cases:
        calc = PM3(charge=q, multiplicity=mult, field=field, options=G.Pm3Options(**TIGHT))
        e = calc.get_potential_energy(s)
        f = calc.get_forces(s)
        z = s.get_atomic_numbers()
        g, r = _oracle_properties(oracle, z, s.get_positions(), q, mult, field_to_engine(field))
        print("case", len(z), mult, field, e - g["total_ev"], flush=True)
        assert abs(e - g["total_ev"]) <= 1e-6
        # forces in eV/Angstrom = -gradient (eV/Bohr) * Bohr per Angstrom (src/python.rs:395-403)
        assert np.abs(f + g["gradient"] * A).max() <= 1e-5
        assert np.abs(calc.get_charges(s) - r["charges"]).max() <= 1e-6
        assert np.abs(calc.get_dipole_moment(s) - r["dipole"] * DEBYE_TO_E_ANGSTROM).max() <= 1e-6
        assert abs(calc.get_property("heat_of_formation_kcal", s) - r["hof_kcal"]) <= 1e-5
        assert calc.get_potential_energy(s, force_consistent=True) == e
        assert np.array_equal(calc.get_gradient(s), -f)


def test_forces_are_the_energy_derivative(G):
    # a central difference of the calculator's own energy along three coordinates (eV/Angstrom)
    from pm3amd.ase import PM3
    s = _structures()[1]
    calc = PM3(options=G.Pm3Options(**TIGHT))
    f = calc.get_forces(s)
    x0 = s.get_positions()
    h = 1e-4
    for i, c in ((0, 0), (1, 2), (4, 1)):
        xp, xm = x0.copy(), x0.copy()
        xp[i, c] += h
        xm[i, c] -= h
        s.set_positions(xp)
        ep = calc.get_potential_energy(s)
        s.set_positions(xm)
        em = calc.get_potential_energy(s)
        assert abs(-(ep - em) / (2 * h) - f[i, c]) <= 2e-4, (i, c)
    s.set_positions(x0)


def test_calculate_many_equals_single_evaluations(G):
    from pm3amd.ase import PM3
    calc = PM3(options=G.Pm3Options(**TIGHT))
    ss = _structures()
    many = calc.calculate_many(ss)
    for s, m in zip(ss, many):
        one = PM3(options=G.Pm3Options(**TIGHT))
        assert abs(one.get_potential_energy(s) - m["energy"]) <= 1e-10
        assert np.abs(one.get_forces(s) - m["forces"]).max() <= 1e-9
        assert np.abs(one.get_charges(s) - m["charges"]).max() <= 1e-10
        assert np.abs(one.get_dipole_moment(s) - m["dipole"]).max() <= 1e-10


def test_md_follows_the_oracle_driven_integrator(G, oracle):
    """20 steps of 0.5 fs for three molecules in one batch, against the same integrator driven by
    the oracle's forces from the same initial momenta."""
    from pm3amd.ase import PM3
    from pm3amd.md import BatchedVelocityVerlet, maxwell_boltzmann
    ss = _structures()
    maxwell_boltzmann(ss, 600.0, np.random.default_rng(7))
    ref = [s.copy() for s in ss]
    opts = oracle.options(**TIGHT)

    def oracle_forces(positions):
        es, fs = [], []
        for s, x in zip(ref, positions):
            st, g = oracle.gradient(s.get_atomic_numbers(), x * A, 0.0, opts)
            assert st == 0
            es.append(g["total_ev"])
            fs.append(-g["gradient"] * A)
        return es, fs

    md = BatchedVelocityVerlet(ss, 0.5, calc=PM3(options=G.Pm3Options(**TIGHT)))
    fr = md.run(20)
    md.close()
    fo = BatchedVelocityVerlet(ref, 0.5, evaluator=oracle_forces).run(20)
    for s, r in zip(ss, ref):
        d = np.abs(s.get_positions() - r.get_positions()).max()
        print("md max|dx|", d, flush=True)
        assert d <= 1e-7
        assert np.abs(s.get_momenta() - r.get_momenta()).max() <= 1e-6
    assert np.abs(fr[-1].potential_ev - fo[-1].potential_ev).max() <= 1e-6


def test_md_conserves_energy(G):
    from pm3amd.ase import PM3
    from pm3amd.md import BatchedVelocityVerlet, maxwell_boltzmann
    ss = _structures()
    maxwell_boltzmann(ss, 300.0, np.random.default_rng(11))
    md = BatchedVelocityVerlet(ss, 0.25, calc=PM3(options=G.Pm3Options(**TIGHT)))
    frames = md.run(200, record_every=10)
    md.close()
    tot = np.array([f.total_ev for f in frames])
    drift = np.abs(tot - tot[0]).max(axis=0)
    print("MD total-energy excursions (eV)", drift, "scf iterations", md.scf_iterations, flush=True)
    assert np.all(drift <= 2e-3)
    assert all(np.all(f.kinetic_ev > 0) for f in frames[1:])
