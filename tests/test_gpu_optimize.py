"""Geometry optimisation on the GPU (pm3_optimize / pm3_optimize_batch): the reference's L-BFGS
(src/optimizer.rs:77-262) over batched SCF + analytic gradient solves, against the oracle's
restatement of the same optimiser on the CPU path; and the position upload it is built on
(pm3_batch_set_positions) against freshly created batches."""
import numpy as np
import pytest

from pm3amd import workloads as W

pytestmark = pytest.mark.gpu

A = 1.0 / 0.529177210903
TIGHT = dict(p_tol=1e-10, e_tol=1e-12)


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def _cases():
    rng = np.random.default_rng(31)
    out = [("water", [8, 1, 1], np.array([[0.0, 0.0, 0.0], [1.05, 0.0, 0.0], [-0.30, 1.02, 0.0]])),
           ("formaldehyde", [6, 8, 1, 1], np.array([[0, 0, 0], [0.05, 0, 1.25], [0.94, 0.1, -0.54], [-0.94, 0, -0.5]])),
           ("methane", [6, 1, 1, 1, 1], np.array([[0, 0, 0], [0.7, 0.6, 0.62], [-0.6, -0.66, 0.6], [-0.62, 0.6, -0.7],
                                                  [0.6, -0.6, -0.6]]))]
    z, x = W.alkane(3)
    out.append(("propane", list(z), W.jitter(np.asarray(x), 0.08, rng)))
    out.append(("methanol", [6, 8, 1, 1, 1, 1], np.array([[0, 0, 0], [1.42, 0.02, 0], [1.75, 0.93, 0.05],
                                                           [-0.37, 1.03, 0.02], [-0.38, -0.5, 0.9],
                                                           [-0.36, -0.52, -0.88]])))
    return [(name, np.asarray(z, np.int32), np.asarray(x, np.float64) * A) for name, z, x in out]


def _dimer():
    z, x = W.water_cluster(2, 6.0, 7)
    return np.asarray(z, np.int32), np.asarray(x, np.float64) * A


def test_optimisations_follow_the_oracle(G, oracle):
    for name, z, x in _cases():
        r = G.optimize(G.Molecule(z, x), scf_options=G.Pm3Options(**TIGHT))
        o = oracle.optimize(z, x, 0.0, oracle.options(**TIGHT))
        print(f"case {name} iterations {r.iterations} / {o['iterations']} evaluations {r.evaluations} "
              f"dHf {r.heat_of_formation_kcal - o['hof_kcal']:.1e} max|dx| "
              f"{np.abs(r.molecule.positions - o['xyz']).max():.1e}", flush=True)
        assert r.converged and o["converged"] and r.max_gradient < 1e-3
        assert r.iterations == o["iterations"], name
        assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= 1e-5, name
        assert np.abs(r.molecule.positions - o["xyz"]).max() <= 1e-5, name
        assert abs(r.first_energy_ev - o["trajectory"][0]) <= 1e-6 and r.energy_ev <= r.first_energy_ev
        # one solve per accepted point less than the reference's energy-then-gradient evaluations
        assert r.evaluations >= r.iterations
        # the trajectory (OptResult::trajectory): the same accepted points, step for step
        assert len(r.trajectory) == len(o["steps"]), name
        # (on soft modes -- methanol's methyl rotation -- the two paths part by up to ~1e-5 Bohr
        # midway and meet again at the minimum, so midway points get the looser bounds)
        for a, b in zip(r.trajectory, o["steps"]):
            assert abs(a.energy_ev - b["energy_ev"]) <= 1e-5
            assert abs(a.heat_of_formation_kcal - b["hof_kcal"]) <= 2.5e-4
            assert abs(a.max_gradient - b["max_gradient"]) <= 1e-4
            assert np.abs(a.positions - b["xyz"]).max() <= 1e-4
        # each step's scalars belong to its own positions: the oracle at the GPU's points
        for a in (r.trajectory[0], r.trajectory[len(r.trajectory) // 2], r.trajectory[-1]):
            st, g = oracle.gradient(z, a.positions, 0.0, oracle.options(**TIGHT))
            assert st == 0 and abs(g["total_ev"] - a.energy_ev) <= 1e-7
            assert abs(np.abs(g["gradient"]).max() - a.max_gradient) <= 1e-6
        assert np.array_equal(r.trajectory[0].positions, x) and np.array_equal(r.trajectory[-1].positions,
                                                                                 r.molecule.positions)
        assert r.trajectory[-1].energy_ev == r.energy_ev and r.trajectory[0].energy_ev == r.first_energy_ev


def test_open_shell_optimisation_follows_the_oracle(G, oracle):
    """A doublet (ethyl radical, distorted) optimised on the unrestricted surface: the UHF
    energy and gradient at every point, the trajectory step for step against the oracle."""
    z = np.array([6, 6, 1, 1, 1, 1, 1], np.int32)
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [-0.4, 1.0, 0.1], [-0.4, -0.5, 0.9], [-0.35, -0.5, -0.85],
                  [2.0, 0.95, 0.1], [2.05, -0.9, -0.1]]) * A
    r = G.optimize(G.Molecule(z, x, multiplicity=2), scf_options=G.Pm3Options(**TIGHT))
    o = oracle.optimize(z, x, 0.0, oracle.options(**TIGHT), multiplicity=2)
    print("ethyl", r.iterations, o["iterations"], r.heat_of_formation_kcal - o["hof_kcal"], flush=True)
    assert r.converged and o["converged"]
    assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= 1e-4
    n = min(len(r.trajectory), len(o["steps"]), 10)
    for a, b in zip(r.trajectory[:n], o["steps"][:n]):
        assert abs(a.energy_ev - b["energy_ev"]) <= 1e-5
        assert np.abs(a.positions - b["xyz"]).max() <= 1e-4


def test_flat_surface_reaches_the_oracles_basin(G, oracle):
    """A water dimer: its intermolecular modes are nearly flat, so gradient differences of 1e-9
    eV/Bohr grow over the hundred-odd steps until an Armijo decision differs and the two
    trajectories part; both still end in the same minimum."""
    z, x = _dimer()
    r = G.optimize(G.Molecule(z, x), scf_options=G.Pm3Options(**TIGHT))
    o = oracle.optimize(z, x, 0.0, oracle.options(**TIGHT))
    print("dimer", r.iterations, o["iterations"], r.heat_of_formation_kcal - o["hof_kcal"], flush=True)
    assert r.converged and o["converged"]
    assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= 1e-3


def test_reference_water_run_on_the_gpu(G):
    # src/optimizer.rs:362-386 with the default SCF options
    z = np.array([8, 1, 1], np.int32)
    x = np.array([[0.0, 0.0, 0.0], [1.05, 0.0, 0.0], [-0.30, 1.02, 0.0]]) * A
    r = G.optimize(G.Molecule(z, x))
    print("water", r.iterations, r.heat_of_formation_kcal, r.max_gradient, flush=True)
    assert r.converged and abs(r.heat_of_formation_kcal - (-53.4330121104622)) < 1e-4


def test_batch_gives_each_molecule_its_single_run(G):
    """Molecules of different sizes finish after different numbers of rounds, so the batch is
    rebuilt from the running ones several times; each result is still the molecule's own run."""
    rng = np.random.default_rng(5)
    mols = [G.Molecule(z, x) for _, z, x in _cases()] + [G.Molecule(*_dimer())]
    for z, x in W.mixed_batch(10, 44, sigma=0.08):
        mols.append(G.Molecule(np.asarray(z, np.int32), W.to_bohr(W.jitter(np.asarray(x), 0.03, rng))))
    opts = G.Pm3Options(**TIGHT)
    st, rs = G.optimize_batch(mols, options=opts)
    assert np.all(st == 0)
    for m, r in zip(mols, rs):
        # (a floppy water cluster may reach max_iter unconverged, as the reference's would)
        one = G.optimize(m, scf_options=opts)
        assert r.converged == one.converged
        assert r.iterations == one.iterations and r.evaluations == one.evaluations
        assert abs(r.energy_ev - one.energy_ev) <= 1e-10
        assert np.abs(r.molecule.positions - one.molecule.positions).max() <= 1e-10
        assert len(r.trajectory) == len(one.trajectory) >= 1
        assert np.abs(r.trajectory[-1].positions - one.trajectory[-1].positions).max() <= 1e-10
    print("iterations", [r.iterations for r in rs], "converged", sum(r.converged for r in rs), flush=True)
    assert all(r.converged for r in rs[:5])  # the rigid molecules
    assert len({r.iterations for r in rs}) > 1


def test_new_positions_equal_a_fresh_batch(G):
    rng = np.random.default_rng(9)
    base = [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(12, 3)]
    moved = [W.jitter(x, 0.05, rng) for _, x in base]
    opts = G.Pm3Options(**TIGHT)
    b = G.Batch([G.Molecule(z, x) for z, x in base], options=opts, want_gradient=True)
    b.solve()
    b.set_positions(moved)
    b.solve()
    s1, r1 = b.results()
    g1 = b.gradients()
    f = G.Batch([G.Molecule(z, x) for (z, _), x in zip(base, moved)], options=opts, want_gradient=True)
    f.solve()
    s2, r2 = f.results()
    g2 = f.gradients()
    f.close()
    assert np.all(s1 == 0) and np.all(s2 == 0)
    for a, c, ga, gc in zip(r1, r2, g1, g2):
        assert a.heat_of_formation_kcal == c.heat_of_formation_kcal and a.iterations == c.iterations
        assert np.array_equal(ga, gc)
    # refused positions leave the batch as it was
    bad = [x.copy() for x in moved]
    bad[4][1] = bad[4][0]
    with pytest.raises(G.InvalidInput):
        b.set_positions(bad)
    b.solve()
    s3, r3 = b.results()
    assert all(a.heat_of_formation_kcal == c.heat_of_formation_kcal for a, c in zip(r3, r2))
    b.close()


def test_warm_start_keeps_the_answers_in_fewer_iterations(G):
    """PM3_BATCH_WARM_START: after a small move, each molecule's SCF starts from its own last
    density; the converged answers are a cold batch's, the SCF paths shorter."""
    rng = np.random.default_rng(12)
    base = [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(16, 21)]
    moved = [x + rng.normal(0.0, 0.02, x.shape) for _, x in base]
    opts = G.Pm3Options(**TIGHT)
    b = G.Batch([G.Molecule(z, x) for z, x in base], options=opts, want_gradient=True, warm_start=True)
    b.solve()
    b.set_positions(moved)
    b.solve()
    s1, r1 = b.results()
    g1 = b.gradients()
    b.close()
    f = G.Batch([G.Molecule(z, x) for (z, _), x in zip(base, moved)], options=opts, want_gradient=True)
    f.solve()
    s2, r2 = f.results()
    g2 = f.gradients()
    f.close()
    assert np.all(s1 == 0) and np.all(s2 == 0)
    warm, cold = sum(r.iterations for r in r1), sum(r.iterations for r in r2)
    print("iterations warm", warm, "cold", cold, flush=True)
    assert warm < cold
    for a, c, ga, gc in zip(r1, r2, g1, g2):
        assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= 1e-7
        assert np.abs(ga - gc).max() <= 1e-6


def test_warm_optimisations_reach_the_oracles_minima(G, oracle):
    opt = G.OptOptions(warm_start=True)
    for name, z, x in _cases():
        r = G.optimize(G.Molecule(z, x), scf_options=G.Pm3Options(**TIGHT), opt=opt)
        o = oracle.optimize(z, x, 0.0, oracle.options(**TIGHT))
        print(f"warm {name} iterations {r.iterations} / {o['iterations']} dHf {r.heat_of_formation_kcal - o['hof_kcal']:.1e}",
              flush=True)
        assert r.converged and r.max_gradient < 1e-3
        assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= 1e-4, name


def test_warm_start_skips_densities_of_unconverged_solves(G):
    """A molecule whose last solve did not converge starts again from the options' guess: with
    max_scf too small to converge, a warm batch at the same geometry gives a cold batch's answer
    bit for bit, solve after solve."""
    mols = [G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(6, 8)]
    opts = G.Pm3Options(max_scf=3, p_tol=1e-10, e_tol=1e-12)
    w = G.Batch(mols, options=opts, want_gradient=True, warm_start=True)
    c = G.Batch(mols, options=opts, want_gradient=True)
    for _ in range(2):
        w.solve()
        c.solve()
        sw, rw = w.results()
        sc, rc = c.results()
        assert np.all(sw == 3) and np.all(sc == 3)  # PM3_SCF_NOT_CONVERGED
        assert all(a.electronic_ev == b.electronic_ev for a, b in zip(rw, rc))
    w.close()
    c.close()


def test_warm_batch_gives_each_molecule_its_warm_single_run(G):
    """With warm starts the densities follow the running molecules into every rebuilt batch
    (pm3_batch_carry_densities), so each result is still the molecule's own warm run."""
    rng = np.random.default_rng(6)
    mols = [G.Molecule(z, x) for _, z, x in _cases()]
    for z, x in W.mixed_batch(6, 45, sigma=0.08):
        mols.append(G.Molecule(np.asarray(z, np.int32), W.to_bohr(W.jitter(np.asarray(x), 0.03, rng))))
    opts = G.Pm3Options(**TIGHT)
    opt = G.OptOptions(warm_start=True)
    st, rs = G.optimize_batch(mols, options=opts, opt=opt)
    assert np.all(st == 0)
    for m, r in zip(mols, rs):
        one = G.optimize(m, scf_options=opts, opt=opt)
        assert r.converged == one.converged and r.iterations == one.iterations
        assert r.evaluations == one.evaluations
        assert abs(r.energy_ev - one.energy_ev) <= 1e-10
        assert np.abs(r.molecule.positions - one.molecule.positions).max() <= 1e-10
    print("warm iterations", [r.iterations for r in rs], flush=True)
