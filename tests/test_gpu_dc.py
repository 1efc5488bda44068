"""Divide-and-conquer SCF on the GPU (pm3_run_dc: run_dc, src/dc/scf.rs:74-160) against the
oracle's restatement (orc_run_dc) on the same inputs: restricted and unrestricted, several
partitions of water chains and a 3D water cluster, to |dHf| <= 1e-5 kcal/mol and max|dP| <= 1e-8
with both sides converged tightly, with the same partition statistics and chemical potential; at
the default tolerances the two trajectories also take the same number of iterations (converged to
1e-10, the density DIIS works on residuals near rounding noise and the counts drift by a few either
way while the answers agree to 1e-11); full coverage against the GPU's own run_pm3; the refusals."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = 1.0 / 0.529177210903
HOF_TOL = 1e-5
P_TOL = 1e-8
TIGHT = dict(p_tol=1e-10, e_tol=1e-12, max_scf=400)


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def water_chain(n, spacing=3.2):
    """n waters along x (the reference's own D&C test system, src/dc/scf.rs:767-777)."""
    z, x = [], []
    for i in range(n):
        o = spacing * i
        z += [8, 1, 1]
        x += [[o, 0.0, 0.0], [o + 0.9584, 0.0, 0.0], [o - 0.24, 0.9278, 0.0]]
    return np.array(z, np.int32), np.array(x) * A


def _both(G, oracle, z, x, charge=0.0, mult=1, dc=None, **opt):
    dc = dc or {}
    o = dict(TIGHT, **opt)
    g = G.run_dc(G.Molecule(z, x, charge, mult), options=G.Pm3Options(**o), dc=G.DcOptions(**dc))
    st, r = oracle.run_dc(z, x, charge, mult, opts=oracle.options(**o), dc=oracle.dc_options(**dc))
    assert st == 0
    return g, r


def _compare(g, r, name):
    dh = abs(g.heat_of_formation_kcal - r["hof_kcal"])
    dp = float(np.abs(g.density - r["density"]).max())
    print(f"{name}: subsystems {g.n_subsystems} largest {g.largest_subsystem} iterations {g.iterations}/"
          f"{r['iterations']} |dHf| {dh:.2e} max|dP| {dp:.2e} mu {g.fermi_ev:.6f}/{r['fermi_ev']:.6f}", flush=True)
    assert g.converged
    assert g.n_subsystems == r["n_subsystems"] and g.largest_subsystem == r["largest_subsystem"]
    assert g.dropped_pairs == r["dropped_pairs"]
    assert g.unrestricted == bool(r["unrestricted"])
    assert dh <= HOF_TOL, name
    assert dp <= P_TOL, name
    assert abs(g.fermi_ev - r["fermi_ev"]) <= 1e-6, name
    assert np.abs(g.charges - r["charges"]).max() <= 1e-7, name
    if g.unrestricted:
        assert np.abs(g.spin_density - r["spin_density"]).max() <= P_TOL, name


@pytest.mark.parametrize("n,dc", [(4, dict(core_radius=3.0, buffer_radius=8.0)),
                                  (6, dict(core_radius=3.0, buffer_radius=6.0)),
                                  (8, dict()),
                                  (8, dict(core_radius=2.0, buffer_radius=4.0, min_core_atoms=2))])
def test_water_chains_match_the_oracle(G, oracle, n, dc):
    z, x = water_chain(n)
    g, r = _both(G, oracle, z, x, dc=dc)
    assert g.n_subsystems > 1
    _compare(g, r, f"chain{n} {dc}")


@pytest.mark.parametrize("charge,mult", [(0.0, 1), (1.0, 2)])
def test_default_tolerances_take_the_same_iterations(G, oracle, charge, mult):
    z, x = water_chain(6)
    dc = dict(core_radius=3.0, buffer_radius=6.0)
    g = G.run_dc(G.Molecule(z, x, charge, mult), options=G.Pm3Options(max_scf=400), dc=G.DcOptions(**dc))
    st, r = oracle.run_dc(z, x, charge, mult, opts=oracle.options(max_scf=400), dc=oracle.dc_options(**dc))
    assert st == 0
    print(f"chain6 q={charge}: iterations {g.iterations}/{r['iterations']} "
          f"|dHf| {abs(g.heat_of_formation_kcal - r['hof_kcal']):.2e}", flush=True)
    assert g.iterations == r["iterations"]
    assert abs(g.heat_of_formation_kcal - r["hof_kcal"]) <= HOF_TOL


def test_a_water_cluster_with_the_default_partition(G, oracle):
    from pm3amd import workloads as W
    z, x = W.water_cluster(24, 9.0, 31)
    z = np.asarray(z, np.int32)
    g, r = _both(G, oracle, z, W.to_bohr(x), dc=dict(core_radius=4.0, buffer_radius=6.0))
    _compare(g, r, "water24")


@pytest.mark.parametrize("charge,mult", [(1.0, 2), (-1.0, 2), (0.0, 3)])
def test_open_shells_match_the_oracle(G, oracle, charge, mult):
    z, x = water_chain(4)
    g, r = _both(G, oracle, z, x, charge, mult, dc=dict(core_radius=3.0, buffer_radius=8.0))
    assert g.unrestricted and abs(np.trace(g.spin_density) - (mult - 1)) < 1e-6
    _compare(g, r, f"chain4 q={charge} m={mult}")


def test_full_coverage_reproduces_run_pm3(G):
    """A buffer reaching every atom leaves no pair dropped: the partitioned SCF converges onto the
    full single point (the reference's test of the same, src/dc/scf.rs:906-930)."""
    z, x = water_chain(3)
    full = G.run_pm3(G.Molecule(z, x), options=G.Pm3Options(**TIGHT))
    g = G.run_dc(G.Molecule(z, x), options=G.Pm3Options(**TIGHT),
                 dc=G.DcOptions(core_radius=2.0, buffer_radius=500.0, smearing_ev=1e-4))
    assert g.dropped_pairs == 0 and g.n_subsystems > 1
    assert abs(g.total_ev - full.total_ev) < 1e-6
    assert np.abs(g.density - full.density).max() < 1e-5


@pytest.mark.parametrize("charge,mult,field", [(0.0, 1, None), (1.0, 2, None), (0.0, 1, (0.01, -0.02, 0.005))])
def test_dc_gradient_matches_the_oracle(G, oracle, charge, mult, field):
    """dc_gradient (src/dc/derivatives.rs:55-121): the fixed-density gradient at the partitioned
    density, the spin form for an open shell, minus f q_A in a field, to 2.295e-7 eV/Bohr."""
    z, x = water_chain(5)
    dc = dict(core_radius=3.0, buffer_radius=6.0)
    o = dict(TIGHT, field=field) if field else dict(TIGHT)
    g = G.dc_gradient(G.Molecule(z, x, charge, mult), options=G.Pm3Options(**o), dc=G.DcOptions(**dc))
    oo = oracle.options(**TIGHT)
    if field:
        oo = oracle.options(**TIGHT, field=field)
    st, r = oracle.dc_gradient(z, x, charge, mult, opts=oo, dc=oracle.dc_options(**dc))
    assert st == 0
    dg = float(np.abs(g.gradient - r["gradient"]).max())
    print(f"dc_gradient q={charge} field={field}: max|dg| {dg:.2e} eV/Bohr, |dHf| "
          f"{abs(g.scf.heat_of_formation_kcal - r['hof_kcal']):.2e}", flush=True)
    assert dg <= 2.295e-7
    assert abs(g.scf.heat_of_formation_kcal - r["hof_kcal"]) <= HOF_TOL
    assert abs(g.gradient.sum(axis=0)).max() < 1e-6 or field


def test_refusals(G):
    z, x = water_chain(2)
    m = G.Molecule(z, x)
    with pytest.raises(G.InvalidInput):
        G.run_dc(m, dc=G.DcOptions(core_radius=0.0))
    with pytest.raises(G.InvalidInput):
        G.run_dc(m, dc=G.DcOptions(buffer_radius=-1.0))
    with pytest.raises(G.InvalidInput):
        G.run_dc(G.Molecule(z, x, 0.0, 2), options=G.Pm3Options(reference="rhf"))
    with pytest.raises(G.ScfNotConverged):
        G.run_dc(m, options=G.Pm3Options(max_scf=2))


@pytest.mark.parametrize("charge,mult", [(0.0, 1), (1.0, 2)])
def test_optimize_dc_follows_the_oracle(G, oracle, charge, mult):
    """optimize_dc (src/optimizer.rs:277-310): the L-BFGS of optimize over dc_gradient, on a distorted
    three-water chain split into three subsystems, against the oracle's restatement of the same
    optimiser over its own dc_gradient, step by step for ten iterations: the same accepted points,
    energies and final geometry.  (The partitioned density is not variational, so on this floppy
    chain neither side reaches max |g| < 1e-3; the reference's docstring warns of exactly that.)"""
    z, x = water_chain(3, spacing=3.0)
    x = x + np.random.default_rng(5).normal(0.0, 0.05 * A, x.shape)
    dc = dict(core_radius=3.0, buffer_radius=8.0)
    r = G.optimize_dc(G.Molecule(z, x, charge, mult), scf_options=G.Pm3Options(**TIGHT), dc=G.DcOptions(**dc),
                      opt=G.OptOptions(max_iter=10))
    o = oracle.optimize_dc(z, x, charge, mult, oracle.options(**TIGHT), oracle.dc_options(**dc), max_iter=10)
    print(f"optimize_dc q={charge}: iterations {r.iterations}/{o['iterations']} evaluations {r.evaluations} "
          f"HoF {r.heat_of_formation_kcal:.6f}/{o['hof_kcal']:.6f} max|g| {r.max_gradient:.2e} "
          f"subsystems {r.scf.n_subsystems}", flush=True)
    assert r.scf.n_subsystems > 1 and r.scf.unrestricted == (mult != 1)
    assert r.converged == o["converged"] and r.iterations == o["iterations"]
    assert r.evaluations >= o["evaluations"]  # every solve, trials included, against gradient calls
    assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= 1e-5
    assert np.abs(r.molecule.positions - o["xyz"]).max() <= 1e-5
    assert abs(r.first_energy_ev - o["trajectory"][0]) <= 1e-6 and r.energy_ev < r.first_energy_ev
    assert len(r.trajectory) == len(o["steps"]) == 11
    for a, b in zip(r.trajectory, o["steps"]):
        assert abs(a.energy_ev - b["energy_ev"]) <= 1e-6
        assert abs(a.max_gradient - b["max_gradient"]) <= 1e-5
        assert np.abs(a.positions - b["xyz"]).max() <= 1e-5
    assert r.scf.total_ev == r.energy_ev


def test_optimize_dc_first_evaluation_failure_raises(G):
    z, x = water_chain(2)
    with pytest.raises(G.ScfNotConverged):
        G.optimize_dc(G.Molecule(z, x), scf_options=G.Pm3Options(max_scf=2))
