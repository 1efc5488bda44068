"""The CPU oracle against everything the reference itself fixes: its own frozen energies and heats
of formation, its MOPAC rows, its unit-test constants, and identities a correct PM3 code obeys."""
import math

import numpy as np
import pytest

from conftest import ANGSTROM_TO_BOHR, atoms_to_zx

PM3_EV = 27.211386245988


def test_internal_consistency_total_energies(pins, oracle):
    # tools/oracle/INTERNAL_CONSISTENCY.json: pm3-rs total energies to 12 decimals
    for h in pins["hydrides"]:
        z, x = atoms_to_zx(h["atoms"])
        st, r = oracle.run(z, x, h["charge"])
        assert st == 0, h["name"]
        assert abs(r["total_ev"] - h["energy_ev"]) < 1e-10, (h["name"], r["total_ev"], h["energy_ev"])


def test_all_elements_heats_of_formation(pins, oracle):
    # tools/oracle/ALL_ELEMENTS_RESULTS.json: pm3-rs heats of formation to 8 decimals
    for h in pins["hydrides"]:
        if h["hof_kcal"] is None:
            continue
        z, x = atoms_to_zx(h["atoms"])
        st, r = oracle.run(z, x, h["charge"])
        assert abs(r["hof_kcal"] - h["hof_kcal"]) < 1e-7, h["name"]


def test_mopac_oracle_rows(pins, oracle):
    # tests/data/mopac_oracle.tsv with the reference's own tolerances (tests/mopac_oracle.rs:33-47)
    bad = []
    for row in pins["mopac_closed_shell"]:
        z, x = atoms_to_zx(row["atoms"])
        st, r = oracle.run(z, x, row["charge"], oracle.options(max_scf=2000))
        ok = st == 0 and abs(r["hof_kcal"] - row["hof_kcal"]) < 2e-3
        ok = ok and np.abs(np.array(row["charges"]) - r["charges"]).max() < 1e-3
        if row["dipole_debye"] is not None:
            ok = ok and abs(r["dipole_debye"] - row["dipole_debye"]) < 5e-3
        if not ok:
            bad.append(row["name"])
    assert not bad, bad


HARD_CASES = ["butyne", "ketene", "tetrafluoroethene", "tetrachloroethene", "methyl_radical"]


def test_the_stability_search_rescues_the_reference_hard_cases(pins, oracle):
    # tests/mopac_oracle.rs:268-308: these converge to an excited solution from the default start
    # and agree with MOPAC only through the stability search.  For the closed shells the search
    # runs only the two unrestricted rungs (src/scf.rs:676: a rung runs when its unrestricted flag
    # equals closed_shell), so it tries 3 paths; the radical runs all seven rungs (8 paths).
    rows = {r["name"]: r for r in pins["mopac_closed_shell"] + pins["mopac_open_shell"]}
    for name in HARD_CASES:
        row = rows[name]
        z, x = atoms_to_zx(row["atoms"])
        mult = row.get("multiplicity", 1)
        st, r = oracle.run_pm3(z, x, row["charge"], mult, oracle.options(max_scf=2000))
        assert st == 0 and abs(r["hof_kcal"] - row["hof_kcal"]) < 2e-3, (name, r["hof_kcal"], row["hof_kcal"])
        assert r["scf_paths_tried"] == (3 if mult == 1 else 8), name
        assert r["scf_improvement_ev"] > 1e-3, name
        assert r["unrestricted"] == (mult != 1), name
        if mult == 1:  # a spin-pure unrestricted winner is handed back restricted
            assert not r["spin_density"].any() and r["n_beta"] == r["n_occ"], name
        st, plain = oracle.run_pm3(z, x, row["charge"], mult, oracle.options(max_scf=2000, stability=0))
        assert st == 0 and abs(plain["hof_kcal"] - row["hof_kcal"]) > 2e-3, name
        assert plain["scf_paths_tried"] == 1, name


def test_the_stability_search_changes_nothing_it_should_not(pins, oracle):
    # tests/mopac_oracle.rs:310-343: a molecule the search does not touch is bit for bit its
    # search-off answer, and the search fires on at most 20 molecules of the set
    searched = 0
    for row in pins["mopac_closed_shell"] + pins["mopac_open_shell"]:
        z, x = atoms_to_zx(row["atoms"])
        mult = row.get("multiplicity", 1)
        st, a = oracle.run_pm3(z, x, row["charge"], mult, oracle.options(max_scf=2000))
        assert st == 0, row["name"]
        if a["scf_paths_tried"] > 1:
            searched += 1
            continue
        _, b = oracle.run_pm3(z, x, row["charge"], mult, oracle.options(max_scf=2000, stability=0))
        assert abs(a["hof_kcal"] - b["hof_kcal"]) < 1e-12, row["name"]
    assert searched <= 20, searched


def test_mopac_open_shell_rows(pins, oracle):
    # the radicals, triplets and the water cation of tests/data/mopac_oracle.tsv through the UHF
    # restatement, with the reference's tolerances (tests/mopac_oracle.rs:38-47, :144-198); the
    # spin density carries exactly the unpaired electrons and the total density all electrons
    bad = []
    rows = pins["mopac_open_shell"]
    assert len(rows) >= 15
    for row in rows:
        z, x = atoms_to_zx(row["atoms"])
        st, r = oracle.run_uhf(z, x, row["charge"], row["multiplicity"], oracle.options(max_scf=500))
        n_elec = sum({1: 1, 6: 4, 7: 5, 8: 6, 14: 4, 15: 5, 16: 6, 17: 7}[int(v)] for v in z) - row["charge"]
        ok = st == 0 and r["converged"] and abs(r["hof_kcal"] - row["hof_kcal"]) < 2e-3
        ok = ok and np.abs(np.array(row["charges"]) - r["charges"]).max() < 1e-3
        if row["dipole_debye"] is not None:
            ok = ok and abs(r["dipole_debye"] - row["dipole_debye"]) < 5e-3
        ok = ok and abs(np.trace(r["spin_density"]) - (row["multiplicity"] - 1)) < 1e-9
        ok = ok and abs(np.trace(r["density"]) - n_elec) < 1e-9
        ok = ok and r["has_homo"] and r["has_lumo"] and r["homo_ev"] < r["lumo_ev"]
        if not ok:
            bad.append(row["name"])
    assert not bad, bad


def test_open_shell_entry_takes_the_restricted_path_for_singlets(oracle):
    # n_alpha == n_beta selects RHF (src/scf.rs:508-511): the same answer bit for bit, zero spin
    z, x = atoms_to_zx([["C", 0, 0, 0], ["O", 0, 0, 1.21], ["H", 0.94, 0, -0.54], ["H", -0.94, 0, -0.54]])
    st1, a = oracle.run(z, x, 0.0)
    st2, b = oracle.run_uhf(z, x, 0.0, 1)
    assert st1 == st2 == 0 and a["total_ev"] == b["total_ev"] and np.array_equal(a["density"], b["density"])
    assert not b["spin_density"].any()
    # parity between the electron count and the multiplicity (src/scf.rs:497-503)
    assert oracle.run_uhf(z, x, 0.0, 2)[0] == 1


def test_water_heat_of_formation_and_dipole(oracle):
    # src/scf.rs:1893-1905: MOPAC -53.2301514568744 kcal/mol and 1.762063533710513 D
    x = np.array([[0, 0, 0], [0.9584, 0, 0], [-0.24, 0.9278, 0]]) * ANGSTROM_TO_BOHR
    st, r = oracle.run([8, 1, 1], x)
    assert st == 0 and r["converged"]
    assert abs(r["hof_kcal"] - (-53.2301514568744)) < 0.005
    assert abs(r["dipole_debye"] - 1.762063533710513) < 0.002
    assert abs(r["charges"].sum()) < 1e-6
    assert r["charges"][0] < 0 < r["charges"][1]


def test_oh_overlap_regression(oracle):
    # src/overlap.rs:688-705: MOPAC S_ss(O-H) at 1.05 A
    w, e1b, e2a, s = oracle.pair_integrals(8, 1, np.array([1.05 * ANGSTROM_TO_BOHR, 0, 0]))
    assert abs(s[0, 0] - 0.2502982386931348) < 1e-10
    assert abs(s[1, 0]) > 0.1 and abs(s[2, 0]) < 1e-14 and abs(s[3, 0]) < 1e-14


def test_ssss_closed_form_and_rotation_invariance(oracle):
    # src/integrals.rs:585-594
    c = oracle.element(6)
    r = 2.6
    wa = oracle.pair_integrals(6, 6, np.array([r, 0, 0]))[0]
    v = np.array([0.3, -0.5, 0.8])
    wb = oracle.pair_integrals(6, 6, v / np.linalg.norm(v) * r)[0]
    assert abs(wa[0] - wb[0]) < 1e-9
    assert abs(wa[0] - PM3_EV / math.sqrt(r * r + (2 * c["rho0"]) ** 2)) < 1e-9
    # HH closed form: PM3_EV / sqrt(r^2 + (rho0_a + rho0_b)^2)
    h = oracle.element(1)
    whh = oracle.pair_integrals(1, 1, np.array([0.0, 1.4, 0.3]))[0]
    assert abs(whh[0] - PM3_EV / math.sqrt(1.4 ** 2 + 0.3 ** 2 + (2 * h["rho0"]) ** 2)) < 1e-12


def test_w_table_rotation_covariance(oracle):
    # The packed W transforms as a fourth-rank tensor: rotating the pair rotates the p indices.
    rng = np.random.default_rng(3)
    d = np.array([1.3, -0.7, 2.1])
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    w1 = oracle.pair_integrals(6, 8, d)[0].reshape(10, 10)
    w2 = oracle.pair_integrals(6, 8, Q @ d)[0].reshape(10, 10)
    R = np.eye(4)
    R[1:, 1:] = Q

    def unpack(w):
        pk = lambda a, b: max(a, b) * (max(a, b) + 1) // 2 + min(a, b)
        t = np.zeros((4, 4, 4, 4))
        for a in range(4):
            for b in range(4):
                for c_ in range(4):
                    for e in range(4):
                        t[a, b, c_, e] = w[pk(a, b), pk(c_, e)]
        return t

    t1 = np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, unpack(w1))
    assert np.abs(t1 - unpack(w2)).max() < 1e-10


def test_derived_multipoles(oracle):
    # src/params.rs:826-882
    h = oracle.element(1)
    assert h["n_orb"] == 1 and h["dd"] == 0.0
    for zz in (6, 7, 8):
        e = oracle.element(zz)
        assert e["rho0"] > 0 and e["dd"] > 0 and e["qq"] > 0 and e["rho1"] > 0 and e["rho2"] > 0
        assert e["po1"] == e["rho0"] and e["po2"] == e["rho1"] and e["po3"] == e["rho2"]
        assert abs(e["ddp3"] - e["qq"] * math.sqrt(2)) < 1e-15 and e["ddp2"] == e["dd"]
    # rho0 closed form 1/2 * EV / Gss
    assert abs(oracle.element(8)["rho0"] - 0.5 * PM3_EV / 15.75576) < 1e-12


def _pattern_density(n):
    P = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            P[i, j] = P[j, i] = 0.37 * ((i * 7 + j * 3) % 11) - 1.1
    return P


def test_packed_fock_equals_four_index(oracle):
    # src/fock.rs:566-602: packed two-centre build vs the literal four-index loops, to 1e-12
    z, x = atoms_to_zx([["C", 0, 0, 0], ["O", 0, 0, 1.21], ["H", 0.94, 0, -0.54], ["H", -0.94, 0, -0.54],
                        ["O", 1.5, 1.2, 0.3]])
    n = oracle.nao(z)
    P = _pattern_density(n)
    F1 = oracle.fock(z, x, P)
    F2 = oracle.fock(z, x, P, four_index=True)
    assert np.abs(F1 - F2).max() < 1e-12
    assert np.abs(F1 - F1.T).max() < 1e-12


def test_eigen_contract(oracle):
    # src/linalg.rs:300-340
    w, V = oracle.eigen(np.array([[2.0, 1.0], [1.0, 2.0]]))
    assert abs(w[0] - 1) < 1e-12 and abs(w[1] - 3) < 1e-12
    rng = np.random.default_rng(7)
    for n in (3, 17, 80):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, V = oracle.eigen(A)
        assert np.all(np.diff(w) >= 0)
        assert np.abs(V @ np.diag(w) @ V.T - A).max() < 1e-9
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-10
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-10


def test_gradient_matches_finite_difference(oracle):
    # src/gradient.rs:682-705: analytic vs central FD of the SCF energy (< 5e-5 eV/Bohr), and
    # translational invariance (tests/api_surface.rs:467-472)
    z, x = atoms_to_zx([["C", 0, 0, 0], ["O", 0.05, 0, 1.21], ["H", 0.94, 0.1, -0.54], ["H", -0.94, 0, -0.5]])
    opts = oracle.options(p_tol=1e-11, e_tol=1e-12)
    st, g = oracle.gradient(z, x, 0.0, opts)
    assert st == 0
    G = g["gradient"]
    assert np.abs(G.sum(axis=0)).max() < 1e-8
    h = 1e-4
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            ep = oracle.run(z, xp, 0.0, opts, want_arrays=False)[1]["total_ev"]
            em = oracle.run(z, xm, 0.0, opts, want_arrays=False)[1]["total_ev"]
            assert abs((ep - em) / (2 * h) - G[a, k]) < 5e-5, (a, k)


def test_threaded_gradient_batch_equals_single_runs(oracle):
    # the CPU-baseline harness for SCF + gradient: molecules of different sizes packed CSR and
    # solved over threads give each molecule's own run bit for bit, in the molecule's atom order
    mols = [atoms_to_zx([["O", 0, 0, 0], ["H", 0.9584, 0, 0], ["H", -0.24, 0.9278, 0]]),
            atoms_to_zx([["C", 0, 0, 0], ["O", 0.05, 0, 1.21], ["H", 0.94, 0.1, -0.54], ["H", -0.94, 0, -0.5]]),
            atoms_to_zx([["H", 0, 0, 0], ["H", 0.74, 0, 0]]),
            atoms_to_zx([["C", 0, 0, 0], ["H", 0.63, 0.63, 0.63], ["H", -0.63, -0.63, 0.63], ["H", -0.63, 0.63, -0.6],
                         ["H", 0.63, -0.63, -0.63]])]
    opts = oracle.options(p_tol=1e-10, e_tol=1e-12)
    tot, hof, it, st, g = oracle.gradient_batch(mols, opts, nthreads=3)
    assert np.all(st == 0)
    for k, (z, x) in enumerate(mols):
        s1, o = oracle.gradient(z, x, 0.0, opts)
        assert s1 == 0 and g[k].shape == (len(z), 3)
        assert hof[k] == o["hof_kcal"] and tot[k] == o["total_ev"] and it[k] == o["iterations"]
        assert np.array_equal(g[k], o["gradient"]), k


def test_optimizer_reproduces_the_reference_water_run(oracle):
    # src/optimizer.rs:362-386: the distorted water reaches dHf = -53.4330121104622 kcal/mol
    # (the reference asserts 5e-3; the restated L-BFGS follows the same trajectory)
    A = oracle.ANGSTROM_TO_BOHR
    z = np.array([8, 1, 1], np.int32)
    x = np.array([[0.0, 0.0, 0.0], [1.05, 0.0, 0.0], [-0.30, 1.02, 0.0]]) * A
    r = oracle.optimize(z, x)
    assert r["converged"] and r["max_gradient"] < 1e-3
    assert abs(r["hof_kcal"] - (-53.4330121104622)) < 1e-5
    assert r["trajectory"][-1] < r["trajectory"][0]
    assert all(b <= a for a, b in zip(r["trajectory"], r["trajectory"][1:]))  # Armijo: never uphill
    # every OptStep holds the energy, heat and gradient of its own positions (a step that paired
    # a point with its predecessor's scalars fails here); the first is the start, the last the end
    steps = r["steps"]
    assert len(steps) == len(r["trajectory"]) and [s["energy_ev"] for s in steps] == r["trajectory"]
    assert np.array_equal(steps[0]["xyz"], x) and np.array_equal(steps[-1]["xyz"], r["xyz"])
    for s in (steps[0], steps[len(steps) // 2], steps[-1]):
        st, g = oracle.gradient(z, s["xyz"])
        assert st == 0 and g["total_ev"] == s["energy_ev"] and g["hof_kcal"] == s["hof_kcal"]
        assert np.abs(g["gradient"]).max() == s["max_gradient"]


def test_optimizer_on_an_open_shell(oracle):
    # a distorted ethyl radical on the UHF surface: the Armijo condition holds at every step and
    # the end point is stationary by a central difference of the energy along each coordinate
    A = oracle.ANGSTROM_TO_BOHR
    z = np.array([6, 6, 1, 1, 1, 1, 1], np.int32)
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [-0.4, 1.0, 0.1], [-0.4, -0.5, 0.9], [-0.35, -0.5, -0.85],
                  [2.0, 0.95, 0.1], [2.05, -0.9, -0.1]]) * A
    opts = oracle.options(p_tol=1e-10, e_tol=1e-12)
    r = oracle.optimize(z, x, 0.0, opts, multiplicity=2)
    assert r["converged"] and r["max_gradient"] < 1e-3
    assert all(b <= a for a, b in zip(r["trajectory"], r["trajectory"][1:]))
    h = 1e-4
    xe = r["xyz"].copy()
    for i in range(len(z)):
        for c in range(3):
            xp, xm = xe.copy(), xe.copy()
            xp[i, c] += h
            xm[i, c] -= h
            ep = oracle.run_pm3(z, xp, 0.0, 2, opts)[1]["total_ev"]
            em = oracle.run_pm3(z, xm, 0.0, 2, opts)[1]["total_ev"]
            assert abs((ep - em) / (2 * h)) < 2e-3


def test_optimizer_finds_the_minimum_of_a_bond_scan(oracle):
    # H2 from 0.9 A: the optimised bond length against a golden-section minimisation of the
    # energy along the bond; stopping at max |g| < gtol leaves |d - d_min| <= gtol / k
    opts = oracle.options(p_tol=1e-11, e_tol=1e-12)
    z = np.array([1, 1], np.int32)

    def energy(d):
        return oracle.run(z, np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]]), 0.0, opts, want_arrays=False)[1]["total_ev"]

    a, b = 1.0, 2.0
    gr = (5 ** 0.5 - 1) / 2
    for _ in range(80):
        c, d = b - gr * (b - a), a + gr * (b - a)
        if energy(c) < energy(d):
            b = d
        else:
            a = c
    d_min = 0.5 * (a + b)
    e_min = energy(d_min)
    h = 1e-3
    k = (energy(d_min + h) - 2 * e_min + energy(d_min - h)) / h ** 2
    r = oracle.optimize(z, np.array([[0.0, 0.0, 0.0], [0.9 * oracle.ANGSTROM_TO_BOHR, 0.1, 0.0]]), 0.0, opts)
    assert r["converged"]
    d_opt = float(np.linalg.norm(r["xyz"][1] - r["xyz"][0]))
    assert abs(d_opt - d_min) <= 1.5e-3 / k
    assert -1e-9 <= r["energy_ev"] - e_min <= 1e-6


def test_accelerators_agree(oracle):
    # src/scf.rs:1938-1963
    z, x = atoms_to_zx([["C", 0, 0, 0], ["O", 0, 0, 1.21], ["H", 0.94, 0, -0.54], ["H", -0.94, 0, -0.54]])
    runs = {a: oracle.run(z, x, 0.0, oracle.options(accelerator=a, max_scf=500))[1] for a in (0, 1, 2)}
    assert abs(runs[2]["total_ev"] - runs[1]["total_ev"]) < 1e-6
    assert abs(runs[2]["total_ev"] - runs[0]["total_ev"]) < 1e-6
    assert runs[2]["iterations"] < runs[0]["iterations"]


def test_depth_budget_does_not_change_the_answer(oracle):
    # src/scf.rs:1776-1801: scf_memory_mb 1 forces depth-2 CDIIS with the same energy
    z, x = atoms_to_zx([["C", 0, 0, 0], ["O", 0, 0, 1.21], ["H", 0.94, 0, -0.54], ["H", -0.94, 0, -0.54]])
    a = oracle.run(z, x, 0.0, oracle.options(p_tol=1e-10, e_tol=1e-12))[1]
    b = oracle.run(z, x, 0.0, oracle.options(p_tol=1e-10, e_tol=1e-12, scf_memory_mb=0))[1]
    assert abs(a["total_ev"] - b["total_ev"]) < 1e-8


def test_scf_not_converged_is_an_error(oracle):
    z, x = atoms_to_zx([["O", 0, 0, 0], ["H", 0.9584, 0, 0], ["H", -0.24, 0.9278, 0]])
    st, r = oracle.run(z, x, 0.0, oracle.options(max_scf=1))
    assert st == 3 and not r["converged"]


WATER = [["O", 0, 0, 0.1173], ["H", 0, 0.7572, -0.4692], ["H", 0, -0.7572, -0.4692]]
METHYL = [["C", 0, 0, 0], ["H", 1.079, 0, 0], ["H", -0.5395, 0.9344, 0], ["H", -0.5395, -0.9344, 0]]


def test_reference_option_selects_the_path(oracle):
    # tests/api_surface.rs:383-410 and src/scf.rs:508-520: Auto picks by shell, Rhf refuses an open
    # shell, Uhf forced on a singlet gives the restricted energy and an unrestricted result
    zc, xc = atoms_to_zx(WATER)
    zr, xr = atoms_to_zx(METHYL)
    st, auto = oracle.run_pm3(zc, xc, 0.0, 1, oracle.options())
    assert st == 0 and not auto["unrestricted"]
    st, rad = oracle.run_pm3(zr, xr, 0.0, 2, oracle.options())
    assert st == 0 and rad["unrestricted"] and abs(np.trace(rad["spin_density"]) - 1.0) < 1e-6
    assert oracle.run_pm3(zr, xr, 0.0, 2, oracle.options(reference="rhf"))[0] == 1
    st, rhf = oracle.run_pm3(zc, xc, 0.0, 1, oracle.options(reference="rhf"))
    assert st == 0 and not rhf["unrestricted"] and rhf["total_ev"] == auto["total_ev"]
    st, forced = oracle.run_pm3(zc, xc, 0.0, 1, oracle.options(reference="uhf"))
    assert st == 0 and forced["unrestricted"] and abs(forced["total_ev"] - auto["total_ev"]) < 1e-6
    assert np.abs(forced["spin_density"]).max() < 1e-8


def test_steering_switches_the_search_off_and_keeps_the_fixed_point(pins, oracle):
    # src/scf.rs:563-607: a level shift or damping steers the first path and switches the stability
    # search off (ketene then stays in its excited solution, paths = 1).  rhf_loop reads neither,
    # so a restricted steered solve is the search-off solve bit for bit; uhf_loop reads both, and a
    # shifted, damped unrestricted path reaches the same radical fixed point (level shift and
    # damping move the path, not the fixed points: src/scf.rs:193-202)
    row = next(r for r in pins["mopac_closed_shell"] if r["name"] == "ketene")
    z, x = atoms_to_zx(row["atoms"])
    st, steered = oracle.run_pm3(z, x, 0.0, 1, oracle.options(level_shift_ev=0.5, damping=0.2))
    _, off = oracle.run_pm3(z, x, 0.0, 1, oracle.options(stability=0))
    assert st == 0 and steered["scf_paths_tried"] == 1 and steered["total_ev"] == off["total_ev"]
    zr, xr = atoms_to_zx(METHYL)
    st, plain = oracle.run_pm3(zr, xr, 0.0, 2, oracle.options(p_tol=1e-10, e_tol=1e-12, stability=0))
    st2, shifted = oracle.run_pm3(zr, xr, 0.0, 2, oracle.options(p_tol=1e-10, e_tol=1e-12, level_shift_ev=1.5,
                                                                damping=0.3, max_scf=500))
    assert st == st2 == 0 and shifted["scf_paths_tried"] == 1
    assert abs(shifted["total_ev"] - plain["total_ev"]) < 1e-8
    assert np.abs(shifted["density"] - plain["density"]).max() < 1e-7
    assert plain["iterations"] != shifted["iterations"]


def test_field_energy_derivative_is_minus_the_dipole(oracle):
    # mu = -dE/df (src/dipole.rs:1-40): central differences of the total energy in a small uniform
    # field against the field-free dipole (a neutral molecule, so the origin does not matter); a
    # sign or a dropped s-p term in the field operator fails this
    z, x = atoms_to_zx(WATER)
    opts = dict(p_tol=1e-11, e_tol=1e-13, max_scf=500)
    st, r0 = oracle.run_pm3(z, x, 0.0, 1, oracle.options(**opts))
    mu = r0["dipole"] / 2.541746473  # e Bohr
    h = 1e-4
    for ax in range(3):
        f = np.zeros(3)
        f[ax] = h
        _, rp = oracle.run_pm3(z, x, 0.0, 1, oracle.options(field=f, **opts))
        _, rm = oracle.run_pm3(z, x, 0.0, 1, oracle.options(field=-f, **opts))
        assert abs(-(rp["total_ev"] - rm["total_ev"]) / (2 * h) - mu[ax]) < 1e-6, ax
    assert abs(mu).max() > 0.3


def test_gradient_in_a_field_matches_finite_differences(oracle):
    # src/gradient.rs:760-810: the analytic gradient in a field (the -q_A f term included) against a
    # central difference of the energy in the same field
    z, x = atoms_to_zx(WATER)
    o = oracle.options(p_tol=1e-11, e_tol=1e-13, max_scf=500, field=(0.02, -0.013, 0.031))
    st, g = oracle.gradient(z, x, 0.0, o)
    assert st == 0
    h = 1e-4
    for a in range(3):
        for ax in range(3):
            xp = x.copy()
            xp[a, ax] += h
            xm = x.copy()
            xm[a, ax] -= h
            ep = oracle.run_pm3(z, xp, 0.0, 1, o)[1]["total_ev"]
            em = oracle.run_pm3(z, xm, 0.0, 1, o)[1]["total_ev"]
            assert abs((ep - em) / (2 * h) - g["gradient"][a, ax]) < 5e-6, (a, ax)
    _, g0 = oracle.gradient(z, x, 0.0, oracle.options(p_tol=1e-11, e_tol=1e-13, max_scf=500))
    assert np.abs(g["gradient"] - g0["gradient"]).max() > 1e-3


def test_unrestricted_gradient_matches_finite_differences(oracle):
    # fixed_density_gradient_uhf (src/gradient.rs:494-599) on the methyl radical against a central
    # difference of the unrestricted energy; translational invariance
    z, x = atoms_to_zx(METHYL)
    x = x + np.array([[0.0, 0.0, 0.05], [0.02, 0.0, -0.03], [0.0, 0.04, 0.0], [0.0, 0.0, 0.0]])
    o = oracle.options(p_tol=1e-11, e_tol=1e-13, max_scf=500, stability=0)
    st, g = oracle.gradient(z, x, 0.0, o, multiplicity=2)
    assert st == 0 and g["unrestricted"]
    h = 1e-4
    for a in range(4):
        for ax in range(3):
            xp = x.copy()
            xp[a, ax] += h
            xm = x.copy()
            xm[a, ax] -= h
            ep = oracle.run_pm3(z, xp, 0.0, 2, o)[1]["total_ev"]
            em = oracle.run_pm3(z, xm, 0.0, 2, o)[1]["total_ev"]
            assert abs((ep - em) / (2 * h) - g["gradient"][a, ax]) < 5e-6, (a, ax)
    assert np.abs(g["gradient"].sum(axis=0)).max() < 1e-8


def test_pair_cache_ceiling(oracle):
    # src/scf.rs:1805-1840 and src/hamiltonian.rs:51-71,202-211: the exact cache size against a
    # ceiling; a 1 MiB budget rejects a cluster whose cache is above it, before any work
    atoms = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                xx, yy, zz = 3.0 * i, 3.0 * j, 3.0 * k
                atoms += [["O", xx, yy, zz], ["H", xx + 0.96, yy, zz], ["H", xx - 0.24, yy + 0.93, zz]]
    z, x = atoms_to_zx(atoms)
    p = np.array([10 if v == 8 else 1 for v in z], dtype=float)
    need = ((p.sum() ** 2 - (p * p).sum()) / 2) * 8 + len(z) * (len(z) - 1) / 2 * 72
    st, r = oracle.run_pm3(z, x, 0.0, 1, oracle.options(integral_memory_mb=1, max_scf=1))
    assert st == 5 and r["limit_mb"] == 1 and r["required_mb"] == np.ceil(need / 2**20) and need > 2 * 2**20
    st, r = oracle.run_pm3(z, x, 0.0, 1, oracle.options(integral_memory_mb=2 + int(need // 2**20), max_scf=1))
    assert st in (0, 3)
