"""Densities by purification (the batched SCF's route when no orbitals are requested) against
the eigenvector route: the purified projector and the Lanczos frontier levels of the device
against numpy's eigh of the same Fock matrices (symmetric molecules with degenerate frontier
orbitals, n_occ = n, a closed gap), then whole SCF runs through the purified route against the
oracle and against the engine's own eigensolver route."""
import ctypes

import numpy as np
import pytest

from conftest import atoms_to_zx
from pm3amd import workloads as W

pytestmark = pytest.mark.gpu

HOF_TOL = 1e-5
TIGHT = dict(p_tol=1e-10, e_tol=1e-12)


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def _stage(G):
    L = G.native()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    L.pm3_stage_purify.argtypes = [ctypes.POINTER(G._CMolecule), ctypes.c_void_p, dp, dp, ip, dp, dp]
    prm = G.Pm3Parameters.standard()

    def run(m, F):
        n = F.shape[0]
        F = np.ascontiguousarray(F)
        P = np.zeros((n, n))
        k = ctypes.c_int32(0)
        h, l = ctypes.c_double(0.0), ctypes.c_double(0.0)
        st = L.pm3_stage_purify(ctypes.byref(m._c()), prm._h, F.ctypes.data_as(dp), P.ctypes.data_as(dp),
                                ctypes.byref(k), ctypes.byref(h), ctypes.byref(l))
        assert st == 0
        return P, k.value, h.value, l.value
    return run


def _cases(pins):
    names = ["water", "methane", "benzene", "ethane", "lithium_fluoride", "chloroform", "hydroxide",
             "silicon_tetrafluoride", "ammonium", "ketene", "tetrafluoroethene"]
    rows = [r for r in pins["mopac_closed_shell"] if r["name"] in names]
    out = [(r["name"], *atoms_to_zx(r["atoms"]), r["charge"]) for r in rows]
    for name, (z, x) in (("C24H50", W.alkane(24)), ("water16", W.water_cluster(16, 7.0, 5)),
                         ("conformer", W.solvated_conformers(1, 9)[0])):
        out.append((name, np.asarray(z, np.int32), W.to_bohr(x), 0.0))
    return out


def test_purified_density_and_frontier_levels(G, pins, oracle):
    run = _stage(G)
    for name, z, x, q in _cases(pins):
        st, o = oracle.run(z, x, q, oracle.options(**TIGHT))
        assert st == 0
        F = oracle.fock(z, x, o["density"])
        nocc = o["n_occ"]
        w, V = np.linalg.eigh(F)
        Pex = 2.0 * V[:, :nocc] @ V[:, :nocc].T
        P, k, homo, lumo = run(G.Molecule(z, x, q), F)
        print(f"case {name} n {len(w)} nocc {nocc} products {k} |dP| {np.abs(P - Pex).max():.1e} "
              f"dHOMO {homo - w[nocc - 1]:.1e} dLUMO {lumo - w[nocc]:.1e}", flush=True)
        assert 0 < k <= 60
        assert np.abs(P - Pex).max() <= 1e-11, name
        assert np.abs(P - P.T).max() == 0.0
        assert abs(homo - w[nocc - 1]) <= 1e-9 and abs(lumo - w[nocc]) <= 1e-9, name


def test_purification_edge_cases(G, oracle):
    run = _stage(G)
    # n_occ = n (fluoride): P = 2 I, the HOMO is the top of the spectrum and there is no LUMO
    z, x = atoms_to_zx([["F", 0, 0, 0]])
    F = np.diag([-30.0, -12.0, -11.5, -11.0])
    F[0, 1] = F[1, 0] = 0.3
    P, k, homo, lumo = run(G.Molecule(z, x, -1.0), F)
    assert k == 0 and np.abs(P - 2.0 * np.eye(4)).max() == 0.0
    assert abs(homo - np.linalg.eigvalsh(F)[-1]) <= 1e-10 and lumo == 0.0
    # a closed gap (HOMO = LUMO) is reported, so the SCF diagonalises instead
    z, x = atoms_to_zx([["O", 0, 0, 0], ["H", 0.9584, 0, 0], ["H", -0.24, 0.9278, 0]])
    st, o = oracle.run(z, x)
    w, V = np.linalg.eigh(oracle.fock(z, x, o["density"]))
    nocc = o["n_occ"]
    w[nocc] = w[nocc - 1]
    P, k, _, _ = run(G.Molecule(z, x), (V * w) @ V.T)
    assert k == -1
    # an open but small gap still converges, to the right projector
    w[nocc] = w[nocc - 1] + 0.05
    Fs = (V * w) @ V.T
    P, k, homo, lumo = run(G.Molecule(z, x), Fs)
    Pex = 2.0 * V[:, :nocc] @ V[:, :nocc].T
    print("small gap products", k, np.abs(P - Pex).max(), flush=True)
    assert k > 0 and np.abs(P - Pex).max() <= 1e-10
    assert abs(homo - w[nocc - 1]) <= 1e-9 and abs(lumo - w[nocc]) <= 1e-9


def test_purified_scf_matches_oracle(G, pins, oracle):
    """Batches take the purified route: the answers, the frontier levels and the stability rungs
    are the oracle's (every density is the exact aufbau projector of its Fock matrix)."""
    cases = _cases(pins)
    mols = [G.Molecule(z, x, q) for _, z, x, q in cases]
    b = G.Batch(mols, options=G.Pm3Options(max_scf=2000, **TIGHT), want_density=True)
    b.solve()
    status, res = b.results()
    for (name, z, x, q), s, r in zip(cases, status, res):
        st, o = oracle.run(z, x, q, oracle.options(max_scf=2000, **TIGHT))
        print("case", name, r.iterations, o["iterations"], r.scf_paths_tried, o["scf_paths_tried"], flush=True)
        assert s == 0 and st == 0
        assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= HOF_TOL, name
        assert np.abs(r.density - o["density"]).max() <= 1e-8, name
        assert abs(r.homo_ev - o["homo_ev"]) <= 1e-7 and abs(r.lumo_ev - o["lumo_ev"]) <= 1e-7, name
        assert np.abs(r.charges - o["charges"]).max() <= 1e-8
        assert np.abs(r.dipole_debye - o["dipole"]).max() <= 1e-7
        assert r.scf_paths_tried == o["scf_paths_tried"], name
        assert r.mo_energies is None or len(r.mo_energies) == 0


def test_purified_and_eigensolver_routes_agree(G, monkeypatch):
    """The same batch through both routes (the eigensolver route at full precision): same
    iteration counts at tolerances where rounding does not decide the stopping iteration."""
    mols = [G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(96, 515)]
    opts = G.Pm3Options(p_tol=1e-8, e_tol=1e-8)
    b = G.Batch(mols, options=opts, want_density=True)
    b.solve()
    s1, r1 = b.results()
    b.close()
    monkeypatch.setenv("PM3_PURIFY", "0")
    monkeypatch.setenv("PM3_EIG_TOL_FACTOR", "0")
    b = G.Batch(mols, options=opts, want_density=True)
    b.solve()
    s2, r2 = b.results()
    b.close()
    assert np.all(s1 == 0) and np.all(s2 == 0)
    same = 0
    for a, c in zip(r1, r2):
        assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= HOF_TOL
        assert np.abs(a.density - c.density).max() <= 1e-7
        assert abs(a.homo_ev - c.homo_ev) <= 1e-6 and abs(a.lumo_ev - c.lumo_ev) <= 1e-6
        same += a.iterations == c.iterations
    print("equal iteration counts", same, "of", len(r1), flush=True)
    assert same >= 0.95 * len(r1)


def test_missed_spectral_bounds_are_caught(G, pins, oracle, monkeypatch):
    """Spectral bounds deliberately set inside the spectrum (negative margin): the checks redo each
    purification with Gershgorin bounds, and a path whose final HOMO lies above its LUMO is run
    again, so densities and answers are those of the default settings."""
    monkeypatch.setenv("PM3_SP2_MARGIN", "-0.3")
    monkeypatch.setenv("PM3_SP2_TIGHT_N", "8")
    run = _stage(G)
    for name, z, x, q in _cases(pins)[-4:]:
        st, o = oracle.run(z, x, q, oracle.options(**TIGHT))
        F = oracle.fock(z, x, o["density"])
        nocc = o["n_occ"]
        w, V = np.linalg.eigh(F)
        P, k, homo, lumo = run(G.Molecule(z, x, q), F)
        print("case", name, "products", k, flush=True)
        assert k > 0 and np.abs(P - 2.0 * V[:, :nocc] @ V[:, :nocc].T).max() <= 1e-11
        assert abs(homo - w[nocc - 1]) <= 1e-9 and abs(lumo - w[nocc]) <= 1e-9
    mols = [G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(48, 616)]
    opts = G.Pm3Options(p_tol=1e-8, e_tol=1e-8)
    b = G.Batch(mols, options=opts, want_density=True)
    b.solve()
    s1, r1 = b.results()
    b.close()
    monkeypatch.delenv("PM3_SP2_MARGIN")
    monkeypatch.delenv("PM3_SP2_TIGHT_N")
    b = G.Batch(mols, options=opts, want_density=True)
    b.solve()
    s2, r2 = b.results()
    b.close()
    assert np.all(s1 == 0) and np.all(s2 == 0)
    for a, c in zip(r1, r2):
        assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= HOF_TOL
        assert np.abs(a.density - c.density).max() <= 1e-7
        assert abs(a.homo_ev - c.homo_ev) <= 1e-6 and abs(a.lumo_ev - c.lumo_ev) <= 1e-6
