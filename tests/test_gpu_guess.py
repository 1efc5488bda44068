"""The learned initial density on the GPU against the CPU restatement (oracle/guess_oracle.cpp):
features bit for bit, the bf16 network's outputs and parameter gradients, the assembled density
and the commutator test of the fallback, the training loss and its full gradient through the
training kernels; then the SCF: the learned start leaves converged answers unchanged and the
trained network cuts SCF iterations."""

import os

import numpy as np
import pytest

from pm3amd import workloads as W

pytestmark = pytest.mark.gpu

HOF_TOL = 1e-5
NF, NO = 96, 16


@pytest.fixture(scope="module")
def G():
    import pm3amd
    import pm3amd.guess
    pm3amd.native()
    return pm3amd


def _mols(count, seed):
    return [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(count, seed)]


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, 1 / np.sqrt(93), (96, 128)) * scale
    w1[93:95] = 0.0
    w2 = rng.normal(0, 1 / np.sqrt(128), (128, 128)) * scale
    b2 = rng.normal(0, 0.1, 128)
    w3 = rng.normal(0, 0.3 / np.sqrt(128), (128, 16)) * scale
    b3 = rng.normal(0, 0.05, 16)
    return np.concatenate([w1.ravel(), w2.ravel(), b2, w3.ravel(), b3]).astype(np.float32)


def _sorted_pairs(src, dst):
    return np.lexsort((dst, src))


def test_features_match_oracle(G, oracle):
    cases = [W.water_cluster(4, 7.0, 3), W.alkane(6)] + W.mixed_batch(6, 21)[:4]
    for z, x in cases:
        z = np.asarray(z, np.int32)
        xb = W.to_bohr(x)
        print("case nat", len(z), flush=True)
        src, dst, xg = G.guess.features(G.Molecule(z, xb))
        so, do = oracle.guess_pairs(z, xb)
        _, xo = oracle.guess_features(z, xb)
        ig, io = _sorted_pairs(src, dst), _sorted_pairs(so, do)
        np.testing.assert_array_equal(src[ig], so[io])
        np.testing.assert_array_equal(dst[ig], do[io])
        np.testing.assert_array_equal(xg[ig], xo[io])


def test_mlp_forward_backward_match_oracle(G, oracle):
    xs = []
    for z, xb in _mols(6, 33):
        xs.append(oracle.guess_features(z, xb)[1])
    x = np.concatenate(xs)[:1000]
    p = _params(1)
    m = G.guess.GuessModel.create(7)
    m.set_params(p)
    rng = np.random.default_rng(2)
    dy = rng.normal(size=(len(x), NO)).astype(np.float32)
    y = m.mlp(x)
    y2, g = m.mlp(x, dy)
    np.testing.assert_array_equal(y, y2)
    yo = oracle.guess_forward(p, x=x, rounded=True)
    go = oracle.guess_backward(p, dy, x=x, rounded=True)
    assert np.abs(y - yo).max() <= 2e-3 * np.abs(yo).max()
    assert np.linalg.norm(y - yo) <= 1e-3 * np.linalg.norm(yo)
    for name, a, b in (("W1", 0, 96 * 128), ("W2", 96 * 128, 96 * 128 + 128 * 128),
                       ("b2", 96 * 128 + 128 * 128, 96 * 128 + 128 * 128 + 128),
                       ("W3", 96 * 128 + 128 * 128 + 128, len(go) - 16), ("b3", len(go) - 16, len(go))):
        rel = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        print(name, rel, flush=True)
        assert rel <= 5e-3, name
    # the padding inputs 93, 94 never receive gradient
    assert np.all(g[93 * 128:95 * 128] == 0.0)


def test_learned_density_and_commutators(G, oracle):
    p = _params(3, 0.7)
    m = G.guess.GuessModel.create(1)
    m.set_params(p)
    for z, xb in _mols(6, 44)[:3] + [(np.asarray(W.water_cluster(5, 7.0, 9)[0], np.int32),
                                      W.to_bohr(W.water_cluster(5, 7.0, 9)[1]))]:
        mol = G.Molecule(z, xb)
        print("case nat", len(z), flush=True)
        P, cl, cs = m.density(mol)
        src, dst, xg = G.guess.features(mol)
        yg = m.mlp(xg)
        so, do = oracle.guess_pairs(z, xb)
        order = _sorted_pairs(src, dst)
        yo = np.zeros((len(so), NO), np.float32)
        yo[_sorted_pairs(so, do)] = yg[order]
        Po = oracle.guess_density(z, xb, yo)
        assert np.abs(P - Po).max() <= 1e-12
        assert np.abs(P - P.T).max() == 0.0
        for Pc, c in ((Po, cl), (None, cs)):
            if Pc is None:
                Pc = oracle.guess_density(z, xb, np.zeros_like(yo))
            F = oracle.fock(z, xb, Pc)
            E = F @ Pc - Pc @ F
            ref = float(np.sum(E * E))
            assert abs(c - ref) <= 1e-9 * max(ref, 1e-12), (c, ref)


def _solved(G, mols, opts, density=True):
    b = G.Batch([G.Molecule(z, x) for z, x in mols], options=opts, want_density=density)
    b.solve()
    return b


def test_training_loss_and_gradient_match_oracle(G, oracle):
    mols = _mols(12, 55)
    b = _solved(G, mols, G.Pm3Options(p_tol=1e-9, e_tol=1e-9))
    status, res = b.results()
    assert np.all(status == 0)
    data = G.guess.GuessData()
    assert data.add(b) == len(mols)
    p = _params(4, 0.7)
    m = G.guess.GuessModel.create(2)
    m.set_params(p)
    g, loss = G.guess.gradient(m, data)
    go = np.zeros_like(g, dtype=np.float64)
    lo, atoms = 0.0, 0
    for (z, xb), r in zip(mols, res):
        _, x = oracle.guess_features(z, xb)
        y = oracle.guess_forward(p, x=x, rounded=True).astype(np.float32)
        l, dy = oracle.guess_loss(z, xb, y, r.density)
        lo += l
        atoms += len(z)
        go += oracle.guess_backward(p, dy, x=x, rounded=True)
    lo /= atoms
    go /= atoms
    assert abs(loss - lo) <= 2e-3 * lo, (loss, lo)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("loss", loss, lo, "grad rel", rel, flush=True)
    assert rel <= 1e-2
    assert abs(G.guess.evaluate(m, data) - loss) <= 1e-6 * loss


def test_training_reduces_loss_and_is_reproducible(G):
    mols = _mols(64, 66)
    b = _solved(G, mols, G.Pm3Options(p_tol=1e-8, e_tol=1e-8))
    data = G.guess.GuessData()
    data.add(b)
    st = data.stats()
    assert st["molecules"] == 64 and st["pairs"] > 0
    runs = []
    for _ in range(2):
        m = G.guess.GuessModel.create(5)
        l0 = G.guess.evaluate(m, data)
        losses, sec = G.guess.train(m, data, steps=120, batch_molecules=16, lr=1e-3, seed=3)
        l1 = G.guess.evaluate(m, data)
        runs.append((l0, l1, losses, m.params()))
        assert m.steps == 120
    l0, l1, losses, p = runs[0]
    print("loss", l0, "->", l1, flush=True)
    assert np.all(np.isfinite(losses))
    assert l1 < 0.5 * l0
    np.testing.assert_array_equal(runs[0][3], runs[1][3])
    np.testing.assert_array_equal(runs[0][2], runs[1][2])


def _model(G):
    if os.path.exists(G.guess.DEFAULT_MODEL):
        return G.guess.GuessModel.default(), True
    m = G.guess.GuessModel.create(11)
    b = _solved(G, _mols(512, 77), G.Pm3Options(p_tol=1e-8, e_tol=1e-8))
    data = G.guess.GuessData()
    data.add(b)
    G.guess.train(m, data, steps=400, batch_molecules=128, seed=1)
    return m, False


def test_learned_scf_keeps_converged_answers(G, oracle, exact_eigensolves):
    model, _ = _model(G)
    mols = _mols(48, 88)
    tight = dict(p_tol=1e-10, e_tol=1e-12, max_scf=500)
    bs = _solved(G, mols, G.Pm3Options(**tight))
    bl = _solved(G, mols, G.Pm3Options(guess="learned", guess_model=model, **tight))
    ss, rs = bs.results()
    sl, rl = bl.results()
    assert np.all(ss == 0) and np.all(sl == 0)
    for a, c in zip(rs, rl):
        assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= HOF_TOL
        assert np.abs(a.density - c.density).max() <= 1e-8
        assert c.guess_used in (1, 2)
    # the oracle started from the device's learned density follows the same SCF (tolerances at
    # which the stopping iteration is not decided by rounding)
    loose = dict(p_tol=1e-8, e_tol=1e-8)
    for z, xb in mols[:6]:
        mol = G.Molecule(z, xb)
        r = G.run_pm3(mol, options=G.Pm3Options(guess="learned", guess_model=model, **loose))
        if r.guess_used != 1:
            continue
        P0, cl, cs = model.density(mol)
        assert cl <= cs
        st, o = oracle.run_from_density(z, xb, P0, 0.0, oracle.options(**loose))
        assert st == 0 and abs(o["hof_kcal"] - r.heat_of_formation_kcal) <= HOF_TOL
        print("iterations", r.iterations, o["iterations"], flush=True)
        assert abs(o["iterations"] - r.iterations) <= 1


def test_trained_guess_cuts_iterations(G):
    model, trained = _model(G)
    mols = _mols(256, 4242)
    opts = dict(p_tol=1e-8, e_tol=1e-8)
    bs = _solved(G, mols, G.Pm3Options(**opts), density=False)
    bl = _solved(G, mols, G.Pm3Options(guess="learned", guess_model=model, **opts), density=False)
    ss, rs = bs.results()
    sl, rl = bl.results()
    assert np.all(ss == 0) and np.all(sl == 0)
    dh = max(abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) for a, c in zip(rs, rl))
    it_s = np.mean([r.iterations for r in rs])
    it_l = np.mean([r.iterations for r in rl])
    used = np.mean([r.guess_used == 1 for r in rl])
    print(f"iterations SAD {it_s:.2f} learned {it_l:.2f}; learned kept {used:.3f}; max|dHf| {dh:.2e}", flush=True)
    assert dh <= HOF_TOL
    assert it_l <= (0.88 if trained else 0.97) * it_s


def test_conformer_batch_with_learned_guess_and_gradients(G, oracle):
    """The C4 configuration at reduced count: solvated C12H26 conformers (98 atoms, 194 AOs) in one
    batch with the packaged learned guess and analytic gradients, at the gradient bench's
    tolerances (p_tol 1e-10, e_tol 1e-12), against the oracle's SCF + gradient of every molecule:
    heats of formation within 1e-5 kcal/mol and gradients within 2.295e-7 eV/Bohr."""
    import pm3amd.guess as GS
    from pm3amd import workloads as W
    model = GS.GuessModel.load(GS.DEFAULT_MODEL)
    mols = [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.solvated_conformers(12, 2027)]
    tight = dict(p_tol=1e-10, e_tol=1e-12)
    b = G.Batch([G.Molecule(z, x) for z, x in mols], options=G.Pm3Options(guess="learned", guess_model=model, **tight),
                want_gradient=True)
    b.solve()
    st, res = b.results()
    grads = b.gradients()
    b.close()
    tot, hof, it, ost, og = oracle.gradient_batch(mols, oracle.options(**tight), nthreads=8)
    assert np.all(st == 0) and np.all(ost == 0)
    dh = max(abs(r.heat_of_formation_kcal - h) for r, h in zip(res, hof))
    dg = max(float(np.abs(g - o).max()) for g, o in zip(grads, og))
    kept = np.mean([r.guess_used == 1 for r in res])
    print(f"conformers: max|dHf| {dh:.2e} max|dg| {dg:.2e} learned kept {kept:.2f} "
          f"iterations {np.mean([r.iterations for r in res]):.2f} (oracle {np.mean(it):.2f})", flush=True)
    assert dh <= HOF_TOL and dg <= 2.295e-7
    assert kept > 0.5  # the learned density is kept for most conformers (H, C and O only)


def test_moved_learned_batch_matches_a_fresh_one(G):
    """pm3_batch_set_positions on a learned-guess batch rebuilds its 4 Angstrom neighbour lists:
    after compressing the molecules (more pairs: the device lists and network buffers regrow),
    stretching them and moving them back, every solve equals that of a learned batch created at
    the same positions bit for bit, with the same directed-pair count."""
    model, _ = _model(G)
    base = _mols(24, 606)
    opts = G.Pm3Options(guess="learned", guess_model=model, p_tol=1e-10, e_tol=1e-12)

    def scaled(f):
        return [x.mean(axis=0) + f * (x - x.mean(axis=0)) for _, x in base]

    b = G.Batch([G.Molecule(z, x) for z, x in base], options=opts, want_density=True, want_gradient=True)
    b.solve()
    counts = [b.guess_timing()[2]]
    for f in (0.9, 1.12, 1.0):
        moved = scaled(f)
        b.set_positions(moved)
        b.solve()
        s1, r1 = b.results()
        g1 = b.gradients()
        n1 = b.guess_timing()[2]
        fresh = G.Batch([G.Molecule(z, x) for (z, _), x in zip(base, moved)], options=opts, want_density=True,
                        want_gradient=True)
        fresh.solve()
        s2, r2 = fresh.results()
        g2 = fresh.gradients()
        n2 = fresh.guess_timing()[2]
        fresh.close()
        print(f"scale {f}: directed pairs {n1} (fresh {n2}), learned kept "
              f"{np.mean([r.guess_used == 1 for r in r1]):.2f}, statuses {np.bincount(s1)}", flush=True)
        counts.append(n1)
        assert n1 == n2 and np.array_equal(s1, s2)
        for a, c, ga, gc in zip(r1, r2, g1, g2):
            assert a.heat_of_formation_kcal == c.heat_of_formation_kcal
            assert a.iterations == c.iterations and a.guess_used == c.guess_used
            assert np.array_equal(a.density, c.density) and np.array_equal(ga, gc)
    b.close()
    assert counts[1] > counts[0] > counts[2] and counts[3] == counts[0]


def test_learned_guess_optimisations_reach_the_plain_minima(G):
    """pm3_optimize_batch with the learned guess (each round's batch moves its atoms and rebuilds
    the guess's neighbour lists) reaches the minima of the superposition-guess optimiser."""
    model, _ = _model(G)
    rng = np.random.default_rng(55)
    mols = [G.Molecule(np.asarray(z, np.int32), W.to_bohr(W.jitter(np.asarray(x), 0.06, rng)))
            for z, x in (W.alkane(n) for n in (2, 3, 4, 5, 6))]
    tight = dict(p_tol=1e-10, e_tol=1e-12)
    st_p, rp = G.optimize_batch(mols, options=G.Pm3Options(**tight), trajectory=False)
    st_l, rl = G.optimize_batch(mols, options=G.Pm3Options(guess="learned", guess_model=model, **tight),
                                trajectory=False)
    assert np.all(st_p == 0) and np.all(st_l == 0)
    for a, c in zip(rp, rl):
        print(f"plain {a.heat_of_formation_kcal:.6f} ({a.iterations} it) learned "
              f"{c.heat_of_formation_kcal:.6f} ({c.iterations} it)", flush=True)
        assert a.converged and c.converged and c.max_gradient < 1e-3
        assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= 1e-4
