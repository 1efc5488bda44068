"""Tuning entries change speed only: the 16 x 16 wave tiles of the purification products give the
same products bit for bit as the 32 x 32 tiles, the in-place LDS placement the same densities bit
for bit as the two-image and HBM-product placements, whole batches agree with the default table under
every entry's extreme values, and the autotuner returns a table that keeps every answer while
leaving it as the process-wide table."""
import numpy as np
import pytest

from pm3amd import workloads as W

pytestmark = pytest.mark.gpu

HOF_TOL = 1e-5


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    start = pm3amd.tuning_get()
    yield pm3amd
    pm3amd.tuning_set(start)


def _mols(G, count, seed):
    return [G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(count, seed)]


def _solve(G, mols, table, **kw):
    G.tuning_set(table)
    b = G.Batch(mols, options=G.Pm3Options(p_tol=1e-8, e_tol=1e-8), want_density=True, **kw)
    b.solve()
    st, res = b.results()
    ms = b.last_kernel_ms()
    b.close()
    return st, res, ms


def test_tile16_purification_products_are_bitwise_equal(G, oracle):
    """One purification per case, 16 x 16 against 32 x 32 wave tiles: each product element is the same
    MFMA chain, so the densities match bit for bit unless a trace sum's rounding flipped a step choice
    (then the projectors still agree to 1e-13)."""
    from test_gpu_purify import _stage
    run = _stage(G)
    d = G.tuning_default()
    exact = 0
    cases = [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(24, 77)]
    for z, x in cases:
        st, o = oracle.run(z, x, 0.0, oracle.options(p_tol=1e-10, e_tol=1e-12))
        F = oracle.fock(z, x, o["density"])
        m = G.Molecule(z, x)
        G.tuning_set(dict(d, sp2_tile16_n=0))
        P32, k32, h32, l32 = run(m, F)
        G.tuning_set(dict(d, sp2_tile16_n=1 << 20))
        P16, k16, h16, l16 = run(m, F)
        # the in-place placement (one LDS image) against the two-image / HBM-product ones
        G.tuning_set(dict(d, sp2_lds=3))
        Pin, kin, hin, lin = run(m, F)
        G.tuning_set(dict(d, sp2_lds=2))
        Pim, kim, him, lim = run(m, F)
        assert kin == kim and np.array_equal(Pin, Pim) and hin == him and lin == lim, F.shape[0]
        print("nao", F.shape[0], "products", k32, k16, "max|dP|", np.abs(P16 - P32).max(), flush=True)
        assert k32 > 0 and k16 > 0
        assert np.abs(P16 - P32).max() <= 1e-13
        assert abs(h16 - h32) <= 1e-12 and abs(l16 - l32) <= 1e-12
        exact += int(np.array_equal(P16, P32))
    G.tuning_set(d)
    assert exact >= len(cases) - 2


def test_partial_lds_rows_give_the_same_bits(G, oracle):
    """Molecules whose X does not fit one LDS image: the placement that keeps X's leading rows in
    LDS gives the HBM placement's densities bit for bit (same MFMA chains, same store order)."""
    from test_gpu_purify import _stage
    run = _stage(G)
    d = G.tuning_default()
    cases = [W.alkane(22), W.alkane(24), W.water_cluster(32, 10.0, 2026), W.solvated_conformers(1, 9)[0]]
    for z, x in cases:
        z = np.asarray(z, np.int32)
        xb = W.to_bohr(x)
        st, o = oracle.run(z, xb, 0.0, oracle.options(p_tol=1e-10, e_tol=1e-12))
        F = oracle.fock(z, xb, o["density"])
        m = G.Molecule(z, xb)
        G.tuning_set(dict(d, sp2_lds=4))
        P4, k4, h4, l4 = run(m, F)
        G.tuning_set(dict(d, sp2_lds=0))
        P0, k0, h0, l0 = run(m, F)
        print("nao", F.shape[0], "products", k4, k0, flush=True)
        assert k4 == k0 > 0 and np.array_equal(P4, P0) and h4 == h0 and l4 == l0, F.shape[0]
    G.tuning_set(d)


def test_split_tiled_image_matches_the_hbm_route(G, oracle):
    """Molecules whose packed tiled image does not fit LDS (194 AOs: 91 blocks): the split placement
    (leading blocks in LDS, the rest in the workgroup's HBM slab, tiles beyond a wave's hold through
    HBM) against the all-HBM route: the same number of products and densities within 1e-13 (the
    tiles sum in another order), and both within 1e-12 of the oracle's converged density."""
    from test_gpu_purify import _stage
    run = _stage(G)
    d = G.tuning_default()
    for seed in (9, 10, 11):
        z, x = W.solvated_conformers(1, seed)[0]
        z = np.asarray(z, np.int32)
        xb = W.to_bohr(x)
        st, o = oracle.run(z, xb, 0.0, oracle.options(p_tol=1e-10, e_tol=1e-12))
        F = oracle.fock(z, xb, o["density"])
        m = G.Molecule(z, xb)
        G.tuning_set(dict(d, sp2_lds=6))
        P6, k6, h6, l6 = run(m, F)
        G.tuning_set(dict(d, sp2_lds=0))
        P0, k0, h0, l0 = run(m, F)
        print("nao", F.shape[0], "products", k6, k0, "max|dP|", np.abs(P6 - P0).max(), flush=True)
        assert F.shape[0] == 194 and k6 == k0 > 0
        assert np.abs(P6 - P0).max() <= 1e-13
        assert np.abs(P6 - o["density"]).max() <= 1e-12
        assert abs(h6 - h0) <= 1e-10 and abs(l6 - l0) <= 1e-10
    G.tuning_set(d)


def test_every_entry_changes_speed_only(G):
    mols = _mols(G, 64, 88)
    d = G.tuning_default()
    s0, r0, _ = _solve(G, mols, d)
    assert np.all(s0 == 0)
    tables = [dict(d, sp2_tile16_n=1 << 20), dict(d, sp2_tile16_n=0), dict(d, pair_block=64, pair_blocks_per_cu=1),
              dict(d, pair_block=128, pair_blocks_per_cu=64), dict(d, sp2_lds=0), dict(d, sp2_lds=1),
              dict(d, sp2_lds=3), dict(d, sp2_lds=3, sp2_tile16_n=0), dict(d, sp2_lds=4), dict(d, sp2_lds=5),
              dict(d, sp2_tight_n=1 << 20), dict(d, sp2_tight_n=8)]
    for t in tables:
        s, r, ms = _solve(G, mols, t)
        dh = max(abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) for a, b in zip(r, r0))
        dp = max(np.abs(a.density - b.density).max() for a, b in zip(r, r0))
        print("table", {k: v for k, v in t.items() if v != d[k]}, "kernel ms %.2f" % ms, "max dHf %.1e" % dh,
              "max dP %.1e" % dp, flush=True)
        assert np.all(s == 0)
        assert dh <= HOF_TOL and dp <= 1e-7
    # the host-driven path's product routine
    big = [G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in [W.water_cluster(40, W.water_box_for(40), 40)]]
    sa, ra, _ = _solve(G, big, dict(d, large_split=0))
    for t in (dict(d, large_syrk=1), dict(d, large_split=2), dict(d, large_split=3)):
        sb, rb, _ = _solve(G, big, t)
        print("large table", {k: v for k, v in t.items() if v != d[k]}, "iterations", ra[0].iterations,
              rb[0].iterations, flush=True)
        assert sa[0] == 0 and sb[0] == 0
        assert abs(ra[0].heat_of_formation_kcal - rb[0].heat_of_formation_kcal) <= HOF_TOL
        assert np.abs(ra[0].density - rb[0].density).max() <= 1e-7
        assert abs(ra[0].homo_ev - rb[0].homo_ev) <= 1e-6 and abs(ra[0].lumo_ev - rb[0].lumo_ev) <= 1e-6
    G.tuning_set(d)


def test_autotune_keeps_answers_and_sets_the_table(G):
    mols = _mols(G, 96, 99)
    d = G.tuning_default()
    G.tuning_set(d)
    table, rep = G.autotune(mols, options=G.Pm3Options(p_tol=1e-8, e_tol=1e-8), reps=1)
    print("autotune", table, rep, flush=True)
    assert set(table) == set(d)
    assert G.tuning_get() == table
    assert rep["candidates"] >= 20 and rep["rejected"] == 0
    assert 0.0 < rep["best_ms"] <= rep["default_ms"]
    s0, r0, _ = _solve(G, mols, d)
    s1, r1, _ = _solve(G, mols, table)
    G.tuning_set(d)
    assert np.all(s0 == 0) and np.all(s1 == 0)
    assert max(abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) for a, b in zip(r0, r1)) <= HOF_TOL
