"""The host-driven multi-kernel path for molecules too large for one workgroup (grid-wide integral,
Fock and reduction kernels, rocBLAS products, purified densities or the rocSOLVER eigensolve):
the same molecule solved on both paths agrees, a cluster beyond the batched kernel's size limit
agrees with the CPU oracle on both routes, spectral bounds set inside the spectrum on purpose are
caught, and the 512-water cluster (3072 AOs) satisfies the size-independent invariants."""

import numpy as np
import pytest

from pm3amd import workloads as W

pytestmark = pytest.mark.gpu

HOF_TOL = 1e-5
TIGHT = dict(p_tol=1e-10, e_tol=1e-12, max_scf=500)


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def _mols(count, seed):
    return [(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in W.mixed_batch(count, seed)]


def test_forced_large_path_matches_batched(G, exact_eigensolves):
    mols = _mols(6, 99)
    ms = [G.Molecule(z, x) for z, x in mols]
    for opts in (G.Pm3Options(**TIGHT), G.Pm3Options(p_tol=1e-8, e_tol=1e-8)):
        a = G.Batch(ms, options=opts, want_density=True, want_gradient=True)
        b = G.Batch(ms, options=opts, want_density=True, want_gradient=True, force_large=True)
        a.solve()
        b.solve()
        sa, ra = a.results()
        sb, rb = b.results()
        ga, gb = a.gradients(), b.gradients()
        assert np.all(sa == 0) and np.all(sb == 0)
        for x, y, g1, g2 in zip(ra, rb, ga, gb):
            print("nao", x.density.shape[0], "iterations", x.iterations, y.iterations, flush=True)
            assert abs(x.heat_of_formation_kcal - y.heat_of_formation_kcal) <= HOF_TOL
            assert np.abs(x.density - y.density).max() <= (1e-8 if opts.p_tol < 1e-9 else 1e-6)
            assert np.abs(x.charges - y.charges).max() <= 1e-6
            assert np.abs(x.dipole_debye - y.dipole_debye).max() <= 1e-5
            assert abs(x.homo_ev - y.homo_ev) <= 1e-6 and abs(x.lumo_ev - y.lumo_ev) <= 1e-6
            assert np.abs(g1 - g2).max() <= 1e-6
            if opts.p_tol >= 1e-9:
                assert abs(x.iterations - y.iterations) <= 1


def test_cluster_beyond_workgroup_limit_matches_oracle(G, oracle):
    z, x = W.water_cluster(70, W.water_box_for(70), 70)
    z = np.asarray(z, np.int32)
    xb = W.to_bohr(x)
    mol = G.Molecule(z, xb)
    print("atoms", len(z), flush=True)
    r = G.run_pm3(mol, options=G.Pm3Options(**TIGHT))
    assert r.density.shape[0] == 420  # beyond the one-workgroup limit of about 190 AOs
    st, o = oracle.run(z, xb, 0.0, oracle.options(**TIGHT))
    assert st == 0 and r.converged
    assert abs(r.heat_of_formation_kcal - o["hof_kcal"]) <= HOF_TOL
    assert np.abs(r.density - o["density"]).max() <= 1e-8
    assert np.abs(r.charges - o["charges"]).max() <= 1e-8
    assert abs(r.homo_ev - o["homo_ev"]) <= 1e-7
    n = r.mo_coeff.shape[0]
    C = r.mo_coeff
    assert np.abs(C.T @ C - np.eye(n)).max() < 1e-10
    assert np.abs(2 * C[:, :r.n_occ] @ C[:, :r.n_occ].T - r.density).max() < 1e-10
    # the purified route (no orbitals requested)
    b = G.Batch([mol], options=G.Pm3Options(**TIGHT), want_density=True)
    b.solve()
    sp, rp = b.results()
    b.close()
    q = rp[0]
    print("purified route iterations", q.iterations, r.iterations, flush=True)
    assert sp[0] == 0 and abs(q.heat_of_formation_kcal - o["hof_kcal"]) <= HOF_TOL
    assert np.abs(q.density - o["density"]).max() <= 1e-8
    assert abs(q.homo_ev - o["homo_ev"]) <= 1e-7 and abs(q.lumo_ev - o["lumo_ev"]) <= 1e-7


def test_large_path_missed_spectral_bounds_are_caught(G, monkeypatch):
    """Purification bounds deliberately set inside the spectrum (negative margin) on a 240-AO
    cluster: each miss is caught and redone, and the answers are those of the default settings."""
    z, x = W.water_cluster(40, W.water_box_for(40), 41)
    mol = G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))
    opts = G.Pm3Options(p_tol=1e-8, e_tol=1e-8)
    out = []
    for env in ({"PM3_SP2_MARGIN": "-0.3", "PM3_SP2_TIGHT_N": "8"}, {}):
        for k in ("PM3_SP2_MARGIN", "PM3_SP2_TIGHT_N"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        b = G.Batch([mol], options=opts, want_density=True)
        b.solve()
        st, res = b.results()
        b.close()
        assert st[0] == 0
        out.append(res[0])
        print("env", env, "nao", res[0].density.shape[0], "iterations", res[0].iterations, flush=True)
    a, c = out
    assert a.density.shape[0] == 240
    assert abs(a.heat_of_formation_kcal - c.heat_of_formation_kcal) <= HOF_TOL
    assert np.abs(a.density - c.density).max() <= 1e-7
    assert abs(a.homo_ev - c.homo_ev) <= 1e-6 and abs(a.lumo_ev - c.lumo_ev) <= 1e-6


def test_water512_invariants(G):
    z, x = W.water_cluster(512, 25.0, 2026)
    z = np.asarray(z, np.int32)
    xb = W.to_bohr(x)
    mol = G.Molecule(z, xb)
    g = G.closed_form_gradient(mol, options=G.Pm3Options(p_tol=1e-8, e_tol=1e-8))
    r = g.scf
    n = r.density.shape[0]
    print("nao", n, "iterations", r.iterations, "HoF", r.heat_of_formation_kcal, flush=True)
    assert n == 3072 and r.converged and r.n_occ == 2048
    P = r.density
    assert np.abs(P - P.T).max() < 1e-12
    # idempotency of the closed-shell density (P/2)^2 = P/2 and the electron count
    assert np.abs((P / 2) @ (P / 2) - P / 2).max() < 1e-9
    assert abs(np.trace(P) - 4096.0) < 1e-8
    # translational invariance of the analytic gradient
    assert np.abs(g.gradient.sum(axis=0)).max() < 1e-8
    assert abs(r.total_ev - (r.electronic_ev + r.core_ev)) < 1e-9


def test_large_path_timing_counters(G):
    """The per-step timers of the multi-kernel path count what ran: one product per purification
    step enqueued (its issued FLOP), at least one Fock build per SCF
    iteration, two frontier Lanczos runs; zeros for batched molecules."""
    z, x = W.water_cluster(40, W.water_box_for(40), 41)
    b = G.Batch([G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))], options=G.Pm3Options(p_tol=1e-8, e_tol=1e-8),
                want_density=True)
    b.solve()
    st, res = b.results()
    t = b.large_timing()
    b.close()
    print("timing", t, "iterations", res[0].iterations, flush=True)
    assert st[0] == 0
    assert t["products"] >= 10 * res[0].iterations and t["product_ms"] > 0.0
    # up to 1024 AOs a step is one fused kernel over the 16 x 16 tiles of the upper triangle
    # (k running over the AOs rounded up to 64); larger matrices take rocBLAS products
    bm = (240 + 15) // 16
    assert t["product_flop"] == bm * (bm + 1) // 2 * 2 * 256 * 256 * t["products"]
    assert t["fock_builds"] >= res[0].iterations and t["fock_ms"] > 0.0
    assert t["lanczos_runs"] == 2 and 8 <= t["lanczos_steps"] <= 2 * (2 * 240 + 16)
    s = G.Batch([G.Molecule(z2, x2) for z2, x2 in _mols(4, 5)])
    s.solve()
    assert s.large_timing() == {"product_ms": 0.0, "products": 0, "product_flop": 0, "fock_ms": 0.0, "fock_builds": 0,
                                "lanczos_steps": 0, "lanczos_runs": 0}
    s.close()


def test_mixed_batch_of_small_and_large_molecules(G):
    """One batch holding molecules for the persistent kernel and one for the host-driven path, with
    gradients: every molecule's answers equal those of its own single-molecule gradient run (the
    queue puts the large molecule last, and its gradient gather starts after the small ones)."""
    small = _mols(3, 7)
    z, x = W.water_cluster(40, W.water_box_for(40), 41)
    mols = small + [(np.asarray(z, np.int32), W.to_bohr(x))]
    ms = [G.Molecule(zz, xx) for zz, xx in mols]
    opts = G.Pm3Options(p_tol=1e-10, e_tol=1e-12, max_scf=500)
    b = G.Batch(ms, options=opts, want_density=True, want_gradient=True)
    b.solve()
    st, res = b.results()
    grads = b.gradients()
    t = b.large_timing()
    b.close()
    assert np.all(st == 0) and t["products"] > 0
    for m, r, g in zip(ms, res, grads):
        one = G.closed_form_gradient(m, options=opts)
        print("nao", r.density.shape[0], "dHf", abs(r.heat_of_formation_kcal - one.scf.heat_of_formation_kcal),
              "dg", np.abs(g - one.gradient).max(), flush=True)
        assert abs(r.heat_of_formation_kcal - one.scf.heat_of_formation_kcal) <= HOF_TOL
        assert np.abs(g - one.gradient).max() <= 1e-6
        assert np.abs(g.sum(axis=0)).max() < 1e-8


def _sym_step(G):
    import ctypes
    L = G.native()
    dp = ctypes.POINTER(ctypes.c_double)
    L.pm3_stage_sym_step.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, dp, dp, dp, dp, ctypes.c_int32,
                                     ctypes.c_int32, dp]

    def run(x, square, slab, reps=0, workgroups=0):
        n = x.shape[0]
        x = np.ascontiguousarray(x)
        y = np.zeros((n, n))
        tr, f2, ms = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = L.pm3_stage_sym_step(n, square, slab, x.ctypes.data_as(dp), y.ctypes.data_as(dp), ctypes.byref(tr),
                                  ctypes.byref(f2), reps, workgroups, ctypes.byref(ms))
        assert st == 0, G.native().pm3_last_error_message()
        return y, tr.value, f2.value, ms.value
    return run


@pytest.mark.parametrize("slab", [8, 16, 32])
def test_tile_kernel_step_matches_numpy_at_ragged_sizes(G, slab):
    """One purification step of the lower-triangle tile kernel (64 x 64 tiles, mirrored store) on
    symmetric matrices of ragged sizes: y = x x and 2 x - x x against numpy to 1e-13 relative,
    exactly symmetric, every element written (y starts as NaN, x's padding is NaN), trace and
    Frobenius sums to 1e-12, and the same bits on a second call."""
    run = _sym_step(G)
    rng = np.random.default_rng(7 + slab)
    for n in (1, 2, 15, 16, 17, 31, 63, 64, 65, 127, 129, 200, 420, 1000):
        a = rng.standard_normal((n, n))
        x = (a + a.T) / (2.0 * np.sqrt(n))
        x2 = x @ x
        scale = np.abs(x2).max() + 1.0
        for square in (1, 0):
            ref = x2 if square else 2.0 * x - x2
            y, tr, f2, _ = run(x, square, slab)
            assert np.all(np.isfinite(y)), n
            assert np.array_equal(y, y.T), n
            err = np.abs(y - ref).max() / scale
            assert err <= 1e-13, (n, square, err)
            assert abs(tr - np.trace(ref)) <= 1e-12 * max(1.0, abs(np.trace(ref))), n
            assert abs(f2 - np.sum(ref * ref)) <= 1e-12 * max(1.0, np.sum(ref * ref)), n
            y2, tr2, f22, _ = run(x, square, slab)
            assert np.array_equal(y, y2) and tr == tr2 and f22 == f2, n
            # a capped grid (each workgroup several tiles): the same elements, bit for bit
            y3, tr3, f23, _ = run(x, square, slab, workgroups=16)
            assert np.array_equal(y, y3), n
            assert abs(tr3 - tr) <= 1e-12 * max(1.0, abs(tr)) and abs(f23 - f2) <= 1e-12 * max(1.0, f2), n
        print("n", n, "slab", slab, "max rel err", err, flush=True)


def test_tile_kernel_large_path_matches_oracle(G, oracle, monkeypatch):
    """The 420-AO cluster through the multi-kernel path with its purification products on the tile
    kernel (large_split 1, fused-kernel ceiling lowered below the size), both slab widths: the
    oracle's answers to the same tolerances as the default route."""
    z, x = W.water_cluster(70, W.water_box_for(70), 70)
    z = np.asarray(z, np.int32)
    xb = W.to_bohr(x)
    st, o = oracle.run(z, xb, 0.0, oracle.options(**TIGHT))
    assert st == 0
    d = G.tuning_default()
    monkeypatch.setenv("PM3_LARGE_STEPS_N", "64")
    try:
        for slab in ("8", "16", "32"):
            monkeypatch.setenv("PM3_SY_K", slab)
            G.tuning_set(dict(d, large_split=1))
            b = G.Batch([G.Molecule(z, xb)], options=G.Pm3Options(**TIGHT), want_density=True)
            b.solve()
            s, r = b.results()
            t = b.large_timing()
            b.close()
            q = r[0]
            print("slab", slab, "iterations", q.iterations, "products", t["products"], flush=True)
            assert s[0] == 0 and abs(q.heat_of_formation_kcal - o["hof_kcal"]) <= HOF_TOL
            assert np.abs(q.density - o["density"]).max() <= 1e-8
            assert abs(q.homo_ev - o["homo_ev"]) <= 1e-7 and abs(q.lumo_ev - o["lumo_ev"]) <= 1e-7
            nt = (420 + 63) // 64
            assert t["product_flop"] == nt * (nt + 1) // 2 * 2 * 64 * 64 * ((420 + int(slab) - 1) // int(slab) * int(slab)) * t["products"]
    finally:
        G.tuning_set(d)
