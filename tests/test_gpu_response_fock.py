"""The CPHF response operator on the GPU (pm3_response_fock, response_fock_rhf of src/fock.rs:
444-485): many right-hand sides per call, with and without the per-pair switch, against the
oracle's restatement; and against the GPU's own Fock build, G(dP) = F(P + dP) - F(P)."""
import numpy as np
import pytest

from conftest import atoms_to_zx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def _sym(rng, n, k):
    a = rng.standard_normal((k, n, n))
    return 0.5 * (a + np.swapaxes(a, -1, -2))


def _cases():
    from pm3amd import workloads as W
    out = [("water", atoms_to_zx([["O", 0, 0, 0], ["H", 0.96, 0, 0], ["H", -0.24, 0.93, 0]])),
           ("hcn_h_first", atoms_to_zx([["H", -1.07, 0, 0], ["C", 0, 0, 0], ["N", 1.16, 0, 0]])),
           ("sih3cl", atoms_to_zx([["Si", 0, 0, 0], ["Cl", 2.05, 0, 0], ["H", -0.5, 1.4, 0], ["H", -0.5, -0.7, 1.2],
                                   ["H", -0.5, -0.7, -1.2]])),
           ("helium_atom", (np.array([2], np.int32), np.zeros((1, 3))))]
    z, x = W.alkane(8)
    out.append(("octane", (np.asarray(z, np.int32), W.to_bohr(np.asarray(x)))))
    z, x = W.water_cluster(40, 10.0 * (40 / 32) ** (1 / 3), 5)  # 240 AOs: beyond one workgroup's LDS
    out.append(("water40", (np.asarray(z, np.int32), W.to_bohr(np.asarray(x)))))
    return out


def test_response_matches_the_oracle(G, oracle):
    rng = np.random.default_rng(12)
    for name, (z, x) in _cases():
        if name == "helium_atom":
            continue  # no PM3 parameters for He here: refused below
        m = G.Molecule(z, x)
        n = oracle.nao(z)
        k = 3 if n > 100 else 9
        dp = _sym(rng, n, k)
        nat = len(z)
        sw = rng.uniform(0.0, 1.0, nat * (nat - 1) // 2)
        sw[::5] = 0.0
        for pair_sw in (None, sw):
            got = G.response_fock(m, dp, pair_sw)
            ref = oracle.response_fock(z, x, dp, pair_sw)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"case {name} n {n} rhs {k} switch {pair_sw is not None} rel {err:.1e}", flush=True)
            assert err <= 1e-13, name


def test_response_is_the_gpu_fock_difference(G):
    from pm3amd import workloads as W
    z, x = W.alkane(5)
    z, x = np.asarray(z, np.int32), W.to_bohr(np.asarray(x))
    m = G.Molecule(z, x)
    import ctypes
    n = G.native().pm3_query_nao(ctypes.byref(m._c()), G.Pm3Parameters.standard()._h)
    rng = np.random.default_rng(13)
    P, dP = _sym(rng, n, 2), _sym(rng, n, 2)
    L = G.native()
    dp_t = ctypes.POINTER(ctypes.c_double)
    L.pm3_stage_fock.argtypes = [ctypes.POINTER(G._CMolecule), ctypes.c_void_p, dp_t, dp_t]
    prm = G.Pm3Parameters.standard()

    def fock(Pm):  # the SCF kernel's own Fock build, H_core + G(P) (pm3_stage_fock)
        Pm = np.ascontiguousarray(Pm)
        assert Pm.shape == (n, n)  # the C entry point writes n x n doubles
        F = np.zeros_like(Pm)
        assert L.pm3_stage_fock(ctypes.byref(m._c()), prm._h, Pm.ctypes.data_as(dp_t), F.ctypes.data_as(dp_t)) == 0
        return F

    for k in range(2):
        D = fock(P[k] + dP[k]) - fock(P[k])
        got = G.response_fock(m, dP[k])
        assert np.abs(got - D).max() <= 1e-11 * np.abs(D).max()


def test_one_call_many_right_hand_sides_equals_single_calls(G):
    from pm3amd import workloads as W
    z, x = W.alkane(6)
    m = G.Molecule(np.asarray(z, np.int32), W.to_bohr(np.asarray(x)))
    n = 6 * 6 + 2
    dp = _sym(np.random.default_rng(14), n, 64)
    many = G.response_fock(m, dp)
    for r in (0, 17, 63):
        assert np.array_equal(many[r], G.response_fock(m, dp[r]))


def test_response_refusals(G):
    z, x = atoms_to_zx([["O", 0, 0, 0], ["H", 0.96, 0, 0], ["H", -0.24, 0.93, 0]])
    m = G.Molecule(z, x)
    with pytest.raises(G.InvalidInput):
        G.response_fock(m, np.zeros((5, 5)))
    with pytest.raises(G.InvalidInput):
        G.response_fock(m, np.zeros((6, 6)), pair_sw=[1.0, 1.0])
    with pytest.raises(G.InvalidInput):
        G.response_fock(m, np.zeros((6, 6)), pair_sw=[1.0, np.nan, 1.0])
    with pytest.raises((G.MissingElement, G.Unsupported)):
        G.response_fock(G.Molecule(np.array([2], np.int32), np.zeros((1, 3))), np.zeros((1, 1)))
