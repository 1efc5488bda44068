"""The oracle's response_fock_rhf (src/fock.rs:444-485, the CPHF response operator) pinned by what
the mathematics fixes: G(dP) = F(P + dP) - F(P) through the separately pinned Fock build, its
linearity and symmetry, the isolated atoms' one-centre blocks when every pair is switched off, and
each pair switch acting linearly and only on its own atoms' blocks (the reference's pair order)."""
import numpy as np
import pytest

from conftest import atoms_to_zx

MOLS = {
    "water": [["O", 0, 0, 0], ["H", 0.96, 0, 0], ["H", -0.24, 0.93, 0]],
    "methanol": [["C", 0, 0, 0], ["O", 1.42, 0.02, 0], ["H", 1.75, 0.93, 0.05], ["H", -0.37, 1.03, 0.02],
                 ["H", -0.38, -0.5, 0.9], ["H", -0.36, -0.52, -0.88]],
    # hydrogens before heavy atoms: the reference's pair order follows the input, not heavy-first
    "hcn_h_first": [["H", -1.07, 0, 0], ["C", 0, 0, 0], ["N", 1.16, 0, 0]],
    "sih3cl": [["Si", 0, 0, 0], ["Cl", 2.05, 0, 0], ["H", -0.5, 1.4, 0], ["H", -0.5, -0.7, 1.2],
               ["H", -0.5, -0.7, -1.2]],
}


def _sym(rng, n, k=None):
    a = rng.standard_normal((n, n) if k is None else (k, n, n))
    return 0.5 * (a + np.swapaxes(a, -1, -2))


@pytest.mark.parametrize("name", sorted(MOLS))
def test_response_is_the_fock_difference(oracle, name):
    z, x = atoms_to_zx(MOLS[name])
    n = oracle.nao(z)
    rng = np.random.default_rng(3)
    P, dP = _sym(rng, n), _sym(rng, n)
    G = oracle.response_fock(z, x, dP)
    D = oracle.fock(z, x, P + dP) - oracle.fock(z, x, P)
    assert np.abs(G - D).max() <= 1e-12 * max(1.0, np.abs(G).max())
    # symmetric up to the order of the one-centre sums, which the reference keeps per element
    assert np.abs(G - G.T).max() <= 1e-14 * max(1.0, np.abs(G).max())


def test_response_is_linear_and_stacks(oracle):
    z, x = atoms_to_zx(MOLS["methanol"])
    n = oracle.nao(z)
    rng = np.random.default_rng(4)
    d1, d2 = _sym(rng, n), _sym(rng, n)
    g1, g2 = oracle.response_fock(z, x, d1), oracle.response_fock(z, x, d2)
    g12 = oracle.response_fock(z, x, 0.7 * d1 - 1.3 * d2)
    assert np.abs(g12 - (0.7 * g1 - 1.3 * g2)).max() <= 1e-12 * np.abs(g12).max()
    st = oracle.response_fock(z, x, np.stack([d1, d2]))
    assert np.array_equal(st[0], g1) and np.array_equal(st[1], g2)


def test_switched_off_pairs_leave_the_isolated_atoms(oracle):
    z, x = atoms_to_zx(MOLS["sih3cl"])
    n, nat = oracle.nao(z), len(z)
    rng = np.random.default_rng(5)
    dP = _sym(rng, n)
    G0 = oracle.response_fock(z, x, dP, np.zeros(nat * (nat - 1) // 2))
    assert np.array_equal(oracle.response_fock(z, x, dP, np.ones(nat * (nat - 1) // 2)), oracle.response_fock(z, x, dP))
    off = np.cumsum([0] + [1 if v == 1 else 4 for v in z])
    for a in range(nat):
        blk = slice(off[a], off[a + 1])
        alone = oracle.response_fock(z[a:a + 1], x[a:a + 1], dP[blk, blk])
        assert np.array_equal(G0[blk, blk], alone), a
        for b in range(nat):
            if b != a:
                assert not G0[blk, off[b]:off[b + 1]].any()


def test_each_switch_acts_linearly_on_its_own_pair(oracle):
    z, x = atoms_to_zx(MOLS["hcn_h_first"])
    n, nat = oracle.nao(z), len(z)
    npairs = nat * (nat - 1) // 2
    rng = np.random.default_rng(6)
    dP = _sym(rng, n)
    off = np.cumsum([0] + [1 if v == 1 else 4 for v in z])
    pairs = [(u, v) for u in range(nat) for v in range(u + 1, nat)]
    G0 = oracle.response_fock(z, x, dP, np.zeros(npairs))
    parts = []
    for p, (u, v) in enumerate(pairs):
        e = np.zeros(npairs)
        e[p] = 1.0
        d = oracle.response_fock(z, x, dP, e) - G0
        parts.append(d)
        mask = np.zeros((n, n), bool)
        for a in (u, v):
            for b in (u, v):
                mask[off[a]:off[a + 1], off[b]:off[b + 1]] = True
        assert not d[~mask].any(), (p, u, v)
        assert np.abs(d[off[u]:off[u + 1], off[v]:off[v + 1]]).max() > 1e-6  # exchange on the pair's block
        e[p] = 0.3
        half = oracle.response_fock(z, x, dP, e) - G0
        assert np.abs(half - 0.3 * d).max() <= 1e-12 * np.abs(d).max()
    full = oracle.response_fock(z, x, dP) - G0
    assert np.abs(sum(parts) - full).max() <= 1e-12 * np.abs(full).max()
