"""The numerical Hessian on the GPU (pm3_numerical_hessian: all 6N displaced SCF + gradient solves
in one batch) against the oracle's central differences, and the harmonic analysis of the
reference's documented water run (tests/api_surface.rs:534-621) on it."""
import numpy as np
import pytest

from conftest import atoms_to_zx

pytestmark = pytest.mark.gpu

WATER = [["O", 0.0, 0.0, 0.0], ["H", 0.9584, 0.0, 0.0], ["H", -0.24, 0.9278, 0.0]]


@pytest.fixture(scope="module")
def G():
    import pm3amd
    pm3amd.native()
    return pm3amd


def test_hessian_matches_the_oracle(G, oracle):
    from pm3amd.hessian import numerical_hessian
    cases = [(WATER, 0.0, 1),
             ([["C", 0, 0, 0], ["O", 0.05, 0, 1.25], ["H", 0.94, 0.1, -0.54], ["H", -0.94, 0, -0.5]], 0.0, 1),
             ([["N", 0, 0, 0], ["H", 1.02, 0, 0], ["H", -0.25, 0.98, 0.02]], 0.0, 2)]  # NH2, a doublet
    for atoms, q, mult in cases:
        z, x = atoms_to_zx(atoms)
        h = numerical_hessian(G.Molecule(z, x, q, mult))
        o = oracle.numerical_hessian(z, x, q, multiplicity=mult)
        d = np.abs(h - o).max()
        print("hessian", len(z), mult, "max|dH|", d, "max|H|", np.abs(o).max(), flush=True)
        assert np.array_equal(h, h.T)
        assert d <= 1e-5 * max(1.0, np.abs(o).max())


def test_water_vibrations(G):
    from pm3amd.hessian import vibrational_analysis
    z, x = atoms_to_zx(WATER)
    opt = G.optimize(G.Molecule(z, x))
    vm = vibrational_analysis(opt.molecule)
    print("water frequencies", vm.frequencies_cm, "residual", vm.rigid_residual_cm, flush=True)
    assert vm.hessian.shape == (9, 9) and vm.frequencies_cm.shape == (9,)
    assert np.all(np.diff(vm.frequencies_cm) >= 0.0) or vm.frequencies_cm.min() < 0.0
    assert vm.n_rigid == 6 and int(np.sum(vm.frequencies_cm == 0.0)) == 6
    vib = np.sort(vm.frequencies_cm)[6:]
    assert np.all(vib != 0.0) and np.all((vib > 1000.0) & (vib < 4200.0)), vib
    assert 0.0 < vm.rigid_residual_cm < 50.0
    assert np.all(np.signbit(vm.frequencies_cm) == np.signbit(vm.eigenvalues))


def test_hessian_refusals(G):
    from pm3amd.hessian import numerical_hessian
    z, x = atoms_to_zx(WATER)
    with pytest.raises(G.InvalidInput):
        numerical_hessian(G.Molecule(z, x), step=0.0)
    with pytest.raises((G.MissingElement, G.Unsupported)):
        numerical_hessian(G.Molecule(np.array([8, 2], np.int32), x[:2]))
