"""The numerical Hessian and the harmonic analysis on the host: the oracle's central-difference
Hessian (numerical_hessian, src/hessian.rs:268-322) against what the mathematics fixes, and the
mode analysis (vibrational_analysis, src/hessian.rs:327-414, src/rigid.rs) against closed forms."""
import numpy as np

from conftest import atoms_to_zx

WATER = [["O", 0.0, 0.0, 0.0], ["H", 0.9584, 0.0, 0.0], ["H", -0.24, 0.9278, 0.0]]


def test_oracle_hessian_is_a_second_derivative(oracle):
    z, x = atoms_to_zx(WATER)
    h = oracle.numerical_hessian(z, x)
    assert np.array_equal(h, h.T)
    # translation invariance: a uniform displacement leaves the gradient unchanged
    for k in range(3):
        assert np.abs(h[:, k::3].sum(axis=1)).max() <= 1e-4 * np.abs(h).max()
    # diagonal elements against a second difference of the energy
    opts = oracle.options(p_tol=1e-11, e_tol=1e-12, max_scf=500)
    step = 2e-3
    for j in (0, 4, 8):
        e = []
        for s in (-1, 0, 1):
            xs = x.reshape(-1).copy()
            xs[j] += s * step
            e.append(oracle.run(z, xs.reshape(-1, 3), 0.0, opts, want_arrays=False)[1]["total_ev"])
        d2 = (e[0] - 2 * e[1] + e[2]) / step ** 2
        assert abs(d2 - h[j, j]) <= 2e-3 * abs(h[j, j]), (j, d2, h[j, j])


def test_diatomic_modes_match_the_closed_form():
    """A spring k between two masses along x: five rigid directions (linear), exactly five zeros,
    one vibration of 521.4709 sqrt(k / mu) cm^-1 with k in eV/Angstrom^2."""
    from pm3amd.hessian import ANGSTROM_TO_BOHR, modes_from_hessian
    k_bohr = 1.7  # eV/Bohr^2
    m = np.array([12.011, 15.999])
    h = np.zeros((6, 6))
    h[0, 0] = h[3, 3] = k_bohr
    h[0, 3] = h[3, 0] = -k_bohr
    x = np.array([[0.0, 0.0, 0.0], [2.1, 0.0, 0.0]])
    vm = modes_from_hessian(h, x, m)
    assert vm.n_rigid == 5
    assert int(np.sum(vm.frequencies_cm == 0.0)) == 5
    mu = m[0] * m[1] / (m[0] + m[1])
    expect = 521.4709 * np.sqrt(k_bohr * ANGSTROM_TO_BOHR ** 2 / mu)
    assert abs(vm.frequencies_cm.max() - expect) <= 1e-9 * expect
    assert vm.rigid_residual_cm <= 1e-4  # the rigid directions are null up to rounding


def test_rigid_basis_rank_and_sign_convention():
    from pm3amd.hessian import rigid_body_basis, signed_wavenumber
    m = np.array([1.008, 15.999, 1.008])
    bent = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [2.3, 1.7, 0.0]])
    line = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [3.6, 0.0, 0.0]])
    for x, rank in ((bent, 6), (line, 5), (bent[:1], 3)):
        b = rigid_body_basis(x, m[:len(x)])
        assert len(b) == rank
        g = np.stack(b) @ np.stack(b).T
        assert np.abs(g - np.eye(rank)).max() <= 1e-12
    assert signed_wavenumber(-4.0) == -2 * 521.4709 and signed_wavenumber(4.0) == 2 * 521.4709


def test_water_modes_from_the_oracle_hessian(oracle):
    # tests/api_surface.rs:561-621 on the oracle's Hessian at the oracle's minimum: six rigid
    # directions, exactly six zeros, three vibrations between 1000 and 4200 cm^-1, a small
    # non-zero residual, signs carried from the eigenvalues
    from pm3amd.hessian import modes_from_hessian
    z, x = atoms_to_zx(WATER)
    xo = oracle.optimize(z, x)["xyz"]
    masses = np.array([15.999, 1.008, 1.008])
    vm = modes_from_hessian(oracle.numerical_hessian(z, xo), xo, masses)
    assert vm.n_rigid == 6 and int(np.sum(vm.frequencies_cm == 0.0)) == 6
    vib = np.sort(vm.frequencies_cm)[6:]
    assert np.all((vib > 1000.0) & (vib < 4200.0)), vib
    assert 0.0 < vm.rigid_residual_cm < 50.0
    assert np.all(np.signbit(vm.frequencies_cm) == np.signbit(vm.eigenvalues))
