"""Second derivatives on the PM3 surface: the Cartesian Hessian by central differences of the
GPU's analytic gradient (numerical_hessian, src/hessian.rs:268-322, all 6N displaced gradients
in one batched solve, pm3_numerical_hessian), and the harmonic vibrational analysis built on it
(vibrational_analysis, src/hessian.rs:327-414, with the rigid-body projection of src/rigid.rs).

The reference's vibrational_analysis takes the CPHF analytic Hessian; this engine has the CPHF
response operator on the device (pm3amd.response_fock) but not the second-derivative integrals,
so the frequencies here come from the numerical Hessian, which the reference offers as the
independent check of its analytic one (tests/api_surface.rs:551-559, agreement within 5e-3
eV/Bohr^2).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import List, Optional

import numpy as np

from . import InvalidInput, MissingElement, Molecule, Pm3Options, Pm3Parameters, _dp, _raise, native

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
# sqrt(eV / (Angstrom^2 amu)) -> cm^-1 (src/hessian.rs:109-110)
SQRT_EV_PER_ANG2_AMU_TO_CM = 521.4709


def numerical_hessian(molecule: Molecule, params: Optional[Pm3Parameters] = None,
                      options: Optional[Pm3Options] = None, step: float = 1.0e-3) -> np.ndarray:
    """The 3N x 3N Cartesian Hessian (eV/Bohr^2): central differences of the analytic gradient,
    symmetrised by averaging, SCF options tightened as tighten_scf_for_hessian does."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    L = native()
    L.pm3_numerical_hessian.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_double)]
    n = 3 * len(molecule.z)
    h = np.zeros((n, n))
    cm = molecule._c()
    _raise(L.pm3_numerical_hessian(ctypes.byref(cm), params._h, ctypes.byref(options._c()), float(step), _dp(h)))
    return h


def element_masses(z, params: Optional[Pm3Parameters] = None) -> np.ndarray:
    """Atomic masses (amu) of the parameter set (pm3_element_mass)."""
    params = params or Pm3Parameters.standard()
    L = native()
    L.pm3_element_mass.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.pm3_element_mass.restype = ctypes.c_double
    m = np.array([L.pm3_element_mass(params._h, int(v)) for v in np.asarray(z).reshape(-1)])
    if not np.all(np.isfinite(m)):
        raise MissingElement("no PM3 parameters in this engine for an element of the molecule")
    return m


def rigid_body_basis(positions, masses, rotations=True) -> List[np.ndarray]:
    """Orthonormal basis of the rigid-body motions in mass-weighted coordinates (rigid_body_basis,
    src/rigid.rs:59-116): sqrt(m) along each axis, then sqrt(m) (e_axis x (r - centre of mass)),
    modified Gram-Schmidt with two passes, a candidate kept when its residual norm exceeds 1e-6.
    Three, five or six vectors: the count of exactly-zero modes."""
    x = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    m = np.asarray(masses, dtype=np.float64).reshape(-1)
    nat = len(m)
    total = m.sum()
    centre = (m[:, None] * x).sum(axis=0) / total if total > 0.0 else np.zeros(3)
    cands = []
    for axis in range(3):
        v = np.zeros(3 * nat)
        v[axis::3] = np.sqrt(m)
        cands.append(v)
    if rotations:
        d = x - centre
        for axis in range(3):
            e = np.zeros(3)
            e[axis] = 1.0
            cross = np.cross(e, d)
            cands.append((np.sqrt(m)[:, None] * cross).reshape(-1))
    basis: List[np.ndarray] = []
    for c in cands:
        c = c.copy()
        for _ in range(2):
            for b in basis:
                c -= float(np.dot(c, b)) * b
        norm = float(np.sqrt(np.dot(c, c)))
        if norm > 1.0e-6:
            basis.append(c / norm)
    return basis


def project_out_symmetric(basis, a: np.ndarray) -> np.ndarray:
    """(1 - P) A (1 - P) for the projector P onto the basis (project_out_symmetric,
    src/rigid.rs:170-211), applied from both sides so the result stays symmetric."""
    if not basis:
        return a.copy()
    b = np.stack(basis, axis=1)
    w = a @ b
    inner = b.T @ w
    return a - b @ w.T - w @ b.T + b @ inner @ b.T


def rigid_mode_indices(basis, modes: np.ndarray) -> np.ndarray:
    """The len(basis) eigenvectors with the largest weight in the rigid subspace, ascending
    (rigid_mode_indices, src/rigid.rs:283-305)."""
    if not basis:
        return np.zeros(0, dtype=int)
    b = np.stack(basis, axis=1)
    weight = ((b.T @ modes) ** 2).sum(axis=0)
    order = sorted(range(modes.shape[1]), key=lambda k: -weight[k])
    return np.array(sorted(order[:len(basis)]), dtype=int)


def signed_wavenumber(eigenvalue):
    """cm^-1 from a mass-weighted eigenvalue (eV / (Angstrom^2 amu)); negative = imaginary
    (src/hessian.rs:116-122)."""
    e = np.asarray(eigenvalue, dtype=np.float64)
    return np.where(e >= 0.0, SQRT_EV_PER_ANG2_AMU_TO_CM * np.sqrt(np.abs(e)),
                    -SQRT_EV_PER_ANG2_AMU_TO_CM * np.sqrt(np.abs(e)))


@dataclasses.dataclass
class VibrationalModes:
    """VibrationalModes (src/hessian.rs:124-146): the Cartesian Hessian (eV/Bohr^2), the number of
    rigid-body directions projected out, how far from zero the Hessian put them before that
    (cm^-1), frequencies (cm^-1, ascending, negative = imaginary), the mass-weighted eigenvalues
    and the modes (columns)."""
    hessian: np.ndarray
    n_rigid: int
    rigid_residual_cm: float
    frequencies_cm: np.ndarray
    eigenvalues: np.ndarray
    modes: np.ndarray


def modes_from_hessian(hessian, positions_bohr, masses, rotations=True) -> VibrationalModes:
    """The harmonic analysis of vibrational_analysis (src/hessian.rs:337-414) for a given Hessian:
    mass weighting to eV/(Angstrom^2 amu) (a massless atom's rows and columns stay zero), the
    rigid-body residual, the projection, the eigen-decomposition (ascending), the rigid modes'
    eigenvalues set to exactly zero, and the signed wavenumbers."""
    h = np.asarray(hessian, dtype=np.float64)
    m = np.repeat(np.asarray(masses, dtype=np.float64), 3)
    mij = np.sqrt(np.outer(m, m))
    with np.errstate(divide="ignore", invalid="ignore"):
        mw = np.where(mij > 0.0, h * ANGSTROM_TO_BOHR ** 2 / np.where(mij > 0.0, mij, 1.0), 0.0)
    basis = rigid_body_basis(positions_bohr, masses, rotations)
    residual = 0.0
    for b in basis:
        residual = max(residual, abs(float(signed_wavenumber(float(b @ mw @ b)))))
    mw = project_out_symmetric(basis, mw)
    eigs, modes = np.linalg.eigh(0.5 * (mw + mw.T))
    eigs = eigs.copy()
    eigs[rigid_mode_indices(basis, modes)] = 0.0
    return VibrationalModes(h, len(basis), residual, signed_wavenumber(eigs), eigs, modes)


def vibrational_analysis(molecule: Molecule, params: Optional[Pm3Parameters] = None,
                         options: Optional[Pm3Options] = None, step: float = 1.0e-3) -> VibrationalModes:
    """Harmonic frequencies at the molecule's geometry (a stationary point is assumed, as in the
    reference) from the GPU's numerical Hessian."""
    params = params or Pm3Parameters.standard()
    if len(molecule.z) < 1:
        raise InvalidInput("empty molecule")
    h = numerical_hessian(molecule, params, options, step)
    return modes_from_hessian(h, molecule.positions, element_masses(molecule.z, params))
