"""Python mirror of the reference's molecular PM3 interface, backed by libpm3gpu.so (MI355X).

Names, argument meaning and error behaviour follow the reference crate:

* ``Molecule`` (src/system.rs:28-40; positions in Bohr, ``from_xyz_str`` converts from Angstrom),
* ``Pm3Parameters.standard()`` (src/params.rs:135),
* ``Pm3Options`` with the ``Pm3Options::default()`` values (src/scf.rs:259-287),
* ``run_pm3(molecule, params, options) -> Pm3Result`` (src/scf.rs:450-848),
* ``closed_form_gradient(...) -> GradientResult`` (src/gradient.rs:201-276),
* errors as ``Pm3Error`` subclasses mirroring ``Pm3Error`` (src/error.rs:12-47); non-convergence raises
  ``ScfNotConverged`` rather than returning ``converged=False`` (src/scf.rs:754-760).

Every call runs on the GPU through the C ABI in include/pm3_gpu.h; there is no CPU fallback.  A
missing or unloadable library raises at import of the native handle.
"""

import ctypes
import dataclasses
import os
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libpm3gpu.so")

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903   # src/constants.rs:33
EV_TO_KCAL = 23.06054783061903           # src/constants.rs:27

SYMBOLS = {"H": 1, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Na": 11, "Mg": 12, "Al": 13,
           "Si": 14, "P": 15, "S": 16, "Cl": 17, "He": 2, "Ne": 10, "Ar": 18, "K": 19, "Ca": 20}


class Pm3Error(RuntimeError):
    status = -1


class InvalidInput(Pm3Error):
    status = 1


class MissingElement(Pm3Error):
    """Pm3Error::MissingElement(z): ``z`` is the atomic number without parameters."""
    status = 2

    def __init__(self, msg, z=0):
        super().__init__(msg)
        self.z = z


class ScfNotConverged(Pm3Error):
    status = 3

    def __init__(self, msg, iterations=0, error=float("nan")):
        super().__init__(msg)
        self.iterations = iterations
        self.error = error


class Unsupported(Pm3Error):
    """An input the reference handles and this engine does not (Pm3Error::UnsupportedDOrbitals(z) for
    an element outside its s/p set; ``z`` is 0 otherwise)."""
    status = 4

    def __init__(self, msg, z=0):
        super().__init__(msg)
        self.z = z


class ResourceLimit(Pm3Error):
    """Pm3Error::ResourceLimit{operation, required_mb, limit_mb}."""
    status = 5

    def __init__(self, msg, required_mb=0.0, limit_mb=0.0):
        super().__init__(msg)
        self.operation = "PM3 two-electron pair cache"
        self.required_mb = required_mb
        self.limit_mb = limit_mb


class DeviceError(Pm3Error):
    status = 6


_ERRORS = {1: InvalidInput, 2: MissingElement, 3: ScfNotConverged, 4: Unsupported, 5: ResourceLimit, 6: DeviceError}

ACCEL = {"none": 0, "cdiis": 1, "adiis_cdiis": 2}
GUESS = {"auto": 0, "sad": 1, "core": 2, "symmetry_broken": 3, "learned": 4}
STABILITY = {"off": 0, "auto": 1, "always": 2}
REFERENCE = {"auto": 0, "rhf": 1, "uhf": 2}


class _COptions(ctypes.Structure):
    _fields_ = [("max_scf", ctypes.c_int32), ("e_tol", ctypes.c_double), ("p_tol", ctypes.c_double),
                ("use_diis", ctypes.c_int32), ("accelerator", ctypes.c_int32), ("adiis_switch", ctypes.c_double),
                ("guess", ctypes.c_int32), ("stability", ctypes.c_int32), ("scf_memory_mb", ctypes.c_int32),
                ("guess_model", ctypes.c_void_p), ("level_shift_ev", ctypes.c_double), ("damping", ctypes.c_double),
                ("reference", ctypes.c_int32), ("has_field", ctypes.c_int32), ("field", ctypes.c_double * 3),
                ("integral_memory_mb", ctypes.c_int32)]


class _CMolecule(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int32), ("z", ctypes.POINTER(ctypes.c_int32)),
                ("xyz_bohr", ctypes.POINTER(ctypes.c_double)), ("charge", ctypes.c_double),
                ("multiplicity", ctypes.c_int32)]


class _CResult(ctypes.Structure):
    _fields_ = [("total_ev", ctypes.c_double), ("electronic_ev", ctypes.c_double), ("core_ev", ctypes.c_double),
                ("heat_of_formation_kcal", ctypes.c_double), ("dipole_debye", ctypes.c_double * 3),
                ("dipole_magnitude", ctypes.c_double), ("homo_ev", ctypes.c_double), ("lumo_ev", ctypes.c_double),
                ("last_change", ctypes.c_double), ("scf_improvement_ev", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("scf_paths_tried", ctypes.c_int32),
                ("n_occ", ctypes.c_int32), ("nao", ctypes.c_int32), ("has_homo", ctypes.c_int32),
                ("has_lumo", ctypes.c_int32), ("guess_used", ctypes.c_int32), ("unrestricted", ctypes.c_int32),
                ("n_beta", ctypes.c_int32), ("missing_z", ctypes.c_int32), ("required_mb", ctypes.c_double),
                ("limit_mb", ctypes.c_double)]


class _CDcOptions(ctypes.Structure):
    _fields_ = [("core_radius", ctypes.c_double), ("buffer_radius", ctypes.c_double),
                ("min_core_atoms", ctypes.c_int32), ("electron_tol", ctypes.c_double), ("smearing_ev", ctypes.c_double)]


class _CDcResult(ctypes.Structure):
    _fields_ = [("total_ev", ctypes.c_double), ("electronic_ev", ctypes.c_double), ("core_ev", ctypes.c_double),
                ("heat_of_formation_kcal", ctypes.c_double), ("fermi_ev", ctypes.c_double),
                ("fermi_beta_ev", ctypes.c_double), ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("unrestricted", ctypes.c_int32), ("dropped_pairs", ctypes.c_int32),
                ("largest_subsystem", ctypes.c_int32), ("n_subsystems", ctypes.c_int32), ("nao", ctypes.c_int32)]


class _COptOptions(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("gtol", ctypes.c_double), ("grad_step", ctypes.c_double),
                ("history", ctypes.c_int32), ("warm_start", ctypes.c_int32)]


class _COptResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("converged", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("evaluations", ctypes.c_int32), ("trajectory_len", ctypes.c_int32), ("energy_ev", ctypes.c_double),
                ("heat_of_formation_kcal", ctypes.c_double), ("max_gradient", ctypes.c_double),
                ("first_energy_ev", ctypes.c_double)]


class _COptStep(ctypes.Structure):
    _fields_ = [("energy_ev", ctypes.c_double), ("heat_of_formation_kcal", ctypes.c_double),
                ("max_gradient", ctypes.c_double)]


EXPORTED = ["pm3_params_standard", "pm3_options_default", "pm3_query_nao", "pm3_run", "pm3_gradient",
            "pm3_run_batch", "pm3_batch_create", "pm3_batch_solve", "pm3_batch_results", "pm3_batch_destroy",
            "pm3_batch_last_kernel_ms", "pm3_batch_gradient", "pm3_batch_last_pair_ms", "pm3_batch_last_guess_ms", "pm3_batch_guess_timing",
            "pm3_batch_phase_ticks", "pm3_stage_hcore", "pm3_stage_fock", "pm3_stage_eigen", "pm3_stage_eigen_warm",
            "pm3_stage_gemm", "pm3_stage_purify", "pm3_last_error_message",
            "pm3_guess_model_create", "pm3_guess_model_load", "pm3_guess_model_save", "pm3_guess_model_get_params",
            "pm3_guess_model_set_params", "pm3_guess_model_steps", "pm3_guess_model_destroy", "pm3_guess_data_create",
            "pm3_guess_data_add", "pm3_guess_data_stats", "pm3_guess_data_destroy", "pm3_guess_train_options_default",
            "pm3_guess_train", "pm3_guess_eval", "pm3_guess_grad", "pm3_guess_mlp", "pm3_guess_features", "pm3_guess_density",
            "pm3_tuning_default", "pm3_tuning_get", "pm3_tuning_set", "pm3_tuning_load", "pm3_tuning_save", "pm3_autotune",
            "pm3_batch_large_timing", "pm3_batch_set_positions", "pm3_opt_options_default", "pm3_optimize_batch",
            "pm3_optimize", "pm3_batch_carry_densities", "pm3_batch_spin_densities", "pm3_run_noalloc",
            "pm3_response_fock", "pm3_numerical_hessian", "pm3_element_mass", "pm3_dc_options_default", "pm3_run_dc", "pm3_dc_gradient",
            "pm3_stage_sym_step", "pm3_optimize_dc"]

_lib = None


def native():
    """The loaded libpm3gpu.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpm3gpu.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
        L.pm3_params_standard.restype = ctypes.c_void_p
        L.pm3_options_default.argtypes = [ctypes.POINTER(_COptions)]
        L.pm3_query_nao.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p]
        L.pm3_run.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                              ctypes.POINTER(_CResult), dp, dp, dp, dp, dp, dp, dp]
        L.pm3_gradient.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                   ctypes.POINTER(_CResult), dp, dp, dp]
        L.pm3_response_fock.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.c_int32, dp, dp, dp]
        L.pm3_dc_options_default.argtypes = [ctypes.POINTER(_CDcOptions)]
        L.pm3_run_dc.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                 ctypes.POINTER(_CDcOptions), ctypes.POINTER(_CDcResult), dp, dp, dp]
        L.pm3_dc_gradient.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                      ctypes.POINTER(_CDcOptions), ctypes.POINTER(_CDcResult), dp, dp, dp, dp]
        L.pm3_run_batch.argtypes = [ctypes.c_int32, ip, ip, dp, dp, ip, ctypes.c_void_p, ctypes.POINTER(_COptions),
                                    ctypes.c_int32, ip, ctypes.POINTER(_CResult), dp]
        L.pm3_batch_create.argtypes = [ctypes.c_int32, ip, ip, dp, dp, ip, ctypes.c_void_p,
                                       ctypes.POINTER(_COptions), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.pm3_batch_spin_densities.argtypes = [ctypes.c_void_p, dp]
        L.pm3_run_noalloc.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                      ctypes.POINTER(_CResult), dp, dp]
        L.pm3_batch_solve.argtypes = [ctypes.c_void_p]
        L.pm3_batch_results.argtypes = [ctypes.c_void_p, ip, ctypes.POINTER(_CResult), dp, dp]
        L.pm3_batch_destroy.argtypes = [ctypes.c_void_p]
        L.pm3_batch_gradient.argtypes = [ctypes.c_void_p, dp]
        L.pm3_batch_last_kernel_ms.argtypes = [ctypes.c_void_p]
        L.pm3_batch_last_kernel_ms.restype = ctypes.c_double
        L.pm3_batch_last_pair_ms.argtypes = [ctypes.c_void_p]
        L.pm3_batch_last_pair_ms.restype = ctypes.c_double
        L.pm3_batch_last_guess_ms.argtypes = [ctypes.c_void_p]
        L.pm3_batch_last_guess_ms.restype = ctypes.c_double
        L.pm3_batch_guess_timing.argtypes = [ctypes.c_void_p, dp, dp, ctypes.POINTER(ctypes.c_int64)]
        L.pm3_batch_phase_ticks.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        L.pm3_batch_large_timing.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64), dp, ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.pm3_last_error_message.restype = ctypes.c_char_p
        vp, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)
        up, i64p = ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_int64)
        L.pm3_guess_model_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp)]
        L.pm3_guess_model_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.pm3_guess_model_save.argtypes = [vp, ctypes.c_char_p]
        L.pm3_guess_model_get_params.argtypes = [vp, fp]
        L.pm3_guess_model_set_params.argtypes = [vp, fp]
        L.pm3_guess_model_steps.argtypes = [vp]
        L.pm3_guess_model_destroy.argtypes = [vp]
        L.pm3_guess_data_create.argtypes = [ctypes.POINTER(vp)]
        L.pm3_guess_data_add.argtypes = [vp, vp, i64p]
        L.pm3_guess_data_stats.argtypes = [vp, i64p, i64p, i64p]
        L.pm3_guess_data_destroy.argtypes = [vp]
        L.pm3_guess_train.argtypes = [vp, vp, vp, fp, dp]
        L.pm3_guess_eval.argtypes = [vp, vp, dp]
        L.pm3_guess_grad.argtypes = [vp, vp, fp, dp]
        L.pm3_guess_mlp.argtypes = [vp, ctypes.c_int32, up, fp, fp, fp]
        L.pm3_guess_features.argtypes = [ctypes.POINTER(_CMolecule), vp, ip, ip, ip, up]
        L.pm3_guess_density.argtypes = [vp, ctypes.POINTER(_CMolecule), vp, dp, dp, dp]
        L.pm3_tuning_default.argtypes = [ctypes.POINTER(_CTuning)]
        L.pm3_tuning_default.restype = None
        L.pm3_tuning_get.argtypes = [ctypes.POINTER(_CTuning)]
        L.pm3_tuning_set.argtypes = [ctypes.POINTER(_CTuning)]
        L.pm3_tuning_load.argtypes = [ctypes.c_char_p]
        L.pm3_tuning_save.argtypes = [ctypes.c_char_p]
        L.pm3_autotune.argtypes = [ctypes.c_int32, ip, ip, dp, dp, ctypes.c_void_p, ctypes.POINTER(_COptions),
                                   ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_CTuning),
                                   ctypes.POINTER(_CTuneReport)]
        L.pm3_batch_set_positions.argtypes = [ctypes.c_void_p, dp]
        L.pm3_batch_carry_densities.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ip]
        L.pm3_opt_options_default.argtypes = [ctypes.POINTER(_COptOptions)]
        L.pm3_opt_options_default.restype = None
        L.pm3_optimize_batch.argtypes = [ctypes.c_int32, ip, ip, dp, dp, ip, ctypes.c_void_p,
                                         ctypes.POINTER(_COptions), ctypes.POINTER(_COptOptions),
                                         ctypes.POINTER(_COptResult), dp, ctypes.POINTER(_COptStep), dp]
        L.pm3_optimize.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                   ctypes.POINTER(_COptOptions), ctypes.POINTER(_COptResult), dp,
                                   ctypes.POINTER(_COptStep), dp]
        L.pm3_optimize_dc.argtypes = [ctypes.POINTER(_CMolecule), ctypes.c_void_p, ctypes.POINTER(_COptions),
                                      ctypes.POINTER(_CDcOptions), ctypes.POINTER(_COptOptions),
                                      ctypes.POINTER(_COptResult), ctypes.POINTER(_CDcResult), dp,
                                      ctypes.POINTER(_COptStep), dp]
        _lib = L
    return _lib


class _CTuning(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("pair_block", "pair_blocks_per_cu", "sp2_tile16_n", "sp2_lds",
                                              "sp2_tight_n", "large_syrk", "large_split", "wg128_nao", "wg256_nao",
                                              "solo_large_nao")]


class _CTuneReport(ctypes.Structure):
    _fields_ = [("default_ms", ctypes.c_double), ("best_ms", ctypes.c_double), ("candidates", ctypes.c_int32),
                ("rejected", ctypes.c_int32)]


TUNING_KEYS = tuple(k for k, _ in _CTuning._fields_)


def tuning_default():
    """The built-in tuning table (launch shapes and tile choices) as a dict."""
    t = _CTuning()
    native().pm3_tuning_default(ctypes.byref(t))
    return {k: getattr(t, k) for k in TUNING_KEYS}


def tuning_get():
    """The process-wide tuning table that batch creation reads, as a dict."""
    t = _CTuning()
    _raise(native().pm3_tuning_get(ctypes.byref(t)))
    return {k: getattr(t, k) for k in TUNING_KEYS}


def tuning_set(table):
    """Set the process-wide table from a dict (missing keys keep their current values)."""
    cur = tuning_get()
    cur.update(table)
    t = _CTuning(**{k: int(cur[k]) for k in TUNING_KEYS})
    _raise(native().pm3_tuning_set(ctypes.byref(t)))


def tuning_load(path):
    _raise(native().pm3_tuning_load(os.fsencode(path)))


def tuning_save(path):
    _raise(native().pm3_tuning_save(os.fsencode(path)))


def autotune(molecules, params=None, options=None, want_density=False, want_gradient=False, force_large=False,
             reps=3):
    """Tune the launch shapes and tile choices on this workload (pm3_autotune): returns
    (table dict, report dict); the process-wide table is left at the best table found."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    ptr = np.zeros(len(molecules) + 1, dtype=np.int32)
    for k, m in enumerate(molecules):
        ptr[k + 1] = ptr[k] + len(m.z)
    z = np.ascontiguousarray(np.concatenate([m.z for m in molecules]).astype(np.int32))
    xyz = np.ascontiguousarray(np.concatenate([m.positions for m in molecules]))
    q = np.ascontiguousarray([float(m.charge) for m in molecules], dtype=np.float64)
    ip = ctypes.POINTER(ctypes.c_int32)
    t, rep = _CTuning(), _CTuneReport()
    flags = (1 if want_density else 0) | (2 if want_gradient else 0) | (4 if force_large else 0)
    _raise(native().pm3_autotune(len(molecules), ptr.ctypes.data_as(ip), z.ctypes.data_as(ip), _dp(xyz), _dp(q),
                                 params._h, ctypes.byref(options._c()), flags, int(reps), ctypes.byref(t),
                                 ctypes.byref(rep)))
    return ({k: getattr(t, k) for k in TUNING_KEYS},
            {"default_ms": rep.default_ms, "best_ms": rep.best_ms, "candidates": rep.candidates,
             "rejected": rep.rejected})


def _raise(status, res=None):
    if status == 0:
        return
    msg = native().pm3_last_error_message().decode()
    cls = _ERRORS.get(status, Pm3Error)
    if cls is ScfNotConverged and res is not None:
        raise ScfNotConverged(msg, res.iterations, res.last_change)
    if cls in (MissingElement, Unsupported) and res is not None:
        raise cls(msg, res.missing_z)
    if cls is ResourceLimit and res is not None:
        raise ResourceLimit(msg, res.required_mb, res.limit_mb)
    raise cls(msg)


@dataclasses.dataclass
class Molecule:
    """Atoms with positions in Bohr (src/system.rs:21-40)."""
    z: np.ndarray
    positions: np.ndarray
    charge: float = 0.0
    multiplicity: int = 1

    def __post_init__(self):
        self.z = np.ascontiguousarray(self.z, dtype=np.int32)
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float64).reshape(-1, 3)

    @classmethod
    def from_xyz_str(cls, text, charge=0.0):
        """Parse an XYZ block in Angstrom (src/system.rs:90-139)."""
        lines = text.lstrip("﻿").splitlines()
        try:
            n = int(lines[0].strip())
        except (IndexError, ValueError):
            raise InvalidInput(f"invalid XYZ atom count: {lines[0] if lines else ''}")
        if len(lines) < n + 2:
            raise InvalidInput("XYZ ended before all atoms")
        z, pos = [], []
        for i in range(n):
            parts = lines[i + 2].split()
            if len(parts) < 4:
                raise InvalidInput(f"XYZ line {i + 3} has fewer than 4 fields")
            if parts[0] not in SYMBOLS:
                raise InvalidInput(f"unknown element on line {i + 3}: {parts[0]}")
            z.append(SYMBOLS[parts[0]])
            pos.append([float(v) * ANGSTROM_TO_BOHR for v in parts[1:4]])
        return cls(np.array(z), np.array(pos), charge)

    def to_xyz_str(self, comment="pm3-rs optimized geometry; coordinates in Angstrom"):
        """XYZ block in Angstrom, the layout of the reference's writer (src/cli.rs:753-790):
        symbol padded to two columns, coordinates signed with ten decimals."""
        names = {v: k for k, v in SYMBOLS.items()}
        lines = [str(len(self.z)), comment]
        for zz, p in zip(self.z.tolist(), (self.positions / ANGSTROM_TO_BOHR).tolist()):
            lines.append(f"{names.get(int(zz), 'X'):2} {p[0]:+.10f} {p[1]:+.10f} {p[2]:+.10f}")
        return "\n".join(lines) + "\n"

    def _c(self):
        m = _CMolecule(len(self.z), self.z.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                       self.positions.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(self.charge),
                       int(self.multiplicity))
        return m


class Pm3Parameters:
    """Handle on the embedded PM3 table (Pm3Parameters::standard(), src/params.rs:135)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def standard(cls):
        h = native().pm3_params_standard()
        if not h:
            raise Pm3Error(native().pm3_last_error_message().decode())
        return cls(h)


@dataclasses.dataclass
class Pm3Options:
    """Pm3Options::default() (src/scf.rs:259-287) for the fields the molecular single point reads.
    field: None or three components in eV / Bohr / e (Pm3Options::field); reference: "auto",
    "rhf" or "uhf"; integral_memory_mb 0 takes the process default ceiling (4096 MiB)."""
    max_scf: int = 200
    e_tol: float = 1.0e-8
    p_tol: float = 1.0e-7
    use_diis: bool = True
    accelerator: str = "adiis_cdiis"
    adiis_switch: float = 0.1
    guess: str = "auto"
    stability: str = "auto"
    scf_memory_mb: int = 512
    guess_model: Optional[object] = None  # pm3amd.guess.GuessModel, required by guess="learned"
    level_shift_ev: float = 0.0
    damping: float = 0.0
    reference: str = "auto"
    field: Optional[tuple] = None
    integral_memory_mb: int = 0

    def _c(self):
        f = (ctypes.c_double * 3)(*(self.field if self.field is not None else (0.0, 0.0, 0.0)))
        return _COptions(self.max_scf, self.e_tol, self.p_tol, int(self.use_diis), ACCEL[self.accelerator],
                         self.adiis_switch, GUESS[self.guess], STABILITY[self.stability], self.scf_memory_mb,
                         self.guess_model.handle if self.guess_model is not None else None,
                         float(self.level_shift_ev), float(self.damping), REFERENCE[self.reference],
                         int(self.field is not None), f, int(self.integral_memory_mb))


@dataclasses.dataclass
class Pm3Result:
    """Pm3Result (src/scf.rs:352-396) for a restricted solution."""
    density: Optional[np.ndarray]
    mo_energies: Optional[np.ndarray]
    mo_coeff: Optional[np.ndarray]
    n_occ: int
    electronic_ev: float
    core_ev: float
    total_ev: float
    heat_of_formation_kcal: float
    charges: np.ndarray
    dipole_debye: np.ndarray
    dipole_magnitude: float
    homo_ev: Optional[float]
    lumo_ev: Optional[float]
    iterations: int
    converged: bool
    scf_paths_tried: int
    scf_improvement_ev: float
    unrestricted: bool = False
    guess_used: int = 0  # 0 superposition / other rule, 1 learned, 2 learned rejected
    n_beta: int = 0
    spin_density: Optional[np.ndarray] = None       # P_alpha - P_beta, unrestricted results
    mo_energies_beta: Optional[np.ndarray] = None
    mo_coeff_beta: Optional[np.ndarray] = None


@dataclasses.dataclass
class GradientResult:
    """GradientResult (src/gradient.rs:122-134): gradient in eV/Bohr, forces = -gradient."""
    scf: Pm3Result
    energy_ev: float
    gradient: np.ndarray
    forces: np.ndarray
    max_gradient: float


def _result(r, P, eps, C, q, S=None, eps_b=None, C_b=None):
    u = bool(r.unrestricted)
    return Pm3Result(P, eps, C, r.n_occ, r.electronic_ev, r.core_ev, r.total_ev, r.heat_of_formation_kcal, q,
                     np.array(r.dipole_debye[:]), r.dipole_magnitude, r.homo_ev if r.has_homo else None,
                     r.lumo_ev if r.has_lumo else None, r.iterations, bool(r.converged), r.scf_paths_tried,
                     r.scf_improvement_ev, u, r.guess_used, r.n_beta, S if u else None, eps_b if u else None,
                     C_b if u else None)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def run_pm3(molecule: Molecule, params: Optional[Pm3Parameters] = None, options: Optional[Pm3Options] = None,
            want_orbitals=True) -> Pm3Result:
    """RHF PM3 single point on the GPU (run_pm3, src/scf.rs:450)."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    if nao < 0:  # an element without parameters here: the library reports which and why
        res = _CResult()
        null = ctypes.POINTER(ctypes.c_double)()
        st = native().pm3_run(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(res), null, null,
                              null, null, null, null, null)
        _raise(st if st else 2, res)
    res = _CResult()
    P = np.zeros((nao, nao))
    S = np.zeros((nao, nao))
    q = np.zeros(len(molecule.z))
    eps = np.zeros(nao) if want_orbitals else None
    C = np.zeros((nao, nao)) if want_orbitals else None
    eps_b = np.zeros(nao) if want_orbitals else None
    C_b = np.zeros((nao, nao)) if want_orbitals else None
    null = ctypes.POINTER(ctypes.c_double)()
    o = lambda a: _dp(a) if a is not None else null  # noqa: E731
    st = native().pm3_run(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(res), _dp(P), o(eps),
                          o(C), _dp(q), _dp(S), o(eps_b), o(C_b))
    _raise(st, res)
    return _result(res, P, eps, C, q, S, eps_b, C_b)


@dataclasses.dataclass
class DcOptions:
    """DcOptions (src/dc/partition.rs:40-97), lengths in Bohr; long_range_cutoff is not offered."""
    core_radius: float = 6.0
    buffer_radius: float = 9.0
    min_core_atoms: int = 1
    electron_tol: float = 1.0e-9
    smearing_ev: float = 0.1

    def _c(self):
        return _CDcOptions(self.core_radius, self.buffer_radius, self.min_core_atoms, self.electron_tol,
                           self.smearing_ev)


@dataclasses.dataclass
class DcResult:
    """DcResult (src/dc/scf.rs:25-46)."""
    density: np.ndarray
    spin_density: Optional[np.ndarray]
    unrestricted: bool
    electronic_ev: float
    core_ev: float
    total_ev: float
    heat_of_formation_kcal: float
    charges: np.ndarray
    fermi_ev: float
    fermi_beta_ev: float
    iterations: int
    converged: bool
    dropped_pairs: int
    largest_subsystem: int
    n_subsystems: int


def run_dc(molecule: Molecule, params: Optional[Pm3Parameters] = None, options: Optional[Pm3Options] = None,
           dc: Optional[DcOptions] = None) -> DcResult:
    """Divide-and-conquer SCF on the GPU (run_dc, src/dc/scf.rs:74): batched subsystem eigensolves,
    one global chemical potential, Pulay DIIS on the density.  Raises ScfNotConverged like the
    reference."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    dc = dc or DcOptions()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    res = _CDcResult()
    null = ctypes.POINTER(ctypes.c_double)()
    if nao < 0:
        st = native().pm3_run_dc(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(dc._c()),
                                 ctypes.byref(res), null, null, null)
        _raise(st if st else 2)
    P = np.zeros((nao, nao))
    S = np.zeros((nao, nao))
    q = np.zeros(len(molecule.z))
    st = native().pm3_run_dc(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(dc._c()),
                             ctypes.byref(res), _dp(P), _dp(S), _dp(q))
    if st == 3:
        raise ScfNotConverged(native().pm3_last_error_message().decode(), res.iterations, float("nan"))
    _raise(st)
    return DcResult(P, S if res.unrestricted else None, bool(res.unrestricted), res.electronic_ev, res.core_ev,
                    res.total_ev, res.heat_of_formation_kcal, q, res.fermi_ev, res.fermi_beta_ev, res.iterations,
                    bool(res.converged), res.dropped_pairs, res.largest_subsystem, res.n_subsystems)


@dataclasses.dataclass
class DcGradient:
    """DcGradient (src/dc/derivatives.rs:34-40): the divide-and-conquer result and its gradient."""
    scf: DcResult
    energy_ev: float
    gradient: np.ndarray  # eV/Bohr, n_atoms x 3
    forces: np.ndarray
    max_gradient: float


def dc_gradient(molecule: Molecule, params: Optional[Pm3Parameters] = None, options: Optional[Pm3Options] = None,
                dc: Optional[DcOptions] = None) -> DcGradient:
    """dc_gradient (src/dc/derivatives.rs:55-121) on the GPU: run_dc, then the fixed-density
    analytic gradient at the partitioned density, minus f q_A in a field."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    dc = dc or DcOptions()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    res = _CDcResult()
    g = np.zeros((len(molecule.z), 3))
    if nao < 0:
        null = ctypes.POINTER(ctypes.c_double)()
        st = native().pm3_dc_gradient(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(dc._c()),
                                      ctypes.byref(res), _dp(g), null, null, null)
        _raise(st if st else 2)
    P = np.zeros((nao, nao))
    S = np.zeros((nao, nao))
    q = np.zeros(len(molecule.z))
    st = native().pm3_dc_gradient(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(dc._c()),
                                  ctypes.byref(res), _dp(g), _dp(P), _dp(S), _dp(q))
    if st == 3:
        raise ScfNotConverged(native().pm3_last_error_message().decode(), res.iterations, float("nan"))
    _raise(st)
    scf = DcResult(P, S if res.unrestricted else None, bool(res.unrestricted), res.electronic_ev, res.core_ev,
                   res.total_ev, res.heat_of_formation_kcal, q, res.fermi_ev, res.fermi_beta_ev, res.iterations,
                   bool(res.converged), res.dropped_pairs, res.largest_subsystem, res.n_subsystems)
    return DcGradient(scf, res.total_ev, g, -g, float(np.abs(g).max()) if g.size else 0.0)


def run_pm3_noalloc(molecule: Molecule, params: Optional[Pm3Parameters] = None,
                    options: Optional[Pm3Options] = None) -> Pm3Result:
    """run_pm3 through pm3_run_noalloc: no device allocation (the library's static per-GPU scratch
    block), at most 128 atoms and 128 AOs, purified densities (no orbitals)."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    if nao < 0:
        raise MissingElement("no PM3 parameters for an element of the molecule")
    res = _CResult()
    P = np.zeros((nao, nao))
    q = np.zeros(len(molecule.z))
    st = native().pm3_run_noalloc(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(res), _dp(P),
                                  _dp(q))
    _raise(st, res)
    return _result(res, P, None, None, q)


def closed_form_gradient(molecule: Molecule, params: Optional[Pm3Parameters] = None,
                         options: Optional[Pm3Options] = None) -> GradientResult:
    """SCF plus the analytic Hellmann-Feynman gradient on the GPU (src/gradient.rs:201)."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    if nao < 0:
        raise MissingElement("no PM3 parameters in this engine for an element of the molecule")
    res = _CResult()
    g = np.zeros((len(molecule.z), 3))
    P = np.zeros((nao, nao))
    q = np.zeros(len(molecule.z))
    st = native().pm3_gradient(ctypes.byref(cm), params._h, ctypes.byref(options._c()), ctypes.byref(res), _dp(g),
                               _dp(P), _dp(q))
    _raise(st, res)
    scf = _result(res, P, None, None, q)
    return GradientResult(scf, res.total_ev, g, -g, float(np.abs(g).max()) if g.size else 0.0)


def response_fock(molecule: Molecule, dp, pair_sw=None, params: Optional[Pm3Parameters] = None) -> np.ndarray:
    """response_fock_rhf (src/fock.rs:444-485) on the GPU: G(dP) = J(dP) - K(dP / 2), no H_core,
    for one symmetric density (nao, nao) or a stack (k, nao, nao) in one call (pm3_response_fock).
    pair_sw: None, or one weight per atom pair in the reference's order (u < v, row by row)."""
    params = params or Pm3Parameters.standard()
    cm = molecule._c()
    nao = native().pm3_query_nao(ctypes.byref(cm), params._h)
    if nao < 0:
        raise MissingElement("no PM3 parameters in this engine for an element of the molecule")
    dp = np.ascontiguousarray(dp, dtype=np.float64)
    one = dp.ndim == 2
    stack = dp.reshape(1, *dp.shape) if one else dp
    if stack.ndim != 3 or stack.shape[1:] != (nao, nao):
        raise InvalidInput(f"dp must be ({nao}, {nao}) or (k, {nao}, {nao})")
    nat = len(molecule.z)
    sw = None
    if pair_sw is not None:
        sw = np.ascontiguousarray(pair_sw, dtype=np.float64).reshape(-1)
        if sw.size != nat * (nat - 1) // 2:
            raise InvalidInput(f"pair_sw needs {nat * (nat - 1) // 2} values, one per atom pair")
    out = np.zeros_like(stack)
    _raise(native().pm3_response_fock(ctypes.byref(cm), params._h, stack.shape[0], _dp(stack),
                                      None if sw is None else _dp(sw), _dp(out)))
    return out[0] if one else out


def single_point(molecule: Molecule, params: Optional[Pm3Parameters] = None, options: Optional[Pm3Options] = None,
                 want_orbitals=True) -> Pm3Result:
    """run_pm3 under the name of the reference's Python layer (src/python.rs:253)."""
    return run_pm3(molecule, params, options, want_orbitals)


class Batch:
    """Device-resident batch of independent molecules (pm3_batch_* in include/pm3_gpu.h)."""

    def __init__(self, molecules: List[Molecule], params: Optional[Pm3Parameters] = None,
                 options: Optional[Pm3Options] = None, want_density=False, want_gradient=False, force_large=False,
                 warm_start=False):
        self.params = params or Pm3Parameters.standard()
        self.options = options or Pm3Options()
        self.molecules = molecules
        ptr = np.zeros(len(molecules) + 1, dtype=np.int32)
        for k, m in enumerate(molecules):
            ptr[k + 1] = ptr[k] + len(m.z)
        self.atom_ptr = ptr
        self._z = np.ascontiguousarray(np.concatenate([m.z for m in molecules]).astype(np.int32))
        self._xyz = np.ascontiguousarray(np.concatenate([m.positions for m in molecules]))
        self._q = np.ascontiguousarray([float(m.charge) for m in molecules], dtype=np.float64)
        self._mult = np.ascontiguousarray([int(m.multiplicity) for m in molecules], dtype=np.int32)
        h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int32)
        st = native().pm3_batch_create(len(molecules), ptr.ctypes.data_as(ip), self._z.ctypes.data_as(ip),
                                       _dp(self._xyz), _dp(self._q), self._mult.ctypes.data_as(ip), self.params._h,
                                       ctypes.byref(self.options._c()),
                                       (1 if want_density else 0) | (2 if want_gradient else 0) |
                                       (4 if force_large else 0) | (8 if warm_start else 0), ctypes.byref(h))
        _raise(st)
        self._h = h
        self.want_density = want_density

    def solve(self):
        _raise(native().pm3_batch_solve(self._h))

    def last_kernel_ms(self):
        """Device time of the last solve's persistent SCF launch (HIP events on its stream)."""
        return native().pm3_batch_last_kernel_ms(self._h)

    def last_pair_ms(self):
        """Device time of the last solve's pair-integral launch."""
        return native().pm3_batch_last_pair_ms(self._h)

    def last_guess_ms(self):
        """Device time of the last solve's learned-guess launches (features + network)."""
        return native().pm3_batch_last_guess_ms(self._h)

    def guess_timing(self):
        """(feature kernel ms, network kernel ms, directed pairs) of the last solve."""
        f, n, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _raise(native().pm3_batch_guess_timing(self._h, ctypes.byref(f), ctypes.byref(n), ctypes.byref(p)))
        return f.value, n.value, p.value

    def large_timing(self):
        """Molecules on the multi-kernel path, last solve: purification products (ms, count, FLOP),
        Fock builds (ms, count), from HIP events around each step, and frontier Lanczos steps/runs."""
        pm, fm = ctypes.c_double(), ctypes.c_double()
        np_, fl, nf = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        ls, lr = ctypes.c_int64(), ctypes.c_int64()
        _raise(native().pm3_batch_large_timing(self._h, ctypes.byref(pm), ctypes.byref(np_), ctypes.byref(fl),
                                               ctypes.byref(fm), ctypes.byref(nf), ctypes.byref(ls), ctypes.byref(lr)))
        return {"product_ms": pm.value, "products": np_.value, "product_flop": fl.value, "fock_ms": fm.value,
                "fock_builds": nf.value, "lanczos_steps": ls.value, "lanczos_runs": lr.value}

    PHASES = ("hcore", "fock", "energy_diis_gram", "products", "eigen_or_purify", "sort_or_frontier", "density",
              "other", "diis_solve", "diis_combine")

    def phase_ticks(self):
        """Per-phase wall-clock ticks (10 ns) of the last solve (PM3_PHASE_TIMERS=1 at creation)."""
        out = (ctypes.c_uint64 * 16)()
        _raise(native().pm3_batch_phase_ticks(self._h, out))
        d = dict(zip(self.PHASES, [int(v) for v in out[:10]]))
        d["counts"] = {"sweeps_or_products": int(out[13]), "eigensolves_or_purifications": int(out[14]),
                       "rounds_or_fallbacks": int(out[15]), "frontier_lanczos_steps": int(out[10]),
                       "frontier_lanczos_runs": int(out[11])}
        return d

    def results(self):
        """(status array, list of Pm3Result-like objects with density when requested)."""
        k = len(self.molecules)
        status = np.zeros(k, dtype=np.int32)
        res = (_CResult * k)()
        q = np.zeros(int(self.atom_ptr[-1]))
        nao = [native().pm3_query_nao(ctypes.byref(m._c()), self.params._h) for m in self.molecules]
        total = sum(max(n, 0) ** 2 for n in nao)
        P = np.zeros(max(total, 1)) if self.want_density else None
        null = ctypes.POINTER(ctypes.c_double)()
        _raise(native().pm3_batch_results(self._h, status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), res, _dp(q),
                                          _dp(P) if self.want_density else null))
        S = None
        if self.want_density:
            S = np.zeros(max(total, 1))
            _raise(native().pm3_batch_spin_densities(self._h, _dp(S)))
        out, moff = [], 0
        for i in range(k):
            a0, a1 = self.atom_ptr[i], self.atom_ptr[i + 1]
            n = max(nao[i], 0)
            Pi = P[moff:moff + n * n].reshape(n, n) if self.want_density else None
            Si = S[moff:moff + n * n].reshape(n, n) if self.want_density else None
            moff += n * n
            out.append(_result(res[i], Pi, None, None, q[a0:a1].copy(), Si))
        return status, out

    def set_positions(self, positions):
        """New positions (Bohr), one (n_atoms, 3) array per molecule, for the next solve
        (pm3_batch_set_positions); the batch is unchanged if they are refused."""
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3) for p in positions]))
        if xyz.shape != self._xyz.shape:
            raise InvalidInput("positions do not match the batch's atoms")
        _raise(native().pm3_batch_set_positions(self._h, _dp(xyz)))
        self._xyz = xyz
        self.molecules = [Molecule(m.z, xyz[self.atom_ptr[i]:self.atom_ptr[i + 1]].copy(), m.charge, m.multiplicity)
                          for i, m in enumerate(self.molecules)]

    def gradients(self):
        """Per-molecule analytic gradients (eV/Bohr) of the last solve (needs want_gradient=True)."""
        g = np.zeros((int(self.atom_ptr[-1]), 3))
        _raise(native().pm3_batch_gradient(self._h, _dp(g)))
        return [g[self.atom_ptr[i]:self.atom_ptr[i + 1]].copy() for i in range(len(self.molecules))]

    def close(self):
        if getattr(self, "_h", None):
            native().pm3_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclasses.dataclass
class OptOptions:
    """OptOptions (src/optimizer.rs:15-35); grad_step is accepted and unused (analytic gradient).
    warm_start (no reference counterpart) starts each round's SCF from the previous density."""
    max_iter: int = 200
    gtol: float = 1.0e-3
    grad_step: float = 5.0e-4
    history: int = 8
    warm_start: bool = False

    def _c(self):
        return _COptOptions(int(self.max_iter), float(self.gtol), float(self.grad_step), int(self.history),
                            int(bool(self.warm_start)))


@dataclasses.dataclass
class OptStep:
    """OptStep (src/optimizer.rs:37-43): one accepted point; positions in Bohr."""
    energy_ev: float
    heat_of_formation_kcal: float
    max_gradient: float
    positions: np.ndarray


@dataclasses.dataclass
class OptResult:
    """OptResult (src/optimizer.rs:45-52): the final molecule, the flags and counts, the energies
    of the first and last steps, and the trajectory (the starting point and every accepted point;
    empty when it was not asked for)."""
    molecule: Molecule
    converged: bool
    iterations: int
    evaluations: int
    energy_ev: float
    heat_of_formation_kcal: float
    max_gradient: float
    first_energy_ev: float
    trajectory: List[OptStep] = dataclasses.field(default_factory=list)


def _opt_result(m, r, xyz, steps=None, traj=None):
    path = []
    if steps is not None:
        n = len(m.z)
        path = [OptStep(steps[j].energy_ev, steps[j].heat_of_formation_kcal, steps[j].max_gradient,
                        traj[3 * n * j:3 * n * (j + 1)].reshape(n, 3).copy()) for j in range(int(r.trajectory_len))]
    return OptResult(Molecule(m.z, xyz, m.charge, m.multiplicity), bool(r.converged), int(r.iterations),
                     int(r.evaluations), r.energy_ev, r.heat_of_formation_kcal, r.max_gradient, r.first_energy_ev,
                     path)


def optimize_batch(molecules: List[Molecule], params: Optional[Pm3Parameters] = None,
                   options: Optional[Pm3Options] = None, opt: Optional[OptOptions] = None, trajectory: bool = True):
    """L-BFGS geometry optimisation of many molecules at once (pm3_optimize_batch): the optimiser
    of src/optimizer.rs:77-262 per molecule, one batched SCF + gradient solve per round.
    Returns (status array, list of OptResult).  trajectory=False skips the per-step record
    (it holds max_iter + 1 geometries per molecule)."""
    params = params or Pm3Parameters.standard()
    options = options or Pm3Options()
    opt = opt or OptOptions()
    k = len(molecules)
    ptr = np.zeros(k + 1, dtype=np.int32)
    for i, m in enumerate(molecules):
        ptr[i + 1] = ptr[i] + len(m.z)
    z = np.ascontiguousarray(np.concatenate([m.z for m in molecules]).astype(np.int32)) if k else np.zeros(0, np.int32)
    xyz = np.ascontiguousarray(np.concatenate([m.positions for m in molecules])) if k else np.zeros((0, 3))
    q = np.ascontiguousarray([float(m.charge) for m in molecules], dtype=np.float64)
    mult = np.ascontiguousarray([int(m.multiplicity) for m in molecules], dtype=np.int32)
    out = np.zeros_like(xyz)
    res = (_COptResult * max(k, 1))()
    cap = max(int(opt.max_iter), 0) + 1
    steps = (_COptStep * max(k * cap, 1))() if trajectory else None
    traj = np.zeros(3 * cap * int(ptr[-1])) if trajectory else None
    ip = ctypes.POINTER(ctypes.c_int32)
    _raise(native().pm3_optimize_batch(k, ptr.ctypes.data_as(ip), z.ctypes.data_as(ip), _dp(xyz), _dp(q),
                                       mult.ctypes.data_as(ip), params._h, ctypes.byref(options._c()),
                                       ctypes.byref(opt._c()), res, _dp(out), steps,
                                       None if traj is None else _dp(traj)))
    status = np.array([res[i].status for i in range(k)], dtype=np.int32)
    results = []
    for i, m in enumerate(molecules):
        st_i = (_COptStep * cap).from_buffer(steps, ctypes.sizeof(_COptStep) * cap * i) if trajectory else None
        tr_i = traj[3 * cap * int(ptr[i]):3 * cap * int(ptr[i + 1])] if trajectory else None
        results.append(_opt_result(m, res[i], out[ptr[i]:ptr[i + 1]].copy(), st_i, tr_i))
    return status, results


def optimize(molecule: Molecule, params: Optional[Pm3Parameters] = None, scf_options: Optional[Pm3Options] = None,
             opt: Optional[OptOptions] = None) -> OptResult:
    """optimize (src/optimizer.rs:233-262) on the GPU: L-BFGS to max |g| < gtol; raises like the
    reference when the first evaluation fails."""
    params = params or Pm3Parameters.standard()
    scf_options = scf_options or Pm3Options()
    opt = opt or OptOptions()
    cm = molecule._c()
    res = _COptResult()
    out = np.zeros((len(molecule.z), 3))
    cap = max(int(opt.max_iter), 0) + 1
    steps = (_COptStep * cap)()
    traj = np.zeros(3 * cap * len(molecule.z))
    _raise(native().pm3_optimize(ctypes.byref(cm), params._h, ctypes.byref(scf_options._c()), ctypes.byref(opt._c()),
                                 ctypes.byref(res), _dp(out), steps, _dp(traj)))
    return _opt_result(molecule, res, out, steps, traj)


@dataclasses.dataclass
class DcOptResult:
    """DcOptResult (src/optimizer.rs:313-320): OptResult's fields with the divide-and-conquer scalars of
    the last accepted point (scf: a DcResult without arrays: density, spin_density and charges are
    None; run_dc at the final geometry gives them)."""
    molecule: Molecule
    scf: "DcResult"
    converged: bool
    iterations: int
    evaluations: int
    energy_ev: float
    heat_of_formation_kcal: float
    max_gradient: float
    first_energy_ev: float
    trajectory: List[OptStep] = dataclasses.field(default_factory=list)


def optimize_dc(molecule: Molecule, params: Optional[Pm3Parameters] = None, scf_options: Optional[Pm3Options] = None,
                dc: Optional[DcOptions] = None, opt: Optional[OptOptions] = None) -> DcOptResult:
    """optimize_dc (src/optimizer.rs:277-310) on the GPU: the L-BFGS of optimize over dc_gradient;
    raises like the reference when the first evaluation fails."""
    params = params or Pm3Parameters.standard()
    scf_options = scf_options or Pm3Options()
    dc = dc or DcOptions()
    opt = opt or OptOptions()
    cm = molecule._c()
    res = _COptResult()
    sres = _CDcResult()
    out = np.zeros((len(molecule.z), 3))
    cap = max(int(opt.max_iter), 0) + 1
    steps = (_COptStep * cap)()
    traj = np.zeros(3 * cap * len(molecule.z))
    st = native().pm3_optimize_dc(ctypes.byref(cm), params._h, ctypes.byref(scf_options._c()), ctypes.byref(dc._c()),
                                  ctypes.byref(opt._c()), ctypes.byref(res), ctypes.byref(sres), _dp(out), steps,
                                  _dp(traj))
    if st == 3:
        raise ScfNotConverged(native().pm3_last_error_message().decode(), sres.iterations, float("nan"))
    _raise(st)
    o = _opt_result(molecule, res, out, steps, traj)
    scf = DcResult(None, None, bool(sres.unrestricted), sres.electronic_ev, sres.core_ev, sres.total_ev,
                   sres.heat_of_formation_kcal, None, sres.fermi_ev, sres.fermi_beta_ev, sres.iterations,
                   bool(sres.converged), sres.dropped_pairs, sres.largest_subsystem, sres.n_subsystems)
    return DcOptResult(o.molecule, scf, o.converged, o.iterations, o.evaluations, o.energy_ev,
                       o.heat_of_formation_kcal, o.max_gradient, o.first_energy_ev, o.trajectory)


def _rust_debug_f64(x):
    # the shortest round-trip digits (Python's repr), exponent written as Rust's Debug writes it
    t = repr(float(x))
    if "e" in t:
        mant, ex = t.split("e")
        t = mant + "e" + str(int(ex))
    return t


def energy_report(result: Pm3Result) -> str:
    """The lines of the reference CLI's `energy` command (src/cli.rs:1358-1371), which its oracle
    scripts parse (`^Total energy:`, `^MO energies \\(eV\\):\\s*\\[(.*)\\]`)."""
    mo = [] if result.mo_energies is None else list(result.mo_energies)
    return (f"Total energy:          {result.total_ev:.12f} eV\n"
            f"Electronic energy:     {result.electronic_ev:.12f} eV\n"
            f"Core-core energy:      {result.core_ev:.12f} eV\n"
            f"Heat of formation:     {result.heat_of_formation_kcal:.8f} kcal/mol\n"
            f"SCF iterations:        {result.iterations}\n"
            f"MO energies (eV):       [{', '.join(_rust_debug_f64(v) for v in mo)}]\n")
