"""ctypes wrapper of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
The product package never imports this module.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
EV_TO_KCAL = 23.06054783061903

STATUS = {0: "ok", 1: "invalid input", 2: "missing element", 3: "SCF not converged", 4: "unsupported",
          5: "resource limit"}


class Options(ctypes.Structure):
    _fields_ = [
        ("max_scf", ctypes.c_int),
        ("e_tol", ctypes.c_double),
        ("p_tol", ctypes.c_double),
        ("use_diis", ctypes.c_int),
        ("accelerator", ctypes.c_int),
        ("adiis_switch", ctypes.c_double),
        ("guess", ctypes.c_int),
        ("stability", ctypes.c_int),
        ("scf_memory_mb", ctypes.c_int),
        ("level_shift_ev", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("reference", ctypes.c_int),
        ("has_field", ctypes.c_int),
        ("field", ctypes.c_double * 3),
        ("integral_memory_mb", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("total_ev", ctypes.c_double),
        ("electronic_ev", ctypes.c_double),
        ("core_ev", ctypes.c_double),
        ("hof_kcal", ctypes.c_double),
        ("dipole", ctypes.c_double * 3),
        ("dipole_debye", ctypes.c_double),
        ("homo_ev", ctypes.c_double),
        ("lumo_ev", ctypes.c_double),
        ("last_change", ctypes.c_double),
        ("scf_improvement_ev", ctypes.c_double),
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("scf_paths_tried", ctypes.c_int),
        ("n_occ", ctypes.c_int),
        ("nao", ctypes.c_int),
        ("has_homo", ctypes.c_int),
        ("has_lumo", ctypes.c_int),
        ("unrestricted", ctypes.c_int),
        ("n_beta", ctypes.c_int),
        ("required_mb", ctypes.c_double),
        ("limit_mb", ctypes.c_double),
    ]


REFERENCE = {"auto": 0, "rhf": 1, "uhf": 2}


def options(max_scf=200, e_tol=1e-8, p_tol=1e-7, use_diis=True, accelerator=2, adiis_switch=0.1, guess=0,
            stability=1, scf_memory_mb=512, level_shift_ev=0.0, damping=0.0, reference=0, field=None,
            integral_memory_mb=0):
    """Mirror of the reference's Pm3Options defaults (src/scf.rs:259-287).  reference: 0 auto,
    1 RHF, 2 UHF (or the names); field: None or three components (eV / Bohr / e)."""
    if isinstance(reference, str):
        reference = REFERENCE[reference]
    f = (ctypes.c_double * 3)(*(field if field is not None else (0.0, 0.0, 0.0)))
    return Options(max_scf, e_tol, p_tol, int(use_diis), accelerator, adiis_switch, guess, stability, scf_memory_mb,
                   float(level_shift_ev), float(damping), int(reference), int(field is not None), f,
                   int(integral_memory_mb))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_element.argtypes = [ctypes.c_int, dp]
        L.orc_pair_integrals.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]
        L.orc_nao.argtypes = [ctypes.c_int, ip]
        L.orc_hcore.argtypes = [ctypes.c_int, ip, dp, dp]
        L.orc_fock.argtypes = [ctypes.c_int, ip, dp, dp, dp]
        L.orc_fock_fourindex.argtypes = [ctypes.c_int, ip, dp, dp, dp]
        L.orc_response_fock.argtypes = [ctypes.c_int, ip, dp, ctypes.c_int, dp, dp, dp]
        L.orc_core_core.argtypes = [ctypes.c_int, ip, dp]
        L.orc_core_core.restype = ctypes.c_double
        L.orc_eigen.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_run.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.POINTER(Options), ctypes.POINTER(Result),
                              dp, dp, dp, dp]
        L.orc_run_from_density.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.POINTER(Options), dp,
                                           ctypes.POINTER(Result), dp]
        L.orc_fixed_density_gradient.argtypes = [ctypes.c_int, ip, dp, dp, dp]
        L.orc_gradient.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.POINTER(Options),
                                   ctypes.POINTER(Result), dp, dp]
        L.orc_run_batch.argtypes = [ctypes.c_int, ip, ip, dp, dp, ctypes.POINTER(Options), ctypes.c_int, dp, dp, ip,
                                    ip]
        L.orc_gradient_batch.argtypes = [ctypes.c_int, ip, ip, dp, dp, ctypes.POINTER(Options), ctypes.c_int, dp,
                                         dp, ip, ip, dp]
        L.orc_run_pm3.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Options),
                                  ctypes.POINTER(Result), dp, dp, dp, dp, dp, dp, dp]
        L.orc_fixed_density_gradient_spin.argtypes = [ctypes.c_int, ip, dp, dp, dp, dp]
        L.orc_gradient_pm3.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Options),
                                       ctypes.POINTER(Result), dp, dp]
        L.orc_run_uhf.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Options),
                                  ctypes.POINTER(Result), dp, dp, dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_run_dc.argtypes = [ctypes.c_int, ip, dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Options), dp,
                                 ctypes.POINTER(DcResult), dp, dp, dp]
        L.orc_dc_partition.argtypes = [ctypes.c_int, ip, dp, dp, ip, ip, dp, ctypes.c_int, ip, ip, ip]
        L.orc_density_diis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, ip]
        fp = ctypes.POINTER(ctypes.c_float)
        up = ctypes.POINTER(ctypes.c_uint16)
        L.orc_guess_pairs.argtypes = [ctypes.c_int, ip, dp, ip, ip]
        L.orc_guess_features.argtypes = [ctypes.c_int, ip, dp, dp, up]
        L.orc_guess_forward.argtypes = [fp, ctypes.c_int, up, dp, ctypes.c_int, dp]
        L.orc_guess_backward.argtypes = [fp, ctypes.c_int, up, dp, dp, ctypes.c_int, dp]
        L.orc_guess_density.argtypes = [ctypes.c_int, ip, dp, fp, dp]
        L.orc_guess_loss.argtypes = [ctypes.c_int, ip, dp, fp, dp, dp]
        L.orc_guess_loss.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _mol(z, xyz_bohr):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    return z, xyz


def element(zz):
    out = np.zeros(32)
    st = lib().orc_element(zz, _d(out))
    if st:
        raise KeyError(zz)
    keys = ["n_orb", "dd", "qq", "rho0", "rho1", "rho2", "po1", "po2", "po3", "po7", "po9", "ddp2", "ddp3", "e_isol",
            "eheat_ev", "tore", "mass", "hsp"]
    return dict(zip(keys, out[: len(keys)]))


def pair_integrals(za, zb, d_bohr):
    d = np.ascontiguousarray(d_bohr, dtype=np.float64)
    w = np.zeros(100)
    e1b = np.zeros(16)
    e2a = np.zeros(16)
    s = np.zeros(16)
    st = lib().orc_pair_integrals(za, zb, _d(d), _d(w), _d(e1b), _d(e2a), _d(s))
    if st:
        raise RuntimeError(STATUS[st])
    return w, e1b.reshape(4, 4), e2a.reshape(4, 4), s.reshape(4, 4)


def nao(z):
    z = np.ascontiguousarray(z, dtype=np.int32)
    return lib().orc_nao(len(z), _i(z))


def hcore(z, xyz_bohr):
    z, xyz = _mol(z, xyz_bohr)
    n = nao(z)
    h = np.zeros((n, n))
    st = lib().orc_hcore(len(z), _i(z), _d(xyz), _d(h))
    if st:
        raise RuntimeError(STATUS[st])
    return h


def fock(z, xyz_bohr, P, four_index=False):
    z, xyz = _mol(z, xyz_bohr)
    P = np.ascontiguousarray(P, dtype=np.float64)
    F = np.zeros_like(P)
    fn = lib().orc_fock_fourindex if four_index else lib().orc_fock
    st = fn(len(z), _i(z), _d(xyz), _d(P), _d(F))
    if st:
        raise RuntimeError(STATUS[st])
    return F


def response_fock(z, xyz_bohr, dp, pair_sw=None):
    """response_fock_rhf (src/fock.rs:444-485) for one density (n, n) or a stack (k, n, n); pair_sw
    one weight per atom pair in the reference's order (u < v, row by row), or None."""
    z, xyz = _mol(z, xyz_bohr)
    dp = np.ascontiguousarray(dp, dtype=np.float64)
    one = dp.ndim == 2
    stack = dp.reshape(1, *dp.shape) if one else dp
    out = np.zeros_like(stack)
    sw = None if pair_sw is None else np.ascontiguousarray(pair_sw, dtype=np.float64)
    st = lib().orc_response_fock(len(z), _i(z), _d(xyz), stack.shape[0], _d(stack),
                                 _d(sw) if sw is not None else ctypes.POINTER(ctypes.c_double)(), _d(out))
    if st:
        raise RuntimeError(STATUS[st])
    return out[0] if one else out


def core_core(z, xyz_bohr):
    z, xyz = _mol(z, xyz_bohr)
    return lib().orc_core_core(len(z), _i(z), _d(xyz))


def eigen(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_eigen(n, _d(A), _d(w), _d(V))
    return w, V


def run(z, xyz_bohr, charge=0.0, opts=None, want_arrays=True):
    """RHF PM3 single point; returns (status, dict)."""
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    n = nao(z)
    res = Result()
    P = np.zeros((n, n)) if want_arrays else None
    eps = np.zeros(n) if want_arrays else None
    C = np.zeros((n, n)) if want_arrays else None
    q = np.zeros(len(z)) if want_arrays else None
    nullp = ctypes.POINTER(ctypes.c_double)()
    st = lib().orc_run(len(z), _i(z), _d(xyz), float(charge), ctypes.byref(opts), ctypes.byref(res),
                       _d(P) if want_arrays else nullp, _d(eps) if want_arrays else nullp,
                       _d(C) if want_arrays else nullp, _d(q) if want_arrays else nullp)
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out["dipole"] = np.array(res.dipole[:])
    if want_arrays:
        out.update(density=P, mo_energies=eps, mo_coeff=C, charges=q)
    return st, out


class DcResult(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in ("total_ev", "electronic_ev", "core_ev", "hof_kcal", "fermi_ev",
                                                "fermi_beta_ev")] + \
               [(f, ctypes.c_int) for f in ("iterations", "converged", "unrestricted", "dropped_pairs",
                                             "largest_subsystem", "n_subsystems")]


def dc_options(core_radius=6.0, buffer_radius=9.0, min_core_atoms=1, electron_tol=1.0e-9, smearing_ev=0.1):
    """DcOptions (src/dc/partition.rs:37-107) as the oracle's five numbers; defaults :96-107."""
    return np.array([core_radius, buffer_radius, float(min_core_atoms), electron_tol, smearing_ev])


def run_dc(z, xyz_bohr, charge=0.0, multiplicity=1, opts=None, dc=None):
    """run_dc (src/dc/scf.rs:76-160): (status, dict of DcResult fields with density, spin_density
    and charges)."""
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    dc = dc_options() if dc is None else np.ascontiguousarray(dc, dtype=np.float64)
    n = nao(z)
    res = DcResult()
    P = np.zeros((n, n))
    S = np.zeros((n, n))
    q = np.zeros(len(z))
    st = lib().orc_run_dc(len(z), _i(z), _d(xyz), float(charge), int(multiplicity), ctypes.byref(opts), _d(dc),
                          ctypes.byref(res), _d(P), _d(S), _d(q))
    out = {f: getattr(res, f) for f, _ in DcResult._fields_}
    out.update(density=P, spin_density=S, charges=q)
    return st, out


def dc_gradient(z, xyz_bohr, charge=0.0, multiplicity=1, opts=None, dc=None):
    """dc_gradient (src/dc/derivatives.rs:55-121): the fixed-density gradient at the partitioned
    density (the unrestricted form for an unrestricted result), minus f q_A in a field."""
    opts = opts or options()
    st, out = run_dc(z, xyz_bohr, charge, multiplicity, opts, dc)
    if st:
        return st, out
    if out["unrestricted"]:
        g = fixed_density_gradient_spin(z, xyz_bohr, out["density"], out["spin_density"])
    else:
        g = fixed_density_gradient(z, xyz_bohr, out["density"])
    if opts.has_field:
        g = g - np.outer(out["charges"], np.array(opts.field[:]))
    out["gradient"] = g
    return st, out


def dc_partition(z, xyz_bohr, dc=None, want_weights=False):
    """partition (src/dc/partition.rs:184-275): core_of, subsystem count, largest subsystem,
    dropped pairs, pattern entries, the orbital lists with their core flags and, on request, the
    summed weights (Partition::weight_sum)."""
    z, xyz = _mol(z, xyz_bohr)
    dc = dc_options() if dc is None else np.ascontiguousarray(dc, dtype=np.float64)
    n = nao(z)
    core_of = np.zeros(len(z), np.int32)
    stats = np.zeros(4, np.int32)
    ws = np.zeros((n, n)) if want_weights else None
    cap = n * len(z) + 1
    ptr = np.zeros(cap + 1, np.int32)
    orb = np.zeros(cap, np.int32)
    inc = np.zeros(cap, np.int32)
    st = lib().orc_dc_partition(len(z), _i(z), _d(xyz), _d(dc), _i(core_of), _i(stats),
                                _d(ws) if ws is not None else None, cap, _i(ptr), _i(orb), _i(inc))
    if st:
        raise RuntimeError(STATUS[st])
    ns = int(stats[0])
    subs = [(orb[ptr[s]:ptr[s + 1]].copy(), inc[ptr[s]:ptr[s + 1]].astype(bool)) for s in range(ns)]
    return dict(core_of=core_of, n_subsystems=ns, largest_subsystem=int(stats[1]), dropped_pairs=int(stats[2]),
                nnz=int(stats[3]), subsystems=subs, weight_sum=ws)


def density_diis(dens, res, memory_mb=0):
    """DensityDiis (src/densitydiis.rs:34-119): push every (density, residual) row, then
    extrapolate; returns (extrapolation or None, depth)."""
    dens = np.ascontiguousarray(dens, dtype=np.float64)
    res = np.ascontiguousarray(res, dtype=np.float64)
    nvec, ln = dens.shape
    out = np.zeros(ln)
    depth = np.zeros(1, np.int32)
    ok = lib().orc_density_diis(ln, int(memory_mb), nvec, _d(dens), _d(res), _d(out), _i(depth))
    return (out if ok else None), int(depth[0])


def run_pm3(z, xyz_bohr, charge=0.0, multiplicity=1, opts=None):
    """run_pm3 with every output of Pm3Result: (status, dict with density, spin_density,
    mo_energies, mo_coeff, mo_energies_beta, mo_coeff_beta, charges)."""
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    n = nao(z)
    res = Result()
    arr = {k: np.zeros((n, n)) for k in ("density", "spin_density", "mo_coeff", "mo_coeff_beta")}
    arr.update(mo_energies=np.zeros(n), mo_energies_beta=np.zeros(n), charges=np.zeros(len(z)))
    st = lib().orc_run_pm3(len(z), _i(z), _d(xyz), float(charge), int(multiplicity), ctypes.byref(opts),
                           ctypes.byref(res), _d(arr["density"]), _d(arr["spin_density"]), _d(arr["mo_energies"]),
                           _d(arr["mo_coeff"]), _d(arr["mo_energies_beta"]), _d(arr["mo_coeff_beta"]),
                           _d(arr["charges"]))
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out["dipole"] = np.array(res.dipole[:])
    out.update(arr)
    return st, out


def run_uhf(z, xyz_bohr, charge=0.0, multiplicity=1, opts=None):
    """run_pm3 with a multiplicity (UHF for unequal spin counts); returns (status, dict with the
    total density, spin density and charges)."""
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    n = nao(z)
    res = Result()
    P = np.zeros((n, n))
    S = np.zeros((n, n))
    q = np.zeros(len(z))
    st = lib().orc_run_uhf(len(z), _i(z), _d(xyz), float(charge), int(multiplicity), ctypes.byref(opts),
                           ctypes.byref(res), _d(P), _d(S), _d(q))
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out["dipole"] = np.array(res.dipole[:])
    out.update(density=P, spin_density=S, charges=q)
    return st, out


def run_from_density(z, xyz_bohr, P0, charge=0.0, opts=None):
    """RHF PM3 single point started from the density P0; returns (status, dict with density)."""
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    n = nao(z)
    P0 = np.ascontiguousarray(P0, dtype=np.float64).reshape(n, n)
    res = Result()
    P = np.zeros((n, n))
    st = lib().orc_run_from_density(len(z), _i(z), _d(xyz), float(charge), ctypes.byref(opts), _d(P0),
                                    ctypes.byref(res), _d(P))
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out["dipole"] = np.array(res.dipole[:])
    out["density"] = P
    return st, out


def fixed_density_gradient(z, xyz_bohr, P):
    z, xyz = _mol(z, xyz_bohr)
    P = np.ascontiguousarray(P, dtype=np.float64)
    g = np.zeros((len(z), 3))
    st = lib().orc_fixed_density_gradient(len(z), _i(z), _d(xyz), _d(P), _d(g))
    if st:
        raise RuntimeError(STATUS[st])
    return g


def fixed_density_gradient_spin(z, xyz_bohr, P, S):
    z, xyz = _mol(z, xyz_bohr)
    P = np.ascontiguousarray(P, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    g = np.zeros((len(z), 3))
    st = lib().orc_fixed_density_gradient_spin(len(z), _i(z), _d(xyz), _d(P), _d(S), _d(g))
    if st:
        raise RuntimeError(STATUS[st])
    return g


def gradient(z, xyz_bohr, charge=0.0, opts=None, multiplicity=1):
    z, xyz = _mol(z, xyz_bohr)
    opts = opts or options()
    n = nao(z)
    res = Result()
    g = np.zeros((len(z), 3))
    P = np.zeros((n, n))
    st = lib().orc_gradient_pm3(len(z), _i(z), _d(xyz), float(charge), int(multiplicity), ctypes.byref(opts),
                                ctypes.byref(res), _d(g), _d(P))
    out = {f: getattr(res, f) for f, _ in Result._fields_}
    out.update(gradient=g, density=P)
    return st, out


def run_batch(mols, opts=None, nthreads=0):
    """Independent single points over host threads; mols = list of (z, xyz_bohr)."""
    opts = opts or options()
    ptr = np.zeros(len(mols) + 1, dtype=np.int32)
    for k, (z, _) in enumerate(mols):
        ptr[k + 1] = ptr[k] + len(z)
    zz = np.ascontiguousarray(np.concatenate([np.asarray(z, dtype=np.int32) for z, _ in mols]))
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3) for _, x in mols]))
    k = len(mols)
    tot = np.zeros(k)
    hof = np.zeros(k)
    it = np.zeros(k, dtype=np.int32)
    st = np.zeros(k, dtype=np.int32)
    nullp = ctypes.POINTER(ctypes.c_double)()
    lib().orc_run_batch(k, _i(ptr), _i(zz), _d(xyz), nullp, ctypes.byref(opts), int(nthreads), _d(tot), _d(hof),
                        _i(it), _i(st))
    return tot, hof, it, st


def numerical_hessian(z, xyz_bohr, charge=0.0, opts=None, step=1.0e-3, multiplicity=1):
    """numerical_hessian (src/hessian.rs:268-322): column j = (g(x + h e_j) - g(x - h e_j)) / 2h of the
    analytic gradient under tighten_scf_for_hessian (:169-175: e_tol <= 1e-12, p_tol <= 1e-11,
    max_scf >= 500), then H = (H + H^T) / 2 (eV/Bohr^2)."""
    z = np.asarray(z, dtype=np.int32)
    x0 = np.array(xyz_bohr, dtype=np.float64).reshape(-1)
    o = opts or options()
    o = options(max_scf=max(o.max_scf, 500), e_tol=min(o.e_tol, 1e-12), p_tol=min(o.p_tol, 1e-11),
                use_diis=bool(o.use_diis), accelerator=o.accelerator, adiis_switch=o.adiis_switch, guess=o.guess,
                stability=o.stability, scf_memory_mb=o.scf_memory_mb)
    n = x0.size
    h = np.zeros((n, n))
    for j in range(n):
        cols = []
        for sgn in (1.0, -1.0):
            x = x0.copy()
            x[j] += sgn * step
            st, g = gradient(z, x.reshape(-1, 3), charge, o, multiplicity)
            if st != 0:
                raise RuntimeError(STATUS[st])
            cols.append(g["gradient"].reshape(-1))
        h[:, j] = (cols[0] - cols[1]) / (2.0 * step)
    return 0.5 * (h + h.T)


def gradient_batch(mols, opts=None, nthreads=0):
    """run_batch with the gradient of each molecule too; returns (tot, hof, it, st, [grad (nat, 3)])."""
    opts = opts or options()
    ptr = np.zeros(len(mols) + 1, dtype=np.int32)
    for k, (z, _) in enumerate(mols):
        ptr[k + 1] = ptr[k] + len(z)
    zz = np.ascontiguousarray(np.concatenate([np.asarray(z, dtype=np.int32) for z, _ in mols]))
    xyz = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3) for _, x in mols]))
    k = len(mols)
    tot = np.zeros(k)
    hof = np.zeros(k)
    it = np.zeros(k, dtype=np.int32)
    st = np.zeros(k, dtype=np.int32)
    g = np.zeros((int(ptr[-1]), 3))
    nullp = ctypes.POINTER(ctypes.c_double)()
    lib().orc_gradient_batch(k, _i(ptr), _i(zz), _d(xyz), nullp, ctypes.byref(opts), int(nthreads), _d(tot), _d(hof),
                             _i(it), _i(st), _d(g))
    return tot, hof, it, st, [g[ptr[m]:ptr[m + 1]] for m in range(k)]


def num_threads():
    return lib().orc_num_threads()


# ---------------------------------------------------------------------------------------------
# Learned initial density (guess_oracle.cpp).

GUESS_NF, GUESS_NH, GUESS_NO = 96, 128, 16
GUESS_NPARAM = GUESS_NF * GUESS_NH + GUESS_NH * GUESS_NH + GUESS_NH + GUESS_NH * GUESS_NO + GUESS_NO


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u16(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def guess_pairs(z, xyz_bohr):
    """Directed pairs (src, dst) inside 4 Angstrom, ordered by src then dst."""
    z, xyz = _mol(z, xyz_bohr)
    n = lib().orc_guess_pairs(len(z), _i(z), _d(xyz), None, None)
    src = np.zeros(n, dtype=np.int32)
    dst = np.zeros(n, dtype=np.int32)
    lib().orc_guess_pairs(len(z), _i(z), _d(xyz), _i(src), _i(dst))
    return src, dst


def guess_features(z, xyz_bohr):
    """(f double before rounding, x bf16 bits) per directed pair, GUESS_NF each."""
    z, xyz = _mol(z, xyz_bohr)
    src, _ = guess_pairs(z, xyz)
    f = np.zeros((len(src), GUESS_NF))
    x = np.zeros((len(src), GUESS_NF), dtype=np.uint16)
    st = lib().orc_guess_features(len(z), _i(z), _d(xyz), _d(f), _u16(x))
    if st:
        raise RuntimeError(STATUS[st])
    return f, x


def guess_forward(params, x=None, xd=None, rounded=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    n = len(x) if x is not None else len(xd)
    y = np.zeros((n, GUESS_NO))
    xa = np.ascontiguousarray(x, dtype=np.uint16) if x is not None else None
    xda = np.ascontiguousarray(xd, dtype=np.float64) if xd is not None else None
    lib().orc_guess_forward(_f(params), n, _u16(xa) if xa is not None else None,
                            _d(xda) if xda is not None else None, int(rounded), _d(y))
    return y


def guess_backward(params, dy, x=None, xd=None, rounded=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    n = len(dy)
    g = np.zeros(GUESS_NPARAM)
    xa = np.ascontiguousarray(x, dtype=np.uint16) if x is not None else None
    xda = np.ascontiguousarray(xd, dtype=np.float64) if xd is not None else None
    lib().orc_guess_backward(_f(params), n, _u16(xa) if xa is not None else None,
                             _d(xda) if xda is not None else None, _d(dy), int(rounded), _d(g))
    return g


def guess_density(z, xyz_bohr, y):
    z, xyz = _mol(z, xyz_bohr)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n = nao(z)
    P = np.zeros((n, n))
    st = lib().orc_guess_density(len(z), _i(z), _d(xyz), _f(y), _d(P))
    if st:
        raise RuntimeError(STATUS[st])
    return P


def guess_loss(z, xyz_bohr, y, T):
    """(loss, dloss/dy) of the assembled guess against the target density T."""
    z, xyz = _mol(z, xyz_bohr)
    y = np.ascontiguousarray(y, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float64)
    dy = np.zeros((len(y), GUESS_NO))
    loss = lib().orc_guess_loss(len(z), _i(z), _d(xyz), _f(y), _d(T), _d(dy))
    return loss, dy


def _dot(a, b):
    # the reference's iterator sum: products added in index order
    s = 0.0
    for u, v in zip(a.tolist(), b.tolist()):
        s += u * v
    return s


def _drive(x, evaluate, energy_only, max_iter, gtol, history):
    """drive (src/optimizer.rs:77-221): L-BFGS with a backtracking Armijo search over an evaluator
    of (energy, gradient, heat of formation, max |g|) and a trial-energy function that returns
    None for a point it cannot evaluate (a rejected step)."""
    energy, g, heat, maxg = evaluate(x)
    evaluations = 1
    s_hist, y_hist, rho_hist = [], [], []
    trajectory = [energy]
    steps = [{"energy_ev": energy, "hof_kcal": heat, "max_gradient": maxg, "xyz": x.reshape(-1, 3).copy()}]
    converged = maxg < gtol
    iterations = 0
    for it in range(max_iter):
        iterations = it + 1
        if converged:
            break
        q = g.copy()
        m = len(s_hist)
        alpha = [0.0] * m
        for i in range(m - 1, -1, -1):
            a = rho_hist[i] * _dot(s_hist[i], q)
            alpha[i] = a
            q += (-a) * y_hist[i]
        if m > 0:
            sy = _dot(s_hist[-1], y_hist[-1])
            yy = _dot(y_hist[-1], y_hist[-1])
            gamma = sy / yy if yy > 0.0 else 1.0
        else:
            gamma = 0.1 / max(maxg, 1.0e-6)
        q *= gamma
        for i in range(m):
            beta = rho_hist[i] * _dot(y_hist[i], q)
            q += (alpha[i] - beta) * s_hist[i]
        d = -q
        if _dot(d, g) > 0.0:
            d = -g
        gdd = _dot(g, d)
        step, ok = 1.0, False
        while True:
            x_new = x + step * d
            trial = energy_only(x_new)
            if trial is not None and trial <= energy + 1.0e-4 * step * gdd:
                ok = True
                break
            step *= 0.5
            if step < 1.0e-8:
                break
        if not ok:
            break
        e_new, g_new, h_new, m_new = evaluate(x_new)
        evaluations += 1
        s = x_new - x
        y = g_new - g
        sy = _dot(s, y)
        if sy > 1.0e-10:
            s_hist.append(s)
            y_hist.append(y)
            rho_hist.append(1.0 / sy)
            if len(s_hist) > history:
                s_hist.pop(0)
                y_hist.pop(0)
                rho_hist.pop(0)
        x, g, energy, heat, maxg = x_new, g_new, e_new, h_new, m_new
        converged = maxg < gtol
        trajectory.append(energy)
        steps.append({"energy_ev": energy, "hof_kcal": heat, "max_gradient": maxg, "xyz": x.reshape(-1, 3).copy()})
    return {"converged": converged, "iterations": iterations, "evaluations": evaluations, "energy_ev": energy,
            "hof_kcal": heat, "max_gradient": maxg, "xyz": x.reshape(-1, 3), "trajectory": trajectory,
            "steps": steps}


def optimize(z, xyz_bohr, charge=0.0, opts=None, max_iter=200, gtol=1.0e-3, history=8, multiplicity=1):
    """optimize / drive (src/optimizer.rs:77-262): L-BFGS on the PM3 surface with a backtracking
    Armijo search.  Trials ask for the energy alone (run_pm3's total_ev, src/optimizer.rs:254),
    accepted points are evaluated again for the gradient (closed_form_gradient, :244-252).
    Returns a dict: converged, iterations, evaluations (gradient calls), energy_ev, hof_kcal,
    max_gradient, xyz (final, Bohr), trajectory (energies of the accepted points, first included),
    steps (OptStep, :37-43, of the same points: energy_ev, hof_kcal, max_gradient, xyz)."""
    z = np.asarray(z, dtype=np.int32)
    opts = opts or options()
    x = np.array(xyz_bohr, dtype=np.float64).reshape(-1)

    def evaluate(xv):
        st, o = gradient(z, xv.reshape(-1, 3), charge, opts, multiplicity)
        if st != 0:
            raise RuntimeError(STATUS[st])
        g = o["gradient"].reshape(-1).copy()
        return o["total_ev"], g, o["hof_kcal"], float(np.abs(g).max()) if g.size else 0.0

    def energy_only(xv):
        if multiplicity == 1:
            st, o = run(z, xv.reshape(-1, 3), charge, opts, want_arrays=False)
        else:
            st, o = run_pm3(z, xv.reshape(-1, 3), charge, multiplicity, opts)
        return o["total_ev"] if st == 0 else None

    return _drive(x, evaluate, energy_only, max_iter, gtol, history)


def optimize_dc(z, xyz_bohr, charge=0.0, multiplicity=1, opts=None, dc=None, max_iter=200, gtol=1.0e-3, history=8):
    """optimize_dc (src/optimizer.rs:277-310): the same drive over dc_gradient, trial energies from
    run_dc's total_ev (any failure a rejected trial).  Returns optimize's dict."""
    z = np.asarray(z, dtype=np.int32)
    opts = opts or options()
    x = np.array(xyz_bohr, dtype=np.float64).reshape(-1)

    def evaluate(xv):
        st, o = dc_gradient(z, xv.reshape(-1, 3), charge, multiplicity, opts, dc)
        if st != 0:
            raise RuntimeError(STATUS[st])
        g = o["gradient"].reshape(-1).copy()
        return o["total_ev"], g, o["hof_kcal"], float(np.abs(g).max()) if g.size else 0.0

    def energy_only(xv):
        st, o = run_dc(z, xv.reshape(-1, 3), charge, multiplicity, opts, dc)
        return o["total_ev"] if st == 0 else None

    return _drive(x, evaluate, energy_only, max_iter, gtol, history)
