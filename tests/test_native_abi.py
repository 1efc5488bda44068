"""The C-ABI library loads and exports every entry point the headers declare (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"\b(pm3_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    import pm3amd
    lib = pm3amd.native()
    names = declared_symbols()
    assert "pm3_run" in names and "pm3_gradient" in names and "pm3_run_batch" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(pm3amd.EXPORTED) <= names


def test_params_and_defaults_without_device():
    import pm3amd as G
    p = G.Pm3Parameters.standard()
    o = G._COptions()
    G.native().pm3_options_default(ctypes.byref(o))
    # Pm3Options::default(), src/scf.rs:259-287
    assert (o.max_scf, o.e_tol, o.p_tol, o.use_diis, o.accelerator, o.adiis_switch, o.scf_memory_mb) == \
        (200, 1e-8, 1e-7, 1, 2, 0.1, 512)
    m = G.Molecule(np.array([8, 1, 1]), np.zeros((3, 3)))
    assert G.native().pm3_query_nao(ctypes.byref(m._c()), p._h) == 6
    m2 = G.Molecule(np.array([6, 18]), np.zeros((2, 3)))
    assert G.native().pm3_query_nao(ctypes.byref(m2._c()), p._h) == -1


def test_host_side_errors_need_no_device():
    import pm3amd as G
    # RHF on an open shell is refused before any device work (src/scf.rs:512-519), with the
    # reference's message; so is a multiplicity below 1
    m = G.Molecule(np.array([6, 1, 1, 1]), np.eye(4, 3) * 2.0, multiplicity=2)
    with pytest.raises(G.InvalidInput, match="RHF requested"):
        G.run_pm3(m, options=G.Pm3Options(reference="rhf"))
    with pytest.raises(G.InvalidInput):
        G.run_pm3(G.Molecule(m.z, m.positions, multiplicity=0))
    with pytest.raises(G.InvalidInput, match="damping"):
        G.run_pm3(m, options=G.Pm3Options(damping=1.0))
    with pytest.raises(G.InvalidInput):
        G.Molecule.from_xyz_str("2\nx\nH 0 0 0\n")
    with pytest.raises(G.InvalidInput):
        G.Molecule.from_xyz_str("1\nx\nXx 0 0 0\n")


def test_parameter_table_matches_reference_csv():
    """The embedded rows equal the reference's CSV values (checked where the reference exists)."""
    ref = "/root/reference/src/data/pm3_parameters.csv"
    if not os.path.exists(ref):
        pytest.skip("reference tree not present")
    import csv
    rows = {int(r["z"]): r for r in csv.DictReader(l for l in open(ref) if not l.startswith("#"))}
    inc = open(os.path.join(ROOT, "pm3-rs_amd", "csrc", "pm3_table.inc")).read()
    keys = ["uss", "upp", "zs", "zp", "betas", "betap", "gss", "gsp", "gpp", "gp2", "hsp", "alp",
            "g1_k", "g1_l", "g1_m", "g2_k", "g2_l", "g2_m"]
    for m in re.finditer(r"PM3_ELEMENT\(([^)]*)\)", inc):
        vals = [float(v) for v in m.group(1).split(",")]
        z = int(vals[0])
        for k, v in zip(keys, vals[1:]):
            assert float(rows[z][k]) == v, (z, k)


def test_tuning_table_round_trip_and_validation(tmp_path):
    """The tuning table is host state: set/get/save/load and range checks need no device."""
    import pm3amd as G
    d = G.tuning_default()
    assert d == {"pair_block": 256, "pair_blocks_per_cu": 8, "sp2_tile16_n": 128, "sp2_lds": 6, "sp2_tight_n": 48,
                 "large_syrk": 0, "large_split": 1, "wg128_nao": 80, "wg256_nao": 96, "solo_large_nao": 128}
    start = G.tuning_get()
    try:
        G.tuning_set({"pair_block": 128, "sp2_tile16_n": 96})
        assert G.tuning_get()["pair_block"] == 128 and G.tuning_get()["sp2_tile16_n"] == 96
        path = str(tmp_path / "t.txt")
        G.tuning_save(path)
        G.tuning_set(d)
        G.tuning_load(path)
        assert G.tuning_get() == dict(d, pair_block=128, sp2_tile16_n=96)
        for bad in ({"pair_block": 96}, {"pair_block": 512}, {"sp2_lds": 7}, {"pair_blocks_per_cu": 0}):
            with pytest.raises(G.InvalidInput):
                G.tuning_set(bad)
        assert G.tuning_get() == dict(d, pair_block=128, sp2_tile16_n=96)  # a rejected table changes nothing
        (tmp_path / "u.txt").write_text("pair_block 64\nno_such_key 1\n")
        with pytest.raises(G.InvalidInput):
            G.tuning_load(str(tmp_path / "u.txt"))
        (tmp_path / "v.txt").write_text("# comment\n  sp2_tight_n 96  # trailing\n\n")
        G.tuning_load(str(tmp_path / "v.txt"))
        assert G.tuning_get() == dict(d, sp2_tight_n=96)  # keys not listed take their defaults
    finally:
        G.tuning_set(start)


def test_tuning_file_from_environment(tmp_path):
    import subprocess
    import sys
    (tmp_path / "t.txt").write_text("pair_blocks_per_cu 4\nsp2_tile16_n 64\n")
    code = ("import sys; sys.path.insert(0, %r); import pm3amd as G; t = G.tuning_get(); "
            "print(t['pair_blocks_per_cu'], t['sp2_tile16_n'], t['pair_block'])" % os.path.join(ROOT, "pm3-rs_amd"))
    env = dict(os.environ, PM3_TUNING_FILE=str(tmp_path / "t.txt"))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.stdout.split() == ["4", "64", "256"], out.stderr


def test_autotune_rejects_bad_arguments_without_device():
    """The tuner validates its workload on the host first (no batch is created) and leaves the
    process-wide table as it was."""
    import pm3amd as G
    start = G.tuning_get()
    ip, dp = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    z = np.array([8, 1, 1], np.int32)
    xyz = np.zeros((3, 3))
    t, rep = G._CTuning(), G._CTuneReport()
    o = G._COptions()
    G.native().pm3_options_default(ctypes.byref(o))
    for n_mol, ptr in ((0, np.array([0], np.int32)), (1, np.array([1, 3], np.int32)), (2, np.array([0, 3, 2], np.int32))):
        st = G.native().pm3_autotune(n_mol, ptr.ctypes.data_as(ip), z.ctypes.data_as(ip), xyz.ctypes.data_as(dp), None,
                                      G.Pm3Parameters.standard()._h, ctypes.byref(o), 0, 1, ctypes.byref(t),
                                      ctypes.byref(rep))
        assert st == 1, (n_mol, ptr)  # PM3_INVALID_INPUT
    assert G.tuning_get() == start


def test_optimizer_defaults_and_host_checks_without_device():
    """OptOptions::default (src/optimizer.rs:26-35), and the optimiser's argument checks, which
    return before any batch is created."""
    import pm3amd as G
    L = G.native()
    o = G._COptOptions()
    L.pm3_opt_options_default(ctypes.byref(o))
    assert (o.max_iter, o.gtol, o.grad_step, o.history, o.warm_start) == (200, 1.0e-3, 5.0e-4, 8, 0)
    ip, dp = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    z = np.array([8, 1, 1], np.int32)
    xyz = np.zeros((3, 3))
    out = np.zeros((3, 3))
    res = (G._COptResult * 1)()
    params = G.Pm3Parameters.standard()._h
    so = G._COptions()
    L.pm3_options_default(ctypes.byref(so))

    def call(n_mol, ptr, opts=so, opt=o):
        return L.pm3_optimize_batch(n_mol, ptr.ctypes.data_as(ip), z.ctypes.data_as(ip), xyz.ctypes.data_as(dp), None,
                                    None, params, ctypes.byref(opts), ctypes.byref(opt), res, out.ctypes.data_as(dp),
                                    None, None)

    assert call(0, np.array([0], np.int32)) == 0  # nothing to do, no device work
    assert call(-1, np.array([0], np.int32)) == 1
    assert call(1, np.array([1, 3], np.int32)) == 1  # atom_ptr[0] != 0
    bad = G._COptOptions(-1, 1e-3, 5e-4, 8, 0)
    assert call(1, np.array([0, 3], np.int32), opt=bad) == 1
    learned = G._COptions()
    L.pm3_options_default(ctypes.byref(learned))
    learned.guess = 4  # PM3_GUESS_LEARNED without a guess model: refused before any device work
    assert call(1, np.array([0, 3], np.int32), opts=learned) == 1
    m = G.Molecule(z, np.eye(3), 0.0, 0)
    r = G._COptResult()
    assert L.pm3_optimize(ctypes.byref(m._c()), params, ctypes.byref(so), ctypes.byref(o), ctypes.byref(r),
                          out.ctypes.data_as(dp), None, None) == 1  # multiplicity below 1
    dco = G._CDcOptions()
    L.pm3_dc_options_default(ctypes.byref(dco))
    sres = G._CDcResult()
    assert L.pm3_optimize_dc(ctypes.byref(m._c()), params, ctypes.byref(so), ctypes.byref(dco), ctypes.byref(o),
                             ctypes.byref(r), ctypes.byref(sres), out.ctypes.data_as(dp), None, None) == 1
    m1 = G.Molecule(z, np.eye(3), 0.0, 1)
    assert L.pm3_optimize_dc(ctypes.byref(m1._c()), params, ctypes.byref(so), ctypes.byref(dco), ctypes.byref(bad),
                             ctypes.byref(r), ctypes.byref(sres), out.ctypes.data_as(dp), None, None) == 1
    assert L.pm3_optimize_dc(ctypes.byref(m1._c()), params, ctypes.byref(so), ctypes.byref(dco), ctypes.byref(o),
                             None, None, out.ctypes.data_as(dp), None, None) == 1  # no result struct
    assert L.pm3_batch_set_positions(None, xyz.ctypes.data_as(dp)) == 1
    assert L.pm3_batch_carry_densities(None, None, None) == 1
