"""Seeded geometry generators (sizes quoted for the benchmark workloads)."""
import numpy as np

from pm3amd import workloads as W


def nao(z):
    return int(np.sum(np.where(z == 1, 1, 4)))


def test_water_cluster_sizes_and_spacing():
    z, x = W.water_cluster(32, 10.0, 11)
    assert len(z) == 96 and nao(z) == 192 and int(np.sum(np.where(z == 1, 1, 6))) == 256
    o = x[z == 8]
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(32) * 99
    assert d.min() >= 2.7
    assert (o >= 0).all() and (o <= 10.0).all()
    # rigid geometry
    for k in range(32):
        oh1, oh2 = x[3 * k + 1] - x[3 * k], x[3 * k + 2] - x[3 * k]
        assert abs(np.linalg.norm(oh1) - 0.9572) < 1e-12 and abs(np.linalg.norm(oh2) - 0.9572) < 1e-12
        ang = np.degrees(np.arccos(oh1 @ oh2 / 0.9572 ** 2))
        assert abs(ang - 104.52) < 1e-9
    z2, x2 = W.water_cluster(32, 10.0, 11)
    assert np.array_equal(x, x2)


def test_large_water_cluster():
    z, x = W.water_cluster(512, 25.0, 5)
    assert len(z) == 1536 and nao(z) == 3072


def test_alkane_geometry():
    for n in (1, 2, 5, 24):
        z, x = W.alkane(n)
        assert (z == 6).sum() == n and (z == 1).sum() == 2 * n + 2
        d = np.linalg.norm(x[:, None] - x[None], axis=-1)
        c = np.where(z == 6)[0]
        h = np.where(z == 1)[0]
        for i in h:  # every H sits 1.09 A from exactly one carbon
            assert np.sum(np.abs(d[i, c] - 1.09) < 1e-9) == 1
        for i in range(n - 1):
            assert abs(d[c[i], c[i + 1]] - 1.54) < 1e-9
    assert nao(W.alkane(24)[0]) == 146


def test_mixed_batch_and_conformers():
    mols = W.mixed_batch(200, 1)
    sizes = [nao(z) for z, _ in mols]
    assert min(sizes) >= 8 and max(sizes) <= 146
    confs = W.solvated_conformers(4, 9)
    z, x = confs[0]
    assert len(z) == 98 and nao(z) == 194
    heavy = x[z != 1]
    # the unperturbed placement respects 2.6 A between heavy atoms of different molecules
    zb, xb = W.solvated_alkane(9)
    o = xb[zb == 8]
    hv = xb[zb != 1]
    for p in o:
        dd = np.linalg.norm(hv - p, axis=1)
        assert np.sum(dd < 1e-9) == 1 and np.sort(dd)[1] >= 2.6


def test_xyz_writer_round_trips_and_keeps_the_reference_layout():
    import pm3amd as G
    from pm3amd import workloads as W
    z, x = W.alkane(4)
    m = G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))
    text = m.to_xyz_str()
    lines = text.splitlines()
    assert lines[0] == str(len(z)) and lines[1] == "pm3-rs optimized geometry; coordinates in Angstrom"
    assert lines[2].startswith("C  ") and all(len(f.split(".")[1]) == 10 for f in lines[2].split()[1:])
    back = G.Molecule.from_xyz_str(text)
    assert np.array_equal(back.z, m.z)
    assert np.abs(back.positions - m.positions).max() < 1e-9  # ten decimals of Angstrom


def test_energy_report_reads_back_with_the_reference_scripts_patterns():
    import re
    import pm3amd as G
    mo = np.array([-40.123456789012345, -1.5, 0.0, 3.25e-07, 12.0])
    r = G.Pm3Result(None, mo, None, 2, -350.25, 123.5, -226.75, -57.123456789, np.zeros(3), np.zeros(3), 0.0,
                    -1.5, 0.0, 11, True, 1, 0.0)
    text = G.energy_report(r)
    # tools/oracle/internal_consistency.py:43-45 and all_element_validation.py's MO pattern
    e = float(re.search(r"^Total energy:\s*([-+0-9.eE]+)", text, re.M).group(1))
    assert e == -226.75
    body = re.search(r"^MO energies \(eV\):\s*\[(.*)\]\s*$", text, re.M).group(1)
    assert [float(v) for v in body.split(",")] == mo.tolist()
    assert "3.25e-7" in body  # Rust's Debug exponent form
    assert re.search(r"^Heat of formation:\s*(-57\.12345679) kcal/mol$", text, re.M)
