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
