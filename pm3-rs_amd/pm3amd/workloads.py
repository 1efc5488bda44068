"""Seeded synthetic geometries for the PM3 benchmark workloads (all deterministic from a seed).

* ``water_cluster(k, box, seed)``: rigid waters (O-H 0.9572 A, H-O-H 104.52 deg), orientations
  uniform on SO(3) (Shoemake quaternions), O positions uniform in a cube by rejection at
  min O-O 2.7 A.  32 waters in a 10 A cube give the 96-atom / 192-AO cluster; 512 in a 25 A cube
  the 1536-atom / 3072-AO cluster.
* ``alkane(n)``: all-trans CnH2n+2 zigzag, C-C 1.54 A, C-H 1.09 A, tetrahedral angles.
* ``mixed_batch(count, seed)``: half alkanes (n uniform in [1, 24]), half water clusters
  (k uniform in [2, 16], cube scaled to the 32-water density), every atom jittered by
  Gaussian sigma 0.05 A.
* ``solvated_conformers(count, seed)``: C12H26 plus 20 waters placed at min heavy-atom distance
  2.6 A (98 atoms, 194 AOs), then ``count`` conformers jittered by sigma 0.1 A.

Coordinates are returned in Angstrom; ``to_bohr`` converts with the reference's factor
(src/constants.rs:33).
"""

import math

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
R_OH = 0.9572
ANG_HOH = math.radians(104.52)
R_CC = 1.54
R_CH = 1.09
TETRA = math.acos(-1.0 / 3.0)


def to_bohr(xyz):
    return np.asarray(xyz, dtype=np.float64) * ANGSTROM_TO_BOHR


def random_rotation(rng):
    """Uniform rotation on SO(3) from a Shoemake unit quaternion."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_frame():
    return np.array([[0.0, 0.0, 0.0], [R_OH, 0.0, 0.0], [R_OH * math.cos(ANG_HOH), R_OH * math.sin(ANG_HOH), 0.0]])


def _place(rng, count, lo, hi, min_dist, fixed=None, max_tries=2_000_000):
    """Uniform points in the box [lo, hi] accepted when at least min_dist from every accepted
    point and every row of `fixed`."""
    lo = np.asarray(lo, dtype=float)
    hi = np.asarray(hi, dtype=float)
    pts = np.zeros((0, 3)) if fixed is None else np.asarray(fixed, dtype=float).reshape(-1, 3)
    nfix = len(pts)
    tries = 0
    d2 = min_dist * min_dist
    while len(pts) - nfix < count:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("rejection sampling did not place every molecule")
        p = lo + (hi - lo) * rng.random(3)
        if len(pts) and np.min(np.sum((pts - p) ** 2, axis=1)) < d2:
            continue
        pts = np.vstack([pts, p])
    return pts[nfix:]


def water_cluster(k, box, seed, min_oo=2.7):
    """(z, xyz_angstrom) of k rigid waters in a cube of edge `box` (A)."""
    rng = np.random.default_rng(seed)
    opos = _place(rng, k, [0, 0, 0], [box, box, box], min_oo)
    frame = _water_frame()
    z, xyz = [], []
    for o in opos:
        R = random_rotation(rng)
        for zi, v in zip((8, 1, 1), frame):
            z.append(zi)
            xyz.append(o + R @ v)
    return np.array(z, dtype=np.int32), np.array(xyz)


def alkane(n):
    """(z, xyz_angstrom) of all-trans CnH2n+2."""
    if n == 1:
        d = R_CH / math.sqrt(3.0)
        xyz = [[0, 0, 0], [d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d]]
        return np.array([6, 1, 1, 1, 1], dtype=np.int32), np.array(xyz, dtype=float)
    half = TETRA / 2.0
    cpos = np.array([[i * R_CC * math.sin(half), (i % 2) * R_CC * math.cos(half), 0.0] for i in range(-1, n + 1)])
    carbons = cpos[1:-1]
    z = [6] * n
    xyz = [c for c in carbons]
    nrm = np.array([0.0, 0.0, 1.0])
    for i in range(n):
        c = cpos[i + 1]
        a = cpos[i] - c
        b = cpos[i + 2] - c
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        bis = -(a + b)
        bis /= np.linalg.norm(bis)
        phi = TETRA / 2.0
        for s in (1.0, -1.0):
            z.append(1)
            xyz.append(c + R_CH * (math.cos(phi) * bis + s * math.sin(phi) * nrm))
        # terminal carbons: the virtual neighbour continuing the zigzag becomes a hydrogen
        if i == 0:
            z.append(1)
            xyz.append(c + R_CH * a)
        if i == n - 1:
            z.append(1)
            xyz.append(c + R_CH * b)
    return np.array(z, dtype=np.int32), np.array(xyz)


def jitter(xyz, sigma, rng):
    return xyz + rng.normal(0.0, sigma, size=xyz.shape)


def water_box_for(k):
    """Cube edge (A) giving k waters the density of 32 waters in a 10 A cube."""
    return (k * 1000.0 / 32.0) ** (1.0 / 3.0)


def mixed_batch(count, seed, sigma=0.05):
    """List of (z, xyz_angstrom): alternating alkanes (n in [1,24]) and water clusters (k in
    [2,16]), jittered."""
    rng = np.random.default_rng(seed)
    mols = []
    for i in range(count):
        if i % 2 == 0:
            n = int(rng.integers(1, 25))
            z, xyz = alkane(n)
        else:
            k = int(rng.integers(2, 17))
            z, xyz = water_cluster(k, water_box_for(k), int(rng.integers(0, 2**31 - 1)))
        mols.append((z, jitter(xyz, sigma, rng)))
    return mols


def solvated_alkane(seed, n_carbon=12, n_water=20, min_heavy=2.6):
    """C12H26 with 20 waters whose O sit at least `min_heavy` A from every heavy atom."""
    rng = np.random.default_rng(seed)
    za, xa = alkane(n_carbon)
    heavy = xa[za != 1]
    lo = xa.min(axis=0) - 4.0
    hi = xa.max(axis=0) + 4.0
    opos = _place(rng, n_water, lo, hi, min_heavy, fixed=heavy)
    frame = _water_frame()
    z, xyz = list(za), list(xa)
    for o in opos:
        R = random_rotation(rng)
        for zi, v in zip((8, 1, 1), frame):
            z.append(zi)
            xyz.append(o + R @ v)
    return np.array(z, dtype=np.int32), np.array(xyz)


def solvated_conformers(count, seed, sigma=0.1):
    z, base = solvated_alkane(seed)
    rng = np.random.default_rng(seed + 1)
    return [(z, jitter(base, sigma, rng)) for _ in range(count)]
