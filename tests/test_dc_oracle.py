"""The oracle's divide-and-conquer SCF (orc_run_dc) pinned by what the reference's own tests of it
assert and by what the mathematics fixes: the partition (src/dc/partition.rs:300-406), the density
DIIS (src/densitydiis.rs:121-213) and the SCF (src/dc/scf.rs:906-1055) -- a buffer reaching the
whole molecule reproduces the full diagonalisation, widening the buffer converges onto it, the
electron count is exact at every charge, the answer does not depend on where the cuts fall, and an
open shell polarises to its moment."""
import numpy as np
import pytest

A = 1.0 / 0.529177210903


def water_chain(n, spacing=3.2):
    """n waters along x (src/dc/scf.rs:767-777): O at spacing * i, H at +0.9584 along x and at
    (-0.24, 0.9278) Angstrom."""
    z, x = [], []
    for i in range(n):
        o = spacing * i
        z += [8, 1, 1]
        x += [[o, 0.0, 0.0], [o + 0.9584, 0.0, 0.0], [o - 0.24, 0.9278, 0.0]]
    return np.array(z, np.int32), np.array(x) * A


METHYL = (np.array([6, 1, 1, 1], np.int32),
          np.array([[0.0, 0.0, 0.0], [0.0, 1.078, 0.0], [0.9336, -0.539, 0.0], [-0.9336, -0.539, 0.0]]) * A)


def _opts(oracle, **kw):
    return oracle.options(max_scf=400, **kw)


# ---- the partition (src/dc/partition.rs:300-406) ----

def test_every_atom_belongs_to_exactly_one_core(oracle):
    z, x = water_chain(8, 6.0)
    p = oracle.dc_partition(z, x)
    off = np.concatenate([[0], np.cumsum(np.where(z == 1, 1, 4))])
    atom_of = np.repeat(np.arange(len(z)), np.diff(off))
    seen = np.zeros(off[-1], int)
    for s, (orb, inc) in enumerate(p["subsystems"]):
        # core atoms come first in every subsystem, and its core orbitals are its core atoms'
        assert inc[0] and np.all(np.diff(inc.astype(int)) <= 0)
        assert np.array_equal(p["core_of"][atom_of[orb]] == s, inc)
        seen[orb[inc]] += 1
    assert np.all(seen == 1) and p["n_subsystems"] > 1


def test_the_weights_sum_to_one_or_zero(oracle):
    z, x = water_chain(6, 6.0)
    w = oracle.dc_partition(z, x, want_weights=True)["weight_sum"]
    assert np.all((np.abs(w - 1.0) < 1e-12) | (np.abs(w) < 1e-12))


def test_a_reaching_buffer_drops_nothing_and_a_small_one_drops_distant_pairs(oracle):
    z, x = water_chain(6, 6.0)
    assert oracle.dc_partition(z, x, oracle.dc_options(buffer_radius=1000.0))["dropped_pairs"] == 0
    z, x = water_chain(8, 6.0)
    p = oracle.dc_partition(z, x, oracle.dc_options(core_radius=3.0, buffer_radius=3.0))
    assert p["dropped_pairs"] > 0 and p["largest_subsystem"] < 6 * 8


def test_subsystems_grow_monotonically_with_the_buffer(oracle):
    z, x = water_chain(8, 6.0)
    prev = 0
    for b in (2.0, 4.0, 8.0, 16.0):
        size = oracle.dc_partition(z, x, oracle.dc_options(core_radius=3.0, buffer_radius=b))["largest_subsystem"]
        assert size >= prev
        prev = size


def test_partition_by_brute_force(oracle):
    """The partition against a direct reading of its rules on a random cluster: cores grown from the
    unassigned atom furthest from the centroid (last index on ties), every atom within the core
    radius of the seed claimed, buffers every other atom within core + buffer radius of a core atom."""
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 14.0, (23, 3))
    z = np.where(rng.random(23) < 0.4, 8, 1).astype(np.int32)
    rc, rb = 4.0, 3.0
    p = oracle.dc_partition(z, x, oracle.dc_options(core_radius=rc, buffer_radius=rb))
    c = x.sum(axis=0) * (1.0 / len(z))
    core_of = -np.ones(len(z), int)
    cores = []
    while np.any(core_of < 0):
        free = np.flatnonzero(core_of < 0)
        d2 = ((x[free] - c) ** 2).sum(axis=1)
        seed = free[np.flatnonzero(d2 == d2.max())[-1]]
        mine = [a for a in free if np.sqrt(((x[a] - x[seed]) ** 2).sum()) <= rc]
        core_of[mine] = len(cores)
        cores.append(mine)
    assert np.array_equal(core_of, p["core_of"])
    off = np.concatenate([[0], np.cumsum(np.where(z == 1, 1, 4))])
    for s, core in enumerate(cores):
        atoms = list(core) + [a for a in range(len(z)) if core_of[a] != s and
                              min(np.sqrt(((x[a] - x[q]) ** 2).sum()) for q in core) <= rc + rb]
        orb = np.concatenate([np.arange(off[a], off[a + 1]) for a in atoms])
        inc = np.concatenate([[core_of[a] == s] * (off[a + 1] - off[a]) for a in atoms])
        assert np.array_equal(p["subsystems"][s][0], orb) and np.array_equal(p["subsystems"][s][1], inc)


# ---- the density DIIS (src/densitydiis.rs:121-213) ----

def test_density_diis_interpolates_to_the_zero_residual_point(oracle):
    out, depth = oracle.density_diis(np.array([[0.0], [4.0]]), np.array([[1.0], [-1.0]]))
    assert depth == 8 and abs(out[0] - 2.0) < 1e-10
    out, _ = oracle.density_diis(np.array([[1.0, 2.0, 3.0]]), np.array([[0.1, 0.2, 0.3]]))
    assert out is None  # one entry: nothing to extrapolate between


def test_density_diis_depth_and_the_linear_model(oracle):
    assert oracle.density_diis(np.zeros((1, 100000)), np.zeros((1, 100000)), memory_mb=1)[1] == 2
    assert oracle.density_diis(np.zeros((1, 100)), np.zeros((1, 100)), memory_mb=512)[1] == 8
    # residuals linear in the iterate, r = M (p* - p), with p* an affine combination of the
    # iterates: the extrapolation recovers p* (the residual of the combination vanishes)
    rng = np.random.default_rng(3)
    M = rng.normal(size=(12, 12)) + 6 * np.eye(12)
    dens = rng.normal(size=(6, 12))
    a = np.array([0.3, -0.2, 0.5, 0.1, 0.6, -0.3])
    target = a @ dens
    out, _ = oracle.density_diis(dens, (target - dens) @ M.T)
    assert np.abs(out - target).max() < 1e-9
    # a history deeper than the depth keeps only the newest eight
    dens = rng.normal(size=(20, 12))
    a = rng.uniform(-0.5, 0.5, 8)
    a[-1] += 1.0 - a.sum()
    target = a @ dens[12:]
    res = (target - dens) @ M.T
    res[:12] += rng.normal(size=(12, 12))  # the dropped entries would spoil the fit
    out, _ = oracle.density_diis(dens, res)
    assert np.abs(out - target).max() < 1e-9


# ---- the SCF (src/dc/scf.rs:906-1055) ----

def test_a_reaching_buffer_reproduces_the_full_result_exactly(oracle):
    z, x = water_chain(3)
    st, full = oracle.run_pm3(z, x, opts=_opts(oracle))
    st2, dc = oracle.run_dc(z, x, opts=_opts(oracle),
                            dc=oracle.dc_options(core_radius=2.0, buffer_radius=500.0, smearing_ev=1e-4))
    assert st == 0 and st2 == 0 and dc["dropped_pairs"] == 0 and dc["n_subsystems"] > 1
    assert abs(dc["total_ev"] - full["total_ev"]) < 1e-6
    assert np.abs(dc["charges"] - full["charges"]).max() < 1e-5
    assert np.abs(dc["density"] - full["density"]).max() < 1e-5


def test_widening_the_buffer_converges_onto_the_full_diagonalization(oracle):
    z, x = water_chain(6)
    ref = oracle.run_pm3(z, x, opts=_opts(oracle))[1]["total_ev"]
    err = []
    for b in (3.0, 6.0, 10.0, 16.0):
        st, r = oracle.run_dc(z, x, opts=_opts(oracle), dc=oracle.dc_options(core_radius=3.0, buffer_radius=b))
        assert st == 0
        err.append(abs(r["total_ev"] - ref))
    assert err[3] < err[0] and err[3] < 5e-4


@pytest.mark.parametrize("charge", [0.0, 1.0, -1.0])
def test_the_electron_count_is_exact(oracle, charge):
    z, x = water_chain(4)
    st, r = oracle.run_dc(z, x, charge=charge, multiplicity=1 if charge == 0 else 2, opts=_opts(oracle),
                          dc=oracle.dc_options(core_radius=3.0, buffer_radius=8.0))
    assert st == 0 and r["unrestricted"] == (charge != 0)
    assert abs(np.trace(r["density"]) - (32.0 - charge)) < 1e-6
    assert abs(r["charges"].sum() - charge) < 1e-6


def test_the_result_does_not_depend_on_where_the_cuts_fall(oracle):
    z, x = water_chain(5)
    e = [oracle.run_dc(z, x, opts=_opts(oracle), dc=oracle.dc_options(core_radius=c, buffer_radius=14.0))[1]["total_ev"]
         for c in (6.0, 2.5)]
    assert abs(e[0] - e[1]) < 1e-4


def test_an_open_shell_system_converges_with_the_right_moment(oracle):
    z, x = METHYL
    st, r = oracle.run_dc(z, x, multiplicity=2, opts=_opts(oracle),
                          dc=oracle.dc_options(core_radius=4.0, buffer_radius=500.0, smearing_ev=1e-4))
    assert st == 0 and r["unrestricted"]
    assert abs(np.trace(r["spin_density"]) - 1.0) < 1e-5


def test_a_reaching_buffer_reproduces_the_unrestricted_result(oracle):
    # a water-chain cation, split into several subsystems whose buffers reach everything: the
    # unrestricted single point (src/scf.rs:1319-1496) to 1e-6 eV
    z, x = water_chain(3)
    st, r = oracle.run_dc(z, x, charge=1.0, multiplicity=2, opts=_opts(oracle),
                          dc=oracle.dc_options(core_radius=2.0, buffer_radius=500.0, smearing_ev=1e-4))
    st2, full = oracle.run_uhf(z, x, charge=1.0, multiplicity=2, opts=_opts(oracle))
    assert st == 0 and st2 == 0 and r["n_subsystems"] > 1
    assert abs(r["total_ev"] - full["total_ev"]) < 1e-6
    assert np.abs(r["spin_density"] - full["spin_density"]).max() < 1e-4


def test_dc_gradient_is_the_derivative_of_the_full_energy_at_full_coverage(oracle):
    """With the buffer reaching everything the partitioned density is the SCF density, so the
    gradient is the full analytic gradient (src/dc/derivatives.rs:55-121), and a central
    difference of the partitioned energy agrees with it."""
    z, x = water_chain(2)
    dc = oracle.dc_options(core_radius=2.0, buffer_radius=500.0, smearing_ev=1e-4)
    tight = _opts(oracle, e_tol=1e-12, p_tol=1e-11)
    st, r = oracle.dc_gradient(z, x, opts=tight, dc=dc)
    st2, full = oracle.gradient(z, x, opts=tight)
    assert st == 0 and st2 == 0
    assert np.abs(r["gradient"] - full["gradient"]).max() < 1e-6
    h = 1e-4
    for a, c in ((0, 0), (2, 1), (4, 2)):
        e = []
        for s in (1, -1):
            xs = x.copy()
            xs[a, c] += s * h
            e.append(oracle.run_dc(z, xs, opts=tight, dc=dc)[1]["total_ev"])
        assert abs((e[0] - e[1]) / (2 * h) - r["gradient"][a, c]) < 2e-5


def test_dc_refusals(oracle):
    z, x = water_chain(2)
    assert oracle.run_dc(z, x, opts=_opts(oracle), dc=oracle.dc_options(core_radius=0.0))[0] == 1
    assert oracle.run_dc(z, x, opts=_opts(oracle), dc=oracle.dc_options(buffer_radius=-1.0))[0] == 1
    assert oracle.run_dc(z, x, multiplicity=2, opts=_opts(oracle, reference=1))[0] == 1  # RHF for an open shell
    assert oracle.run_dc(z, x, opts=oracle.options(max_scf=2))[0] == 3
