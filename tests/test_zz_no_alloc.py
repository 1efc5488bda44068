"""The allocation-free entry point (pm3_run_noalloc) against pm3_run on the same molecules: two
calls in a row on the shared scratch block, the second smaller than the first, each to the parity
bar (|dHf| <= 1e-5 kcal/mol, max|dP| <= 1e-8), and the size refusal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TIGHT = dict(p_tol=1e-10, e_tol=1e-12)


def test_noalloc_matches_pm3_run_on_two_calls_in_a_row():
    import pm3amd as G
    from pm3amd import workloads as W
    z1, x1 = W.alkane(16)  # C16H34: 50 atoms, 98 AOs
    z2, x2 = W.water_cluster(3, 5.0, 11)  # 9 atoms, 18 AOs
    opts = G.Pm3Options(**TIGHT)
    for z, x in ((z1, x1), (z2, x2)):
        m = G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))
        print("case noalloc nat", len(m.z), flush=True)
        a = G.run_pm3_noalloc(m, options=opts)
        b = G.run_pm3(m, options=opts, want_orbitals=False)
        assert a.converged and b.converged
        assert abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) <= 1e-5
        assert np.abs(a.density - b.density).max() <= 1e-8
        assert np.abs(a.charges - b.charges).max() <= 1e-8
        assert a.iterations == b.iterations and a.scf_paths_tried == b.scf_paths_tried


def test_noalloc_refuses_molecules_beyond_its_maximum():
    import pm3amd as G
    from pm3amd import workloads as W
    z, x = W.alkane(24)  # 146 AOs
    with pytest.raises(G.ResourceLimit, match="at most"):
        G.run_pm3_noalloc(G.Molecule(np.asarray(z, np.int32), W.to_bohr(x)))
