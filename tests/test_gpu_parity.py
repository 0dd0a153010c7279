"""GPU parity: the HIP path (through the C-ABI) against the oracle.

Bars: cell ids and neighbour sets bit-exact; FP64 forces within 1e-10 of
max|F| (summation order differs); energies within 1e-12 relative; the mixed
path within 1e-5 relative force error (floor 1e-3 F_rms) of FP64.
"""
import numpy as np
import pytest

from tests import gpu_common as gc

pytestmark = pytest.mark.gpu


def _sets_equal(o, g):
    o2, olj = o.half_sets()
    g2, glj = g.half_sets()
    bad2 = [i for i in range(len(o2)) if o2[i] != g2[i]]
    badlj = [i for i in range(len(olj)) if olj[i] != glj[i]]
    return bad2, badlj


@pytest.mark.parametrize("case", ["bulk", "ions", "slab", "small", "sige"])
def test_cells_and_lists_bit_exact(oracle_lib, case):
    if case == "bulk":
        X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.12e-10, 3, 0)
    elif case == "ions":
        X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.12e-10, 5, 4)
    elif case == "sige":
        X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.12e-10, 6, 3)
        at = gc.with_ge(at, 0.4, 7)
    elif case == "slab":
        X, at, L, nsg, nlj = gc.reference_slab((4, 4, 3), 3)
    else:
        X, at, L, nsg, nlj = gc.crystal_with_ions(2, 0.1e-10, 9, 1)  # Nc = 2
    print(f"case={case} natm={len(X)}", flush=True)
    o = gc.oracle_system(X, at, L, nsg, nlj)
    g = gc.gpu_system(X, at, L, nsg, nlj)
    np.testing.assert_array_equal(g.bin_cells(), o.atomcell(X))
    bad2, badlj = _sets_equal(o, g)
    assert not bad2 and not badlj, (bad2[:5], badlj[:5])


def test_binning_edge_positions(oracle_lib):
    """Coordinates at 0, -tiny and within ulps of Lb (X*iLb rounding to 1)."""
    X, at, L, nsg, nlj = gc.crystal_with_ions(3, 0.1e-10, 13, 0)
    Lx = L[0]
    X[0, 0] = 0.0
    X[1, 0] = np.nextafter(Lx, 0)
    X[2, 1] = np.nextafter(np.nextafter(L[1], 0), 0)
    X[3, 2] = -1e-30
    X[4, 0] = Lx * (1 - 2.2e-16)
    X[5, :] = np.nextafter(np.array(L), 0)
    o = gc.oracle_system(X, at, L, nsg, nlj)
    g = gc.gpu_system(X, at, L, nsg, nlj)
    np.testing.assert_array_equal(g.bin_cells(), o.atomcell(X))
    bad2, badlj = _sets_equal(o, g)
    assert not bad2 and not badlj


@pytest.mark.parametrize("case", ["bulk", "ions", "slab", "sige"])
def test_force_fp64(oracle_lib, case):
    if case == "bulk":
        X, at, L, nsg, nlj = gc.crystal_with_ions(5, 0.12e-10, 21, 0)
    elif case == "ions":
        X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.12e-10, 23, 4)
    elif case == "sige":
        # Si/Ge mixing of the pair and three-body terms, Ge-Ga Moliere pairs
        X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.12e-10, 25, 4)
        at = gc.with_ge(at, 0.4, 9)
    else:
        X, at, L, nsg, nlj = gc.reference_slab((4, 4, 3), 3)
    impact = nlj  # all ions fired
    print(f"case={case} natm={len(X)}", flush=True)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact)
    Fo, Wo = o.si_force(X, virial=True)
    Uo, uo = o.si_potential(X)
    g = gc.gpu_system(X, at, L, nsg, nlj)
    r = g.force(impact, energy=True, virial=True)
    fmax = np.abs(Fo).max()
    assert np.abs(r["F"] - Fo).max() <= 1e-10 * fmax
    assert abs(r["utot"] - Uo) <= 1e-12 * abs(Uo)
    np.testing.assert_allclose(r["u"], uo, rtol=0, atol=1e-12 * np.abs(uo).max())
    assert np.abs(r["virial"] - Wo).max() <= 1e-10 * np.abs(Wo).max()


def test_force_unfired_ions_skipped(oracle_lib):
    X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.1e-10, 31, 3)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact=1)
    Fo = o.si_force(X)
    g = gc.gpu_system(X, at, L, nsg, nlj)
    r = g.force(1)
    assert np.abs(r["F"] - Fo).max() <= 1e-10 * np.abs(Fo).max()
    assert np.all(r["F"][nsg + 1:] == 0.0)


def test_force_mixed_vs_fp64(oracle_lib):
    from fib_md_amd import configs
    X, at, L = configs.force_microbench(6, seed=1002)
    n = len(X)
    o = gc.oracle_system(X, at, L, n, 0)
    Fo, Wo = o.si_force(X, virial=True)
    Uo, _ = o.si_potential(X)
    g = gc.gpu_system(X, at, L, n, 0, precision="mixed")
    r = g.force(0, energy=True, virial=True)
    emax, erms, enorm = gc.rel_err(r["F"], Fo)
    print(f"mixed: max {emax:.3e} rms {erms:.3e} norm {enorm:.3e}")
    assert emax <= 1e-5
    assert abs(r["utot"] - Uo) <= 1e-6 * abs(Uo)
    assert np.abs(r["virial"] - Wo).max() <= 1e-5 * np.abs(Wo).max()


def test_force_mixed_sige_with_ions(oracle_lib):
    """Mixed path on a Si/Ge crystal with fired Ga: the species-mixed SW
    terms (length unit and eps/sigma of each pair, Ge legs in the three-body
    pass) and the FP32 Moliere pairs against the oracle's FP64, to the same
    bounds as the pure-Si case."""
    X, at, L, nsg, nlj = gc.crystal_with_ions(5, 0.12e-10, 27, 4)
    at = gc.with_ge(at, 0.4, 11)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact=nlj)
    Fo, Wo = o.si_force(X, virial=True)
    Uo, _ = o.si_potential(X)
    g = gc.gpu_system(X, at, L, nsg, nlj, precision="mixed")
    r = g.force(nlj, energy=True, virial=True)
    emax, erms, enorm = gc.rel_err(r["F"], Fo)
    print(f"mixed SiGe+Ga: max {emax:.3e} rms {erms:.3e} norm {enorm:.3e}")
    assert emax <= 1e-5
    assert abs(r["utot"] - Uo) <= 1e-6 * abs(Uo)
    assert np.abs(r["virial"] - Wo).max() <= 1e-5 * np.abs(Wo).max()
