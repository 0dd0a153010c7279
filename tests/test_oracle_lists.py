"""Pins of the oracle's binning and neighbour lists against brute force (CPU).

- cell ids: away from cell faces the reference formula must give the exact
  floor(x N / Lb) + 1 (checked with rational arithmetic); every id in [1, N+1];
- pair and ion lists: the set of j in full(i) equals the O(N^2) minimum-image
  predicate r^2 < rc^2 (resp. rclj^2, with a Ga in the pair) for boxes with
  N >= 3 cells, where the 27-cell stencil covers the cutoff sphere;
- half lists = {j in full(i): j > i}; triplets = connected triples of the
  buffered graph, each exactly once (si.f90:191-232).
"""
from fractions import Fraction
import itertools

import numpy as np
import pytest

from tests import gpu_common as gc


def _brute(X, at, L, rc, rclj):
    L = np.asarray(L)
    n = len(X)
    sw = [set() for _ in range(n)]
    io = [set() for _ in range(n)]
    for i in range(n):
        d = X[i] - X
        d = d - np.round(d / L) * L  # away from ties NINT == round
        r2 = (d * d).sum(1)
        for j in np.nonzero(r2 < rclj * rclj)[0]:
            if j == i:
                continue
            if max(at[i], at[j]) < 3:
                if r2[j] < rc * rc:
                    sw[i].add(j + 1)
            else:
                io[i].add(j + 1)
    return sw, io


@pytest.mark.parametrize("case", ["bulk", "ions", "slab"])
def test_full_lists_equal_brute_force(oracle_lib, case):
    if case == "bulk":
        X, at, L, nsg, nlj = gc.crystal_with_ions(3, 0.2e-10, 3, 0)
    elif case == "ions":
        X, at, L, nsg, nlj = gc.crystal_with_ions(3, 0.2e-10, 4, 3)
    else:
        X, at, L, nsg, nlj = gc.reference_slab((3, 3, 2), 2)
    o = gc.oracle_system(X, at, L, nsg, nlj)
    assert min(o.geti("nc1"), o.geti("nc2"), o.geti("nc3")) >= 3
    sw, io = _brute(X, at, L, o.getd("rc"), o.getd("rclj"))
    c2, f2, clj, flj = o.full_lists()
    p2 = np.concatenate([[0], np.cumsum(c2)])
    plj = np.concatenate([[0], np.cumsum(clj)])
    cell = o.getd("lb1") / o.geti("nc1")
    for i in range(len(X)):
        assert set(f2[p2[i]:p2[i + 1]]) == sw[i]
        assert len(f2[p2[i]:p2[i + 1]]) == len(sw[i])  # no duplicates
        # ion list is stencil-limited: equal to brute force within one cell
        got = set(flj[plj[i]:plj[i + 1]])
        assert got <= io[i]
        near = {j for j in io[i]
                if np.sqrt((((X[i] - X[j - 1]) - np.round((X[i] - X[j - 1]) / np.asarray(L)) * np.asarray(L)) ** 2).sum()) < cell}
        assert near <= got


def test_half_lists_and_triplets(oracle_lib):
    X, at, L, nsg, nlj = gc.crystal_with_ions(2, 0.25e-10, 8, 1)
    # 2x2x2 cells -> Nc = 2: all atoms are scanned (no periodic halo needed)
    o = gc.oracle_system(X, at, L, nsg, nlj)
    h2, hlj = o.half_sets()
    c2, f2, clj, flj = o.full_lists()
    p2 = np.concatenate([[0], np.cumsum(c2)])
    full = [set(f2[p2[i]:p2[i + 1]]) for i in range(len(X))]
    for i in range(len(X)):
        assert h2[i] == {j for j in full[i] if j > i + 1}
    trip = o.nlists()["nlist3"]
    got = [tuple(t) for t in trip]
    assert len(got) == len(set(got))
    adj = [set(full[i]) for i in range(len(X))]
    want = set()
    n = nsg
    for a, b, c in itertools.combinations(range(1, n + 1), 3):
        e = (b in adj[a - 1]) + (c in adj[a - 1]) + (c in adj[b - 1])
        if e >= 2:
            want.add((a, b, c))
    assert set(got) == want


def test_cell_ids_exact_floor(oracle_lib):
    X, at, L, nsg, nlj = gc.crystal_with_ions(4, 0.3e-10, 12, 0)
    o = gc.oracle_system(X, at, L, nsg, nlj)
    ic = o.atomcell(X)
    N = [o.geti("nc1"), o.geti("nc2"), o.geti("nc3")]
    for l in range(3):
        assert ic[:, l].min() >= 1 and ic[:, l].max() <= N[l] + 1
        for i in range(len(X)):
            x = Fraction(X[i, l]) / Fraction(L[l]) * N[l]
            fl = x.numerator // x.denominator
            if abs(x - fl) > Fraction(1, 10 ** 9) and abs(x - fl - 1) > Fraction(1, 10 ** 9):
                assert ic[i, l] == fl + 1
