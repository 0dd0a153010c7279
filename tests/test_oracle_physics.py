"""Pins of the oracle's force field (CPU, no GPU).

Each test checks the oracle against something the mathematics fixes rather
than a retyped formula:
- SW pair minimum: phi2(2^(1/6) sigma) = -eps, phi2'(2^(1/6) sigma) = 0 (the
  published SW parametrisation puts the dimer minimum there);
- perfect diamond at a = 4 * 2^(1/6) sigma / sqrt(3): E/atom = -2 eps, F = 0
  (3-body term vanishes at cos = -1/3);
- Moliere: V(r) r -> Z1 Z2 e^2/(4 pi eps0) = 14.3996 eV A Z1 Z2 for r -> 0
  (screening function sums to 1), Firsov length a_F = 0.885 a0 (Z1^1/2+Z2^1/2)^-2/3;
- derivative table (si.f90:575-597) within O(du^3) of the analytic phi2';
- F = -grad E by central differences on a perturbed crystal with Ga ions;
- virial W_aa = -dE/d(strain_aa) by finite strain;
- sum F = 0 and invariance under translation / rigid rotation.
"""
import math

import numpy as np
import pytest

from fib_md_amd import configs
from tests import gpu_common as gc

EPS_SI = 3.473928e-19
EPS_GE = 3.085e-19  # prms.f90 Ge SW energy
SIGMA = 2.0951e-10
EV = 1.60217646e-19
COULOMB_EV_A = 14.399645  # e^2/(4 pi eps0) in eV*A (CODATA)


def _single(o_lib, X, at, L, nsg, nlj, impact=0):
    return gc.oracle_system(X, at, L, nsg, nlj, impact)


def _cluster(X, at, nsg, nlj, box=60e-10, impact=None):
    """Atoms in a large box (no periodic image interacts)."""
    L = (box, box, box)
    c = X.mean(0)
    Xc = X - c + box / 2
    return gc.oracle_system(Xc, at, L, nsg, nlj, nlj if impact is None else impact), Xc, L


def test_sw_dimer_minimum(oracle_lib):
    rmin = 2 ** (1 / 6) * SIGMA
    for r, expect in ((rmin, -EPS_SI),):
        X = np.array([[0, 0, 0], [r, 0, 0]], float)
        o, Xc, _ = _cluster(X, np.ones(2, np.int32), 2, 0)
        U, u = o.si_potential(Xc)
        assert abs(U - expect) <= 1e-9 * EPS_SI
        F = o.si_force(Xc)
        # force ~ 0 at the minimum (table interpolation error only)
        assert np.abs(F).max() <= 1e-6 * EPS_SI / SIGMA


def test_perfect_diamond_energy_and_forces(oracle_lib):
    a = 4 * 2 ** (1 / 6) * SIGMA / math.sqrt(3)
    n = 4
    X = configs.diamond(n, n, n, a)
    N = len(X)
    o = gc.oracle_system(X, np.ones(N, np.int32), (n * a,) * 3, N, 0)
    U, u = o.si_potential(X)
    assert abs(U / N + 2 * EPS_SI) <= 1e-9 * EPS_SI
    np.testing.assert_allclose(u, -2 * EPS_SI, rtol=1e-9)
    F = o.si_force(X)
    assert np.abs(F).max() <= 1e-9 * EPS_SI / SIGMA
    # 92 connected triples per atom in the buffered crystal graph
    assert o.geti("nbrs") == 92 * N


def test_perfect_diamond_sige_bond_energies(oracle_lib):
    """Random Si/Ge occupation of the perfect diamond at the SW minimum
    (sigma_Ge = sigma_Si): every bond sits at the pair minimum and every angle
    is tetrahedral, so all forces vanish and each bond carries -eps_ij, with
    the mixing eps_SiGe = (eps_Si + eps_Ge) / 2 (prms.f90:452-480).  The per-atom
    energy is half of its four bonds."""
    a = 4 * 2 ** (1 / 6) * SIGMA / math.sqrt(3)
    n = 3
    X = configs.diamond(n, n, n, a)
    N = len(X)
    L = n * a
    at = gc.with_ge(np.ones(N, np.int32), 0.4, 17)
    eps = {1: EPS_SI, 2: EPS_GE}
    o = gc.oracle_system(X, at, (L,) * 3, N, 0)
    U, u = o.si_potential(X)
    # nearest neighbours by brute force (distance a sqrt(3)/4)
    d = X[:, None, :] - X[None, :, :]
    d -= np.round(d / L) * L
    r = np.sqrt((d * d).sum(2))
    nn = np.abs(r - a * math.sqrt(3) / 4) < 1e-3 * a
    assert (nn.sum(1) == 4).all()
    e_ij = np.array([[0.5 * (eps[int(at[i])] + eps[int(at[j])]) for j in range(N)] for i in range(N)])
    u_ref = -0.5 * (e_ij * nn).sum(1)
    np.testing.assert_allclose(u, u_ref, rtol=1e-9)
    assert abs(U - u_ref.sum()) <= 1e-9 * abs(u_ref.sum())
    F = o.si_force(X)
    assert np.abs(F).max() <= 1e-9 * EPS_SI / SIGMA


def test_moliere_coulomb_limit_and_firsov(oracle_lib):
    r = 1e-4 * 1e-10  # 1e-4 A
    X = np.array([[0, 0, 0], [r, 0, 0]], float)
    at = np.array([1, 3], np.int32)
    o, Xc, _ = _cluster(X, at, 1, 1)
    U, _ = o.si_potential(Xc)
    k = U / EV * (r * 1e10) / (14.0 * 31.0)
    assert abs(k - COULOMB_EV_A) <= 2e-3 * COULOMB_EV_A
    aF = 0.885 * 0.53 * (math.sqrt(14) + math.sqrt(31)) ** (-2 / 3) * 1e-10
    assert abs(o.getd("aamol13") - aF) <= 1e-12 * aF
    # screening function at r = aF: sum c_k exp(-b_k) (Moliere 1947)
    X2 = np.array([[0, 0, 0], [aF, 0, 0]], float)
    o2, Xc2, _ = _cluster(X2, at, 1, 1)
    U2, _ = o2.si_potential(Xc2)
    phi = 0.35 * math.exp(-0.3) + 0.55 * math.exp(-1.2) + 0.10 * math.exp(-6.0)
    v = 14.0 * 31.0 * COULOMB_EV_A / (aF * 1e10) * phi
    assert abs(U2 / EV - v) <= 1e-4 * v


def test_table_matches_analytic_derivative(oracle_lib):
    o = gc.oracle_system(np.zeros((1, 3)), np.ones(1, np.int32), (10e-10,) * 3, 1, 0)
    A, B, al = 7.049556277, 0.6022245584, 1.8

    def dphi(x):
        ex = math.exp(1 / (x - al))
        return A * ex * (-4 * B * x ** -5 - (B * x ** -4 - 1) / (x - al) ** 2)

    # 3-point Lagrange remainder on equal spacing du:
    #   |err| <= du^3 / (9 sqrt 3) * max |f'''|  with f = phi2'
    du = 2.5 * 1.1 * 1.8 / 11999

    def d3(x, h=1e-3):  # third derivative of f = phi2' (central differences)
        return (dphi(x + 2 * h) - 2 * dphi(x + h) + 2 * dphi(x - h) - dphi(x - 2 * h)) / (2 * h ** 3)

    for x in np.linspace(1.0, 1.75, 57):
        got = o.itable2bd(x)
        ref = dphi(x)
        m3 = max(abs(d3(x + s)) for s in np.linspace(-2 * du, 2 * du, 9))
        bound = du ** 3 / (9 * math.sqrt(3)) * m3
        assert abs(got - ref) <= 1.5 * bound + 1e-12, (x, got - ref, bound)


def _perturbed_with_ions(seed=3):
    return gc.crystal_with_ions(3, 0.15e-10, seed, 2)


@pytest.mark.parametrize("species", ["si", "sige"])
def test_force_is_minus_gradient(oracle_lib, species):
    X, at, L, nsg, nlj = _perturbed_with_ions()
    if species == "sige":
        at = gc.with_ge(at, 0.4, 23)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact=nlj)
    F = o.si_force(X)
    h = 1e-15
    fmax = np.abs(F).max()
    rng = np.random.default_rng(0)
    picks = list(rng.choice(nsg, 6, replace=False)) + [nsg, nsg + 1]
    for i in picks:
        for c in range(3):
            Xp = X.copy(); Xp[i, c] += h
            Xm = X.copy(); Xm[i, c] -= h
            fd = -(o.si_potential(Xp)[0] - o.si_potential(Xm)[0]) / (2 * h)
            assert abs(fd - F[i, c]) <= 1e-7 * fmax, (i, c, fd, F[i, c])


def test_virial_is_strain_derivative(oracle_lib):
    X, at, L, nsg, nlj = _perturbed_with_ions(5)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact=nlj)
    _, W = o.si_force(X, virial=True)
    e = 1e-7
    for a in range(3):
        vals = []
        for s in (+1, -1):
            S = np.eye(3); S[a, a] += s * e
            Ls = list(L); Ls[a] *= (1 + s * e)
            Xs = X @ S.T
            os_ = gc.oracle_system(Xs, at, Ls, nsg, nlj, impact=nlj)
            vals.append(os_.si_potential(Xs)[0])
        dEde = (vals[0] - vals[1]) / (2 * e)
        assert abs(-dEde - W[a, a]) <= 1e-6 * np.abs(np.diag(W)).max()
    # the total virial is symmetric (no net torque)
    assert np.abs(W - W.T).max() <= 1e-9 * np.abs(W).max()


def test_newton3_translation_rotation(oracle_lib):
    X, at, L, nsg, nlj = _perturbed_with_ions(9)
    o = gc.oracle_system(X, at, L, nsg, nlj, impact=nlj)
    F = o.si_force(X)
    fmax = np.abs(F).max()
    assert np.abs(F.sum(0)).max() <= 1e-9 * fmax
    # translation by an arbitrary vector (periodic box)
    sh = np.array([1.234e-10, -0.77e-10, 2.5e-10])
    Xt = np.mod(X + sh, np.array(L))
    ot = gc.oracle_system(Xt, at, L, nsg, nlj, impact=nlj)
    assert np.abs(ot.si_force(Xt) - F).max() <= 1e-9 * fmax
    # rigid rotation of an isolated cluster
    Xc = configs.diamond(2, 2, 2) + np.random.default_rng(1).normal(0, 0.1e-10, (64, 3))
    oc, Xcc, Lc = _cluster(Xc, np.ones(64, np.int32), 64, 0, box=80e-10)
    Fc = oc.si_force(Xcc)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    c = np.array(Lc) / 2
    Xr = (Xcc - c) @ R.T + c
    orr = gc.oracle_system(Xr, np.ones(64, np.int32), Lc, 64, 0)
    Fr = orr.si_force(Xr)
    assert np.abs(Fr - Fc @ R.T).max() <= 1e-8 * np.abs(Fc).max()
