"""Pins of the oracle's RNG, initial state and time integration (CPU).

- ran1 (ran1.f:10-36): its LCG is the Park-Miller minimal standard, pinned by
  the published check value x_10000 = 1043618065 from x_0 = 1; the shuffle
  conserves values: outputs plus the 32-entry table are exactly the LCG states
  x_9 .. x_(40+n);
- pureepi (prms.f90:484-557): 8 n1 n2 n3 sites of the diamond lattice up to
  the E20.10 text round trip; xsides = atoms on the x/y boundary planes;
- randvel (data.f90:169-204): zero total momentum, T = m Uo^2 / (12 kB) for
  uniform velocities in [-Uo/2, Uo/2];
- initranions: Gaussian spot with sigma = fwhm/2.35, six significant digits;
- onestep: free flight is X0 + n V Ts (mod Lb); total energy conserved to
  O(Ts^2) without thermostat; Berendsen factor on a state of known side
  temperature; the side mask's precedence (x-side OR (y-side AND z-window));
  the regime switch at max KE < ke_limit_1.
"""
import math

import numpy as np
import pytest

from fib_md_amd import configs
from oracle.oracle import Oracle, REFERENCE_SIGA, siga_from_values
from tests import gpu_common as gc

IM = 2147483647
KB = 1.381e-23
MSI = 46.637063e-27


def lcg(n, x0=1):
    xs = [x0]
    x = x0
    for _ in range(n):
        x = (16807 * x) % IM
        xs.append(x)
    return xs


def test_ran1_minimal_standard_and_shuffle(oracle_lib):
    xs = lcg(10000)
    assert xs[10000] == 1043618065  # Park & Miller (1988) check value
    o = Oracle()
    o.seti("ranseed", -1)
    n = 3000
    outs = []
    for k in range(n):
        v = o.ran1()
        outs.append(int(round(v * IM)))
        assert o.geti("ranseed") == xs[40 + k + 1]
    # outputs above RNMX = 1 - 1.2e-7 are clamped (ran1.f: min(AM*iy, RNMX))
    rnmx = int(round((1.0 - 1.2e-7) * IM))
    clamped = [v for v in outs if v == rnmx]
    kept = set(v for v in outs if v != rnmx)
    assert len(kept) == n - len(clamped)
    pool = set(xs[9:40 + n + 1])
    assert kept <= pool
    rest = pool - kept
    # the 32 table entries plus the clamped draws, which all lie above RNMX
    assert len(rest) == 32 + len(clamped)
    assert sum(1 for v in rest if v > (1.0 - 1.2e-7) * IM) >= len(clamped)
    # seed 0 behaves as seed 1 (idum = max(-idum, 1))
    o2 = Oracle()
    o2.seti("ranseed", 0)
    assert int(round(o2.ran1() * IM)) == outs[0]


def _ref(v_over, seed=-4321):
    v = dict(REFERENCE_SIGA)
    v.update(v_over)
    o = Oracle()
    o.readprms(siga_from_values(v), seed)
    return o, v


def test_pureepi_lattice_and_xsides(oracle_lib):
    o, v = _ref(dict(nc=(3, 4, 2), nlj=2, nrand=50))
    o.initranions()
    o.initdata()
    X, V, _ = o.state()
    a = 5.431073e-10
    nsg = 8 * 3 * 4 * 2
    assert o.geti("nsg") == nsg and o.natm == nsg + 2
    assert abs(o.getd("lb1") - 3 * a) < 1e-25 and abs(o.getd("lb3") - 20 * a) < 1e-24
    ref = configs.diamond(3, 4, 2, a)
    assert np.abs(X[:nsg] - ref).max() <= 5e-10 * ref.max() * 1e-9 + 1e-19
    xs = o.xsides()
    on_side = (ref[:, 0] < 1e-10) | (ref[:, 1] < 1e-10) | \
        (ref[:, 0] > ref[:, 0].max() - 1e-10) | (ref[:, 1] > ref[:, 1].max() - 1e-10)
    np.testing.assert_array_equal(xs[:nsg] > 0, on_side)
    at, ms = o.types()
    assert (at[:nsg] == 1).all() and (at[nsg:] == 3).all()
    # ions stored at ionz * Lb(3)/10 with lateral separation >= sigma
    assert np.allclose(X[nsg:, 2], 5.5 * o.getd("lb3") / 10)


def test_randvel_momentum_and_temperature(oracle_lib):
    o, v = _ref(dict(nc=(8, 8, 4), nlj=1, nrand=50))
    o.initranions()
    o.initdata()
    X, V, _ = o.state()
    at, ms = o.types()
    nsg = o.geti("nsg")
    p = (ms[:nsg, None] * V[:nsg]).sum(0)
    assert np.abs(p).max() <= 1e-12 * MSI * v["uo"] * nsg
    T = (ms[:nsg] * (V[:nsg] ** 2).sum(1)).sum() / 3 / nsg / KB
    T0 = MSI * v["uo"] ** 2 / (12 * KB)
    assert abs(T - T0) <= 0.05 * T0
    assert np.all(V[nsg:] == 0)


def test_initranions_gaussian_six_digits(oracle_lib):
    o, v = _ref(dict(nc=(2, 2, 2), nlj=1, nrand=10000))
    o.initranions()
    xr = o.xrand()
    s = v["fwhm"] / 2.35 * 1e-9
    assert abs(xr[:, 0].std() - s) <= 0.03 * s and abs(xr[:, 1].std() - s) <= 0.03 * s
    for x in xr[:50].ravel():
        assert float("%.5E" % x) == x


def _free_system(n=20, seed=2):
    rng = np.random.default_rng(seed)
    L = (200e-10, 200e-10, 200e-10)
    X = rng.random((n, 3)) * 150e-10 + 20e-10
    # spread: keep all pairs beyond the Moliere list cutoff
    X[:, 0] = np.arange(n) * 9e-10 + 5e-10
    V = rng.normal(0, 2e3, (n, 3))
    at = np.ones(n, np.int32)
    return X, V, at, L


def _setup_run(o, v):
    o.set_run(siga_from_values(v))


def test_free_flight_closed_form(oracle_lib):
    X, V, at, L = _free_system()
    n = len(X)
    o = gc.oracle_system(X, at, L, n, 0, V=V)
    v = dict(REFERENCE_SIGA)
    v.update(ts_i=1e-15, atlist_i=10 ** 6, ntlist_i=10 ** 6, dslist_i=10 ** 6, sidewidth=0.0)
    _setup_run(o, v)
    o.set_xsides(np.zeros(n, np.int32))
    o.set_xi(X)
    o.set_force(np.zeros((n, 3)))
    for lt in range(1, 41):
        o.onestep(2, 300.0, lt)
    Xn, Vn, _ = o.state()
    expect = np.mod(X + 40 * 1e-15 * V, np.array(L))
    assert np.abs(Xn - expect).max() <= 1e-21
    np.testing.assert_array_equal(Vn, V)


def test_energy_conservation_nve(oracle_lib):
    Xc = configs.perturbed_crystal(3, 0.05e-10, 4)
    n = len(Xc)
    L = (3 * configs.A_SI,) * 3
    V = configs.maxwell_boltzmann(np.full(n, MSI), 300.0, 5)
    at = np.ones(n, np.int32)
    o = gc.oracle_system(Xc, at, L, n, 0, V=V)
    v = dict(REFERENCE_SIGA)
    v.update(ts_i=2e-16, atlist_i=10 ** 6, ntlist_i=10, dslist_i=10 ** 6, sidewidth=0.0)
    _setup_run(o, v)
    o.set_xsides(np.zeros(n, np.int32))
    o.set_xi(Xc)
    o.tint_init()

    def etot():
        Xs, Vs, _ = o.state()
        U, _ = o.si_potential(Xs)
        return U + 0.5 * MSI * (Vs ** 2).sum()

    e0 = etot()
    ke0 = 1.5 * n * KB * 300.0
    for lt in range(1, 301):
        o.onestep(2, 300.0, lt)
    assert abs(etot() - e0) <= 2e-3 * ke0


def test_side_thermostat_factor_and_mask(oracle_lib):
    n = 6
    L = (100e-10, 100e-10, 100e-10)
    T = L[2] / 10
    # 0: x-side; 1: y-side with z below outz(1); 2: y-side, z inside (not
    # selected: .and. binds before .or.); 3: interior; 4: x-side Ga (never);
    # 5: y-side above outz(4)
    X = np.array([[5e-10, 50e-10, 50e-10], [50e-10, 5e-10, 0.5 * T],
                  [50e-10, 5e-10, 5.0 * T], [50e-10, 50e-10, 50e-10],
                  [5e-10, 60e-10, 60e-10], [50e-10, 95e-10, 9.9 * T]])
    at = np.array([1, 1, 1, 1, 3, 1], np.int32)
    T0 = 150.0
    ms = gc.masses(at)
    vmag = np.sqrt(3 * KB * T0 / ms)
    V = np.zeros((n, 3)); V[:, 0] = vmag
    o = gc.oracle_system(X, at, L, 5, 1, V=V)
    v = dict(REFERENCE_SIGA)
    v.update(ts_i=1e-15, tau_i=2e-15, sidewidth=1.5)
    _setup_run(o, v)
    o.controlsidetemp(300.0)
    _, Vn, _ = o.state()
    s = math.sqrt(1 + 0.5 * (300.0 / T0 - 1))
    sel = [0, 1, 5]
    for k in range(n):
        ratio = Vn[k, 0] / V[k, 0]
        if k in sel:
            assert abs(ratio - s) <= 1e-12
        else:
            assert ratio == 1.0


def test_regime_switch_on_max_ke(oracle_lib):
    X, V, at, L = _free_system(4)
    n = len(X)
    o = gc.oracle_system(X, at, L, n, 0, V=V)
    v = dict(REFERENCE_SIGA)
    _setup_run(o, v)
    o.seti("atom_is_fast", 1)
    ke_ev = 0.5 * MSI * (V ** 2).sum(1).max() / 1.60217646e-19
    v2 = dict(v, ke_limit_1=ke_ev * 0.999)
    _setup_run(o, v2)
    o.seti("atom_is_fast", 1)
    o.calc_atm()
    assert o.getd("ts") == v["ts_i"] and o.geti("atom_is_fast") == 1
    v3 = dict(v, ke_limit_1=ke_ev * 1.001)
    _setup_run(o, v3)
    o.seti("atom_is_fast", 1)
    o.calc_atm()
    assert o.getd("ts") == v["ts_r"] and o.geti("atom_is_fast") == 0
    assert o.geti("ntlist") == v["ntlist_r"] and o.geti("atm_out") == v["atm_out_r"]


def test_sputter_layer_rules(oracle_lib):
    """frzsptrdatms (temp.f90:99-127) with windows in units of T = Lb(3)/10:
    (outz1, outz2) -> moved to atomz1, (outz3, outz4) -> moved to atomz2, both
    with V = F = 0; within 0.3 T of atomz1, atomz2 or ionz -> V = F = 0 only;
    everything else untouched.  Numbers worked by hand for Lb(3) = 100 A."""
    A = 1e-10
    z = np.array([20.0, 80.0, 46.0, 64.0, 57.0, 10.0, 40.0, 99.0, 13.1, 34.9]) * A
    n = len(z)
    X = np.column_stack([np.arange(n) * 5 * A + 2 * A, np.full(n, 50 * A), z])
    L = (60 * A, 100 * A, 100 * A)
    at = np.ones(n, np.int32)
    V = np.full((n, 3), 123.0)
    o = gc.oracle_system(X, at, L, n, 0, V=V)
    v = dict(REFERENCE_SIGA)  # outz (1.3, 3.5, 7.5, 9.8), atomz (4.5, 6.5), ionz 5.5
    assert tuple(v["outz"]) == (1.3, 3.5, 7.5, 9.8) and tuple(v["atomz"]) == (4.5, 6.5)
    o.set_run(siga_from_values(v))
    o.set_force(np.full((n, 3), 7.0))
    o.frzsptrdatms()
    Xn, Vn, Fn = o.state()
    moved = {0: 45.0, 1: 65.0, 8: 45.0, 9: 45.0}
    zeroed = {0, 1, 2, 3, 4, 8, 9}
    for k in range(n):
        exp_z = moved.get(k, z[k] / A)
        assert abs(Xn[k, 2] / A - exp_z) < 1e-9, (k, Xn[k, 2] / A, exp_z)
        assert Xn[k, 0] == X[k, 0] and Xn[k, 1] == X[k, 1]
        if k in zeroed:
            assert np.all(Vn[k] == 0) and np.all(Fn[k] == 0), k
        else:
            assert np.all(Vn[k] == 123.0) and np.all(Fn[k] == 7.0), k
