"""Pins of the oracle's knock, temperature, thermostat, freeze and of the order
of the steps inside onestep / TI (CPU).

Each expectation is worked out by hand from the physics, not by calling the
oracle twice: the Stillinger-Weber pair force comes from the published
potential (Stillinger & Weber, Phys. Rev. B 31, 5262 (1985)) written out
here, velocities and positions from the velocity-Verlet formulas, and the
scenarios are built so that a plausible mistake in the order of onestep
(thermostat after the drift, the sputter freeze after the velocity update,
the list rebuild after the force, the new timestep in the drift) moves the
result by far more than the tolerance.

- knock (knock.f90:21-55): 1/2 m v^2 = eV, direction (polar phiz, azimuth
  phixy, pointing down), start point above the impact spot drawn by ran1;
- temperature / adjusttemp (temp.f90:160-205): the divisor is Nsg even for
  frozen and resting atoms (quirk Q12), ions are left out, the Berendsen
  factor sqrt(1 + Ts/Tau (Ttar/T - 1));
- freezsidesN (tools.f90:38-51): pinned atoms back to Xi with V = 0;
- onestep (timeint.f90:158-242): equilibration thermostat before the drift,
  freeze before rebox, frzsptrdatms before the force and velocity update
  (Q7), rebuild before the force, the regime switch after the drift but
  before the velocity update (Q6);
- TI (timeint.f90:123-155): the end tests and the step counter.
"""
import math

import numpy as np

from oracle.oracle import Oracle, REFERENCE_SIGA, siga_from_values
from tests import gpu_common as gc

A = 1e-10
KB = 1.381e-23
EC = 1.60217646e-19
MSI = 46.637063e-27
MGA = 1.157777417e-25
# Stillinger-Weber silicon (Phys. Rev. B 31, 5262): A, B, p = 4, q = 0, a
SW_A, SW_B, SW_a = 7.049556277, 0.6022245584, 1.8
EPS, SIG = 3.473928e-19, 2.0951e-10


def sw_pair_force(r):
    """-dphi2/dr (N) of the SW pair term at r (m); positive = repulsive."""
    s = r / SIG
    if s >= SW_a:
        return 0.0
    e = math.exp(1.0 / (s - SW_a))
    dphi_ds = EPS * SW_A * (-4.0 * SW_B * s ** -5 * e
                            - (SW_B * s ** -4 - 1.0) * e / (s - SW_a) ** 2)
    return -dphi_ds / SIG


def pair_forces(X, pairs):
    """Forces (N) of the listed SW pairs at positions X (no periodic image)."""
    F = np.zeros_like(X)
    for i, j in pairs:
        d = X[i] - X[j]
        r = float(np.sqrt((d * d).sum()))
        f = sw_pair_force(r) * d / r
        F[i] += f
        F[j] -= f
    return F


def _system(X, V, at=None, nsg=None, nlj=0, L=(100 * A,) * 3, run=None, xsides=None, Xi=None):
    n = len(X)
    at = np.ones(n, np.int32) if at is None else np.asarray(at, np.int32)
    o = gc.oracle_system(X, at, L, n if nsg is None else nsg, nlj, V=V)
    v = dict(REFERENCE_SIGA)
    v.update(sidewidth=0.0, ntlist_i=10 ** 6, atlist_i=10 ** 6, dslist_i=10 ** 6,
             atm_out_i=10 ** 6, res_out=-1)
    v.update(run or {})
    o.set_run(siga_from_values(v))
    o.set_xsides(np.zeros(n, np.int32) if xsides is None else xsides)
    o.set_xi(X if Xi is None else Xi)
    return o, v


# ---------------------------------------------------------------------------
def test_knock_energy_direction_and_start_point(oracle_lib):
    """knock.f90:21-55.  Two oracles with the same seed and call sequence:
    one fires the ion, the other draws the same ran1 value, which gives the
    impact spot index NINT(u Nrand)."""
    for phiz, phixy in ((0.0, 0.0), (45.0, 90.0), (60.0, 180.0)):
        vals = dict(REFERENCE_SIGA)
        vals.update(nc=(2, 2, 2), nlj=2, nrand=500, eV=2000.0, knockz=1.25,
                    phiz=phiz, phixy=phixy)
        o1, o2 = Oracle(), Oracle()
        for o in (o1, o2):
            o.readprms(siga_from_values(vals), -777)
            o.initranions()
            o.initdata()
        u = o2.ran1()
        o1.knock(1)
        X, V, _ = o1.state()
        nsg = o1.geti("nsg")
        ion = nsg  # knock(1) places atom Nsg+1
        v = V[ion]
        # energy: 1663.6334 m/s per sqrt(eV) is sqrt(2 e / m_Ga) to 8 digits
        e_ev = 0.5 * MGA * float((v * v).sum()) / EC
        assert abs(e_ev / 2000.0 - 1.0) < 1e-7, e_ev
        d = v / np.linalg.norm(v)
        want = {(0.0, 0.0): (0.0, 0.0, -1.0),
                (45.0, 90.0): (0.0, 0.7071067811865476, -0.7071067811865476),
                (60.0, 180.0): (-0.8660254037844386, 0.0, -0.5)}[(phiz, phixy)]
        assert np.allclose(d, want, atol=1e-12), (phiz, phixy, d)
        ri = max(1, int(math.floor(u * 500 + 0.5)))
        xr = o1.xrand()
        L1, L2, L3 = o1.getd("lb1"), o1.getd("lb2"), o1.getd("lb3")
        T = L3 / 10
        h = (1.25 - 1.0) * T  # height of the start point above the slab top
        # the reference's lateral offset: tan(phiz) h along the azimuth, times
        # sin(phiz) once more (knock.f90:31-33)
        off = {(0.0, 0.0): (0.0, 0.0),
               (45.0, 90.0): (0.0, h * 1.0 * 0.7071067811865476),
               (60.0, 180.0): (-h * 1.7320508075688772 * 0.8660254037844386, 0.0)}[(phiz, phixy)]
        assert abs(X[ion, 0] - (xr[ri - 1, 0] + L1 / 2 - off[0])) < 1e-12 * L1
        assert abs(X[ion, 1] - (xr[ri - 1, 1] + L2 / 2 - off[1])) < 1e-12 * L2
        assert X[ion, 2] == 1.25 * L3 / 10.0
        # the other ion stays where it is stored
        X2, V2, _ = o2.state()
        np.testing.assert_array_equal(X[nsg + 1], X2[nsg + 1])
        np.testing.assert_array_equal(V[nsg + 1], 0.0)


def test_temperature_divisor_and_berendsen_factor(oracle_lib):
    """temp.f90:183-205 and :160-179.  Ten SW atoms (one Ge), of which six
    move with speed u, four rest (frozen or stored), plus a fast Ga ion:
    T = sum m v^2 / (3 Nsg kB) with Nsg = 10 as divisor (Q12) and without
    the ion.  u is chosen so that T = 300 K exactly."""
    n = 11
    X = np.column_stack([np.arange(n) * 9 * A + 5 * A, np.full(n, 50 * A), np.full(n, 50 * A)])
    at = np.ones(n, np.int32)
    at[3] = 2
    at[10] = 3
    ms = gc.masses(at)
    V = np.zeros((n, 3))
    moving = [0, 1, 2, 3, 4, 5]
    # sum over moving atoms of m u^2 = 3 * 10 * kB * 300
    u2 = 3 * 10 * KB * 300.0 / ms[moving].sum()
    for k, c in zip(moving, (0, 1, 2, 0, 1, 2)):
        V[k, c] = math.sqrt(u2) * (1 if k % 2 else -1)
    V[10] = (3e4, -2e4, -1e5)
    o, _ = _system(X, V, at=at, nsg=10, nlj=1)
    assert abs(o.temperature() - 300.0) < 1e-9
    # Berendsen: Ts/Tau = 0.5, Ttar/T = 3 -> s = sqrt(2); the ion keeps V
    o.setd("ts", 1e-15)
    o.setd("Tau", 2e-15)
    o.adjusttemp(300.0, 900.0)
    _, Vn, _ = o.state()
    s = math.sqrt(2.0)
    for k in range(10):
        for c in range(3):
            if V[k, c] != 0:
                assert abs(Vn[k, c] / V[k, c] - s) < 1e-14
            else:
                assert Vn[k, c] == 0
    np.testing.assert_array_equal(Vn[10], V[10])
    assert abs(o.temperature() - 600.0) < 1e-9


def test_freezsides_pins_marked_atoms(oracle_lib):
    """tools.f90:38-51: atoms with xsides > 0 go back to Xi with V = 0, the
    others keep X and V bit for bit."""
    rng = np.random.default_rng(3)
    n = 8
    X = rng.random((n, 3)) * 80 * A + 10 * A
    Xi = X + rng.normal(0, 0.3 * A, (n, 3))
    V = rng.normal(0, 1e3, (n, 3))
    xs = np.array([1, 0, 0, 2, 0, 1, 0, 0], np.int32)
    o, _ = _system(X, V, xsides=xs, Xi=Xi)
    o.freezsides()
    Xn, Vn, _ = o.state()
    for k in range(n):
        if xs[k] > 0:
            np.testing.assert_array_equal(Xn[k], Xi[k])
            np.testing.assert_array_equal(Vn[k], 0.0)
        else:
            np.testing.assert_array_equal(Xn[k], X[k])
            np.testing.assert_array_equal(Vn[k], V[k])


# ---------------------------------------------------------------------------
# onestep, scenario by scenario.  A Si dimer (atoms 0, 1) along x, free atoms
# elsewhere; forces of the previous step are set by hand.

def test_onestep_equilibration_thermostat_freeze_rebox(oracle_lib):
    """whichctrl 0: V is scaled by the Berendsen factor of the previous
    temperature before the drift, temp is recomputed from the scaled V,
    pinned atoms go back to their sites after the drift, then rebox wraps."""
    X = np.array([[20.0, 30.0, 40.0], [60.0, 20.0, 70.0], [40.0, 80.0, 20.0],
                  [99.99, 50.0, 50.0]]) * A
    V = np.array([[1000.0, -500.0, 200.0], [0.0, 0.0, 0.0], [-300.0, 800.0, 0.0],
                  [2e4, 0.0, 0.0]])
    Xi = X.copy()
    X[1] += 1.0 * A  # pinned atom away from its site
    xs = np.array([0, 1, 0, 0], np.int32)
    o, v = _system(X, V, xsides=xs, Xi=Xi, run=dict(ts_i=1e-15, tau_i=2e-15))
    o.set_force(np.zeros((4, 3)))
    t_prev = 100.0
    o.setd("temp", t_prev)
    o.setd("Tau", 2e-15)
    o.onestep(0, 3 * t_prev, 1)
    Xn, Vn, _ = o.state()
    s = math.sqrt(1 + 0.5 * (3.0 - 1.0))
    ts = 1e-15
    for k in (0, 2):
        np.testing.assert_allclose(Vn[k], s * V[k], rtol=1e-14)
        np.testing.assert_allclose(Xn[k], X[k] + s * V[k] * ts, rtol=1e-14)
    # back on its site, then through rebox's fmod(X + 2 Lb, Lb), which rounds
    np.testing.assert_array_equal(Xn[1], np.fmod(Xi[1] + 2 * 100 * A, 100 * A))
    assert np.abs(Xn[1] - Xi[1]).max() < 1e-23
    np.testing.assert_array_equal(Vn[1], 0.0)
    x3 = 99.99 * A + s * 2e4 * ts - 100 * A  # left the box through x = Lb
    assert abs(Xn[3, 0] - x3) < 1e-22 and 0 < Xn[3, 0] < 1 * A
    # temp of the scaled velocities (before the pinned atom's reset; it was at rest)
    T = sum(MSI * (s * V[k]) @ (s * V[k]) for k in range(4)) / 3 / 4 / KB
    assert abs(o.getd("temp") - T) < 1e-9 * T
    assert abs(o.getd("time") - ts) < 1e-30


def test_onestep_sputter_freeze_before_velocity_update(oracle_lib):
    """atlist step (timeint.f90:195-203, temp.f90:99-127): the dimer drifts,
    lands in the (outz1, outz2) window, is moved to atomz1 with V = F = 0,
    and only then gets the force of its new positions: V = F' Ts / (2m)
    (quirk Q7).  A free atom outside every window drifts on untouched."""
    ts = 1e-15
    r0 = 2.6 * A
    X = np.array([[40.0 * A, 50 * A, 20 * A], [40.0 * A + r0, 50 * A, 20 * A],
                  [70 * A, 50 * A, 60 * A]])
    V = np.array([[500.0, 0.0, 0.0], [-500.0, 0.0, 0.0], [1000.0, 2000.0, -500.0]])
    o, v = _system(X, V, run=dict(ts_i=ts, atlist_i=1))
    F0 = pair_forces(X, [(0, 1)])
    o.set_force(F0)
    o.onestep(2, 300.0, 1)
    Xn, Vn, Fn = o.state()
    X1 = X + V * ts + 0.5 * F0 * ts * ts / MSI
    X1[0:2, 2] = v["atomz"][0] * 10 * A  # atomz1 * Lb(3)/10
    np.testing.assert_allclose(Xn, X1, rtol=1e-14, atol=1e-24)
    F1 = pair_forces(X1, [(0, 1)])
    np.testing.assert_allclose(Vn[0:2], 0.5 * F1[0:2] * ts / MSI, rtol=1e-7, atol=1e-9)
    assert abs(Vn[0, 0]) > 1.0  # the force of the new positions acted
    np.testing.assert_array_equal(Vn[2], V[2])
    np.testing.assert_allclose(Fn, F1, rtol=1e-7, atol=1e-20)


def test_onestep_rebuild_before_force(oracle_lib):
    """ntlist step (timeint.f90:206-211): a dimer 4.4 A apart (beyond the
    list radius, so no pair in the lists) closes to 3.2 A within one step;
    the rebuild at this step comes before the force, so the pair acts on
    the velocities at once."""
    ts = 1e-15
    X = np.array([[40.0, 50, 50], [44.4, 50, 50], [70, 20, 20]]) * A
    V = np.array([[6e4, 0.0, 0.0], [-6e4, 0.0, 0.0], [0.0, 100.0, 0.0]])
    o, _ = _system(X, V, run=dict(ts_i=ts, ntlist_i=1))
    assert o.nlists()["nlist2"].sum() == 0 or o.geti("len2") == 0
    o.set_force(np.zeros((3, 3)))
    o.onestep(2, 300.0, 1)
    Xn, Vn, Fn = o.state()
    X1 = X + V * ts
    np.testing.assert_allclose(Xn, X1, rtol=1e-14)
    assert abs(X1[1, 0] - X1[0, 0] - 3.2 * A) < 1e-22
    F1 = pair_forces(X1, [(0, 1)])
    np.testing.assert_allclose(Vn, V + 0.5 * F1 * ts / MSI, rtol=1e-9, atol=1e-6)
    assert abs(Vn[0, 0] - V[0, 0]) > 10.0  # 37 m/s: the pair acted


def test_onestep_regime_switch_timestep_order(oracle_lib):
    """dslist step in an impact (timeint.f90:165,187,214-233): time and X
    advance with Ts_i, calc_atm switches to Ts_r because the largest kinetic
    energy is below ke_limit_1, and the velocity update already uses Ts_r
    (quirk Q6); the cadences switch with it."""
    ts_i, ts_r = 1e-16, 1e-15
    r0 = 2.6 * A
    X = np.array([[40.0 * A, 50 * A, 60 * A], [40.0 * A + r0, 50 * A, 60 * A],
                  [70 * A, 20 * A, 60 * A]])
    V = np.array([[300.0, 0.0, 0.0], [-300.0, 0.0, 0.0], [0.0, 4000.0, 0.0]])
    run = dict(ts_i=ts_i, ts_r=ts_r, dslist_i=1, ke_limit_1=1000.0, ntlist_r=7, atlist_r=9,
               dslist_r=11, atm_out_r=13)
    o, v = _system(X, V, run=run)
    o.seti("atom_is_fast", 1)
    o.setd("time", 5e-15)
    F0 = pair_forces(X, [(0, 1)])
    o.set_force(F0)
    o.onestep(1, 300.0, 1)
    Xn, Vn, _ = o.state()
    X1 = X + V * ts_i + 0.5 * F0 * ts_i * ts_i / MSI
    np.testing.assert_allclose(Xn, X1, rtol=1e-14)
    F1 = pair_forces(X1, [(0, 1)])
    np.testing.assert_allclose(Vn, V + 0.5 * (F0 + F1) * ts_r / MSI, rtol=1e-9, atol=1e-9)
    # with Ts_i in the velocity update the dimer's velocity change would be 10x smaller
    assert abs((Vn[0, 0] - V[0, 0]) / (0.5 * (F0[0, 0] + F1[0, 0]) * ts_i / MSI) - 10.0) < 1e-6
    assert o.getd("time") == 5e-15 + ts_i and o.getd("restart_time") == o.getd("time")
    assert o.getd("ts") == ts_r and o.geti("atom_is_fast") == 0
    assert (o.geti("ntlist"), o.geti("atlist"), o.geti("dslist"), o.geti("atm_out")) == \
        (7, 9, 11, 13)


def test_ti_end_tests_and_step_counter(oracle_lib):
    """TI (timeint.f90:130-147): equilibration ends before a step once
    |temp - Ttar| <= Teps; an impact ends when time > impact * dti, tested
    after lt is incremented and before the step (the counter runs one
    ahead of the steps at the end of an impact)."""
    X = np.array([[20.0, 30.0, 40.0], [60.0, 20.0, 70.0]]) * A
    V = np.array([[100.0, 0.0, 0.0], [0.0, -100.0, 0.0]])
    o, v = _system(X, V, run=dict(ts_i=1e-15, teps=30.0, dti=1e-13))
    o.set_force(np.zeros((2, 3)))
    o.seti("lt", 10)
    o.setd("temp", 310.0)
    assert o.ti_iter(0, 300.0) == 1 and o.geti("lt") == 10
    o.setd("temp", 340.0)
    o.setd("Tau", 1e-13)
    assert o.ti_iter(0, 300.0) == 0 and o.geti("lt") == 11
    o.seti("impact", 2)
    o.setd("time", 2e-13 - 1e-16)
    X0, _, _ = o.state()
    assert o.ti_iter(1, 300.0) == 0 and o.geti("lt") == 12
    assert abs(o.getd("time") - (2e-13 - 1e-16 + 1e-15)) < 1e-28
    X1, _, _ = o.state()
    assert np.abs(X1 - X0).max() > 0
    assert o.ti_iter(1, 300.0) == 1 and o.geti("lt") == 13 and o.geti("atom_is_slow") == 1
    X2, _, _ = o.state()
    np.testing.assert_array_equal(X2, X1)
