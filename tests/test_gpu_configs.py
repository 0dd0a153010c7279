"""The benchmark configurations at their own sizes, through the MD-step mode
(checks that need no oracle run; the oracle follows the same integrator on
small slabs in test_gpu_md_oracle.py).

- Single impact (8x8x8-cell Si(100) slab, 4,096 Si + 1 Ga, 500 eV normal,
  300 K, seed 1001, 2000 steps), FP64 and mixed: the timestep never exceeds
  1 fs nor lets the fastest atom move more than 0.05 A; the clock is the sum
  of the steps; every removed atom sits more than 10 A above the surface,
  at rest, and no live atom does; the energy ledger stays within 1e-3 of
  the ion's energy; on the FP64 run's last configuration, every atom's
  mixed force within 1e-5 of FP64 (floor 1e-3 F_rms).
- Single large impact (60x60x80 cells, 2,304,000 Si, 30 keV at 7 deg), the
  first 300 steps in both precisions: ledger within 1e-3 of the ion energy,
  the smallest dt at most 0.05 A over the ion's launch speed (a Si recoil
  can outrun the ion by up to 2 m_Ga / (m_Ga + m_Si) = 1.43x); mixed forces
  against FP64 on the FP64 run's configuration: the norm-wise relative
  error below 1e-5, and every atom's error below 1e-5 of the larger of its
  own force and 1e-3 F_rms.
- Cumulative dose (16 replicas of 40^3 cells, 8,192,000 Si + 32 Ga), two
  sequential 2 keV impacts per replica (all replicas fired at once), shells
  on: per replica the removed count equals the atoms found removed in the
  state, cumulative yields never decrease, every ion fired brought 2 keV,
  the ledger stays within 1e-3 of the energy fired.
"""
import numpy as np
import pytest

from fib_md_amd import configs

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19
A = 1e-10


def _handle(instances, precision, **md):
    from fib_md_amd import FibGPU
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"]) for s in instances], precision=precision)
    for k, s in enumerate(instances):
        g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"], inst=k)
    g.md_begin([s["surface_z"] for s in instances], [s["bottom_z"] for s in instances], **md)
    return g


def _removal_consistent(g, s, inst=0):
    """Removed atoms are at rest more than 10 A above the surface; live atoms
    are not (the box holds 40 A of vacuum, so nothing wraps past it)."""
    X, V, _ = g.state(inst)
    r = g.md_report(inst)
    above = X[:, 2] - s["surface_z"] > 10 * A
    at_rest = np.all(V == 0.0, axis=1)
    gone = above & at_rest
    nsg = s["nsg"]
    assert int(gone[:nsg].sum()) == r["removed"], (int(gone[:nsg].sum()), r["removed"])
    assert int(gone[nsg:].sum()) == r["ions_removed"]
    assert not np.any(above & ~at_rest)
    return r


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_single_impact_2000_steps(precision):
    s = configs.si_slab(8, 1001)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 500.0, 0.0, 0.0,
                              np.random.default_rng(1001))
    print(f"\nsingle impact {precision}: {len(s['X'])} atoms", flush=True)
    g = _handle([s], precision)
    g.md_fire(0, 1, x, v)
    L0 = g.md_ledger()[0]
    worst_dx = 0.0
    # the first 100 steps one at a time: each advances the clock by the dt
    # reported before it, which respects the cap and the 0.05 A rule
    for k in range(138):
        r0 = g.md_report()
        _, V0, _ = g.state()
        g.md_run(1 if k < 100 else 50)
        r1 = g.md_report()
        assert r0["dt"] <= 1e-15 * (1 + 1e-12)
        vmax = float(np.sqrt((V0 ** 2).sum(1)).max())
        assert r0["dt"] * vmax <= 0.05 * A * (1 + 1e-9)
        worst_dx = max(worst_dx, r0["dt"] * vmax)
        if k < 100:
            assert r1["time"] - r0["time"] == pytest.approx(r0["dt"], rel=1e-12)
    r = _removal_consistent(g, s)
    L1 = g.md_ledger()[0]
    X1 = g.state()[0]
    g.close()
    assert r["steps"] == 2000
    assert 0 < r["dt_min"] <= 1e-15 and r["time"] >= 2000 * r["dt_min"] * (1 - 1e-12)
    drift = (L1 - L0) / EV
    print(f"{precision}: drift {drift:.4f} eV, removed {r['removed']}, builds {r['builds']}, "
          f"smallest dt {r['dt_min'] * 1e15:.4f} fs, largest step {worst_dx / A:.4f} A",
          flush=True)
    assert abs(drift) < 1e-3 * 500.0
    if precision == "fp64":
        _mixed_force_error_within(s, X1, impact=1)


def _mixed_force_error_within(s, X, impact):
    """Forces of one configuration in both precisions: the norm-wise error
    below 1e-5 and every atom's error below 1e-5 of the larger of its own
    force and 1e-3 F_rms."""
    from fib_md_amd import FibGPU
    F = {}
    for prec in ("fp64", "mixed"):
        g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"])], precision=prec)
        g.set_state(X, atype=s["atype"])
        g.build_nlist()
        F[prec] = g.force(impact=impact)["F"]
        g.close()
    d = np.sqrt(((F["mixed"] - F["fp64"]) ** 2).sum(1))
    mag = np.sqrt((F["fp64"] ** 2).sum(1))
    frms = np.sqrt((mag ** 2).mean())
    e = d / np.maximum(mag, 1e-3 * frms)
    rms = float(np.sqrt((e ** 2).mean()))
    norm = float(np.linalg.norm(F["mixed"] - F["fp64"]) / np.linalg.norm(F["fp64"]))
    print(f"force error mixed vs fp64: rms {rms:.2e}, norm {norm:.2e}, max {e.max():.2e}, "
          f"median {np.median(e):.2e}", flush=True)
    assert norm < 1e-5
    assert e.max() <= 1e-5


def test_big_impact_first_300_steps():
    s = configs.big_impact()
    print(f"\n30 keV impact: {len(s['X'])} atoms", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        g = _handle([s], prec)
        g.md_fire(0, 1, s["ion_x"], s["ion_v"])
        L0 = g.md_ledger()[0]
        g.md_run(300)
        r = g.md_report()
        res[prec] = dict(drift=(g.md_ledger()[0] - L0) / EV, r=r, X=g.state()[0])
        g.close()
        print(f"{prec}: drift {res[prec]['drift']:.3f} eV, dt_min {r['dt_min'] * 1e15:.5f} fs, "
              f"builds {r['builds']}", flush=True)
        v0 = float(np.linalg.norm(s["ion_v"]))
        assert 0.05 * A / (1.43 * v0) <= r["dt_min"] <= 0.05 * A / v0 * (1 + 1e-9)
        assert abs(res[prec]["drift"]) < 1e-3 * 30000.0
    # forces of the FP64 run's configuration in both precisions
    _mixed_force_error_within(s, res["fp64"]["X"], impact=1)


def test_dose_two_impacts():
    reps, shots = configs.dose_replicas(16, 40, 2)
    print(f"\ndose: {sum(len(s['X']) for s in reps)} atoms", flush=True)
    g = _handle(reps, "mixed", shell_width=3 * A, ttar=300.0, tau=1e-13)
    L0 = g.md_ledger()
    prev = np.zeros(16)
    for m in range(2):
        g.md_fire_many(list(range(16)), [m + 1] * 16, [shots[k][m][0] for k in range(16)],
                       [shots[k][m][1] for k in range(16)])
        g.md_run(600)
        rr = [_removal_consistent(g, reps[k], k) for k in range(16)]
        y = np.array([r["removed"] for r in rr], float)
        assert np.all(y >= prev)
        prev = y
        fired = np.array([r["e_fired"] for r in rr]) / EV
        # 1663.6334 sqrt(eV) m/s for Ga gives the energy to 1e-7; a fired
        # ion also books its (small) interaction energy at the start point
        assert np.all(np.abs(fired - 2000.0 * (m + 1)) < 0.5 * (m + 1))
        drift = (g.md_ledger() - L0) / EV
        print(f"impact {m + 1}: yield {y.mean():.2f} per replica, drift mean {drift.mean():.3f} eV, "
              f"max |drift| {np.abs(drift).max():.3f} eV", flush=True)
        assert np.all(np.abs(drift) < 1e-3 * fired)
    g.close()
