"""Cascade statistics: mixed precision against FP64 on a batched ensemble.

Individual cascades are chaotic, so mixed and FP64 runs of the same impact
part ways within a few hundred steps; what must agree are the ensemble
statistics.  A seeded ensemble of 32 small Si(100) slabs, each hit by one Ga
ion (1-3 keV, 0-30 deg), runs 1500 MD steps in both precisions.  Sputter
yield, ion projected range and displaced-atom count (> 1.5 A from the
initial site) agree when the difference of the two means is within 3.5
standard errors of that difference (Welch).  The runs are deterministic, so
the outcome does not vary between runs.

The check has power: the same ensemble in FP64 with the SW bond energy
epsilon raised by half (stiffer, more strongly bound Si) moves the displaced
count by more than 3.5 standard errors.
"""
import math

import numpy as np
import pytest

from fib_md_amd import configs

pytestmark = pytest.mark.gpu

STEPS = 1500
Z_MAX = 3.5


def _run(ens, precision, params=None):
    from fib_md_amd import FibGPU
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"]) for s in ens], precision=precision,
               params=params)
    for k, s in enumerate(ens):
        g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"], inst=k)
    g.md_begin([s["surface_z"] for s in ens], [s["bottom_z"] for s in ens])
    for k, s in enumerate(ens):
        g.md_fire(k, 1, s["ion_x"], s["ion_v"])
    g.md_run(STEPS)
    reps = [g.md_report(k) for k in range(len(ens))]
    g.close()
    y = np.array([r["removed"] for r in reps], float)
    d = np.array([r["displaced"] for r in reps], float)
    rg = np.array([(s["surface_z"] - r["ion_z"]) * 1e10 if r["ion_z"] < 1e299 else np.nan
                   for s, r in zip(ens, reps)])
    return {"yield": y, "displaced": d, "range_A": rg}, reps


def _welch_z(a, b):
    a = a[np.isfinite(a)]
    b = b[np.isfinite(b)]
    se = math.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return (a.mean() - b.mean()) / se if se > 0 else 0.0


def test_mixed_matches_fp64_cascade_statistics():
    ens = configs.ensemble(32, seed=4000, nmin=4e3, nmax=8e3, emin=1e3, emax=3e3,
                           polar_max=30.0)
    f64, r64 = _run(ens, "fp64")
    mix, rmx = _run(ens, "mixed")
    print(f"\n{'':10s} {'fp64 mean':>10s} {'mixed mean':>11s} {'z':>6s}", flush=True)
    for key in ("yield", "range_A", "displaced"):
        z = _welch_z(mix[key], f64[key])
        print(f"{key:10s} {np.nanmean(f64[key]):10.3f} {np.nanmean(mix[key]):11.3f} {z:6.2f}",
              flush=True)
        assert abs(z) < Z_MAX, (key, z)
    # the cascades did happen: atoms displaced and sputtered, ions stopped
    # below the surface
    assert f64["displaced"].mean() > 20 and mix["displaced"].mean() > 20
    assert f64["yield"].sum() > 0 and mix["yield"].sum() > 0
    assert np.nanmean(f64["range_A"]) > 5 and np.nanmean(mix["range_A"]) > 5
    # every instance advanced the same number of steps
    assert all(r["steps"] == STEPS for r in r64 + rmx)
    # power: a different potential is told apart
    from fib_md_amd.gpu import default_params
    p = default_params()
    p.eps_si *= 1.5
    stiff, _ = _run(ens, "fp64", params=p)
    z = _welch_z(stiff["displaced"], f64["displaced"])
    print(f"displaced, eps x1.5 vs fp64: {np.mean(stiff['displaced']):.1f} vs "
          f"{np.mean(f64['displaced']):.1f}, z = {z:.2f}", flush=True)
    assert abs(z) > Z_MAX, z
