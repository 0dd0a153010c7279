"""GPU parity of the benchmark configurations' integrator (MD-step mode,
fib_md_begin / fib_md_fire / fib_md_run) against the oracle's restatement of
it (orc_md_begin / orc_md_fire / orc_md_step), step by step.

Two Si(100) slabs in one batch, each hit by a Ga ion, each followed by its
own oracle instance: one with Berendsen shells on the sides and bottom and a
surface atom launched upwards so that removal, the removal ledger and the
shells all act; one oblique impact.  After every step the positions agree
within 1e-18 m (FP64) and the velocities within 1e-8 of the largest; the
timestep, clock, rebuild count, removed / displaced counts and the energy
ledger's terms agree.  The mixed-precision path follows the same oracle
within a looser, stated bound over the first steps of the cascade.
"""
import numpy as np
import pytest

from fib_md_amd import configs
from tests import gpu_common as gc

pytestmark = pytest.mark.gpu

A = 1e-10
MD = dict(dt_max=1e-15, dx_max=0.05 * A, remove_height=10 * A, shell_width=3 * A, ttar=300.0,
          tau=2e-14, displaced_r=1.5 * A)


def _cases():
    out = []
    s = configs.si_slab(3, 41, nz=3)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 1000.0, 0.0, 0.0,
                              np.random.default_rng(41))
    # a surface atom lifted 9.5 A and sent up at 1e4 m/s: removed within
    # the first steps
    top = int(np.argmax(s["X"][: s["nsg"], 2]))
    s["X"][top, 2] += 9.5 * A
    s["V"][top] = (0.0, 0.0, 1.0e4)
    s.update(ion_x=x, ion_v=v, shells=True)
    out.append(s)
    s = configs.si_slab(4, 42, nz=3)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 2000.0, 30.0, 70.0,
                              np.random.default_rng(42))
    s.update(ion_x=x, ion_v=v, shells=False)
    out.append(s)
    return out


def _oracle(s, md):
    o = gc.oracle_system(s["X"], s["atype"], s["L"], s["nsg"], s["nlj"], V=s["V"])
    o.set_xsides(s["xsides"])
    o.set_xi(s["X"])
    o.md_begin(s["surface_z"], s["bottom_z"], **md)
    o.md_fire(1, s["ion_x"], s["ion_v"])
    return o


def _gpu(cases, precision):
    from fib_md_amd import FibGPU
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"]) for s in cases], precision=precision)
    for k, s in enumerate(cases):
        g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"], inst=k)
    return g


@pytest.mark.parametrize("use_graph", [False, True])
def test_md_fp64_lockstep_with_oracle(oracle_lib, use_graph):
    cases = _cases()
    # shells on, then off (a handle-wide parameter), both cases in one batch
    for shells in (True, False):
        md = dict(MD, shell_width=MD["shell_width"] if shells else 0.0)
        g = _gpu(cases, "fp64")
        g.md_begin([s["surface_z"] for s in cases], [s["bottom_z"] for s in cases],
                   use_graph=use_graph, **md)
        # build counts from the start of MD-step mode on (the oracle counts
        # the lists of its own start as its first build)
        b0 = [g.md_report(k)["builds"] - 1 for k in range(len(cases))]
        orc = []
        for k, s in enumerate(cases):
            g.md_fire(k, 1, s["ion_x"], s["ion_v"])
            orc.append(_oracle(s, md))
        worst_x = 0.0
        nsteps = 300 if use_graph else 80
        for step in range(nsteps):
            g.md_run(1)
            for k, o in enumerate(orc):
                o.md_step()
                Xo, Vo, _ = o.state()
                Xg, Vg, _ = g.state(k)
                dx = float(np.abs(Xg - Xo).max())
                worst_x = max(worst_x, dx)
                assert dx <= 1e-18, (shells, k, step, dx)
                assert np.abs(Vg - Vo).max() <= 1e-8 * np.abs(Vo).max(), (shells, k, step)
                rg, ro = g.md_report(k), o.md_report()
                assert abs(rg["dt"] - ro["dt"]) <= 1e-10 * ro["dt"], (k, step)
                assert abs(rg["time"] - ro["time"]) <= 1e-10 * ro["time"]
                assert rg["steps"] == ro["steps"]
                assert rg["builds"] - b0[k] == ro["builds"], (k, step)
                assert rg["removed"] == ro["removed"] and rg["displaced"] == ro["displaced"]
        for k, o in enumerate(orc):
            rg, ro = g.md_report(k), o.md_report()
            scale = rg["e_fired"]
            for key in ("e_bath", "e_removed", "e_fired", "ke"):
                assert abs(rg[key] - ro[key]) <= 1e-9 * scale, (shells, k, key, rg[key], ro[key])
            ep = g.md_energy()[k]
            assert abs(ep - ro["epot"]) <= 1e-9 * scale
        if shells:
            r0 = g.md_report(0)
            assert r0["removed"] >= 1 and r0["e_removed"] > 0 and r0["e_bath"] != 0.0
        print(f"shells={shells} graph={use_graph}: worst |dX| {worst_x:.2e} m over {nsteps} "
              f"steps, builds {[g.md_report(k)['builds'] - b0[k] for k in range(2)]}", flush=True)
        g.close()


def test_md_mixed_follows_oracle(oracle_lib):
    """Mixed precision (FP32 terms, FP64 integration) against the FP64 oracle
    over the first 40 steps of both cascades: positions within 1e-15 m,
    the same rebuild and removal decisions."""
    cases = _cases()
    md = dict(MD)
    g = _gpu(cases, "mixed")
    g.md_begin([s["surface_z"] for s in cases], [s["bottom_z"] for s in cases], **md)
    b0 = [g.md_report(k)["builds"] - 1 for k in range(len(cases))]
    orc = []
    for k, s in enumerate(cases):
        g.md_fire(k, 1, s["ion_x"], s["ion_v"])
        orc.append(_oracle(s, md))
    worst = 0.0
    for step in range(40):
        g.md_run(1)
        for k, o in enumerate(orc):
            o.md_step()
            Xo, _, _ = o.state()
            Xg, _, _ = g.state(k)
            dx = float(np.abs(Xg - Xo).max())
            worst = max(worst, dx)
            assert dx <= 1e-15, (k, step, dx)
            rg, ro = g.md_report(k), o.md_report()
            assert rg["builds"] - b0[k] == ro["builds"] and rg["removed"] == ro["removed"]
    print(f"mixed: worst |dX| {worst:.2e} m over 40 steps", flush=True)
    g.close()
