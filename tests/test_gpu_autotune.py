"""The MD-step list autotuner (fib_md_amd/autotune.py).

The list radius factor only moves time: with every pair within rc listed
and rc beyond the SW cutoff, the mixed path's forces (analytic pair term)
are the same for any factor, up to the summation order of the FP32 terms.
The tuner runs every candidate on the GPU and returns one of them.
"""
import numpy as np
import pytest

from fib_md_amd import configs
from tests import gpu_common as gc

pytestmark = pytest.mark.gpu


def _force(X, at, L, nsg, nlj, listfac):
    from fib_md_amd import FibGPU
    from fib_md_amd.gpu import default_params
    p = default_params()
    p.listfac = listfac
    g = FibGPU([(len(X), nsg, nlj, L)], precision="mixed", params=p)
    g.set_state(X, atype=at)
    g.build_nlist()
    r = g.force(impact=nlj, energy=True)
    g.close()
    return r


def test_listfac_leaves_mixed_forces_unchanged():
    X, at, L, nsg, nlj = gc.crystal_with_ions(6, 0.15e-10, 5, 3)
    a = _force(X, at, L, nsg, nlj, 1.10)
    b = _force(X, at, L, nsg, nlj, 1.25)
    emax, erms = gc.rel_err(b["F"], a["F"])[:2]
    print(f"\nlistfac 1.25 vs 1.10: max rel {emax:.2e}, rms {erms:.2e}", flush=True)
    assert emax < 1e-5 and erms < 1e-6
    assert abs(b["utot"] - a["utot"]) <= 1e-6 * abs(a["utot"])


def test_tune_listfac_small_batch():
    from fib_md_amd import autotune
    ens = configs.ensemble(4, seed=5000, nmin=4e3, nmax=6e3, emin=500.0, emax=1000.0)
    best, table = autotune.tune_listfac(ens, "mixed", candidates=(1.10, 1.20), steps=6, warmup=2)
    print(f"\n{table} -> {best}", flush=True)
    assert best in (1.10, 1.20)
    assert len(table) == 2 and all(ms is not None and ms > 0 for _, ms, _ in table)


def test_tiles_leave_trajectory_unchanged():
    """Launch tiles change no result: an MD run of a cascade in non-default
    tiles (vertex-pass stage of 128 legs, 2 loads in flight; 16 cells per
    list-build wavefront) equals the default tiles' run bit for bit, in both
    precisions; the tuner returns one of its candidates and leaves the handle
    in it."""
    from fib_md_amd import FibGPU, autotune
    s = configs.si_slab(5, 61, nz=4)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 2000.0, 7.0, 30.0,
                              np.random.default_rng(61))
    for prec in ("fp64", "mixed"):
        out = []
        for tiles in ((0, 0, 0), (128, 2, 16), (256, 2, 4)):
            g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"])], precision=prec)
            g.set_tiles(*tiles)
            g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"])
            g.md_begin([s["surface_z"]], [0.0])
            g.md_fire(0, 1, x, v)
            g.md_run(150)
            out.append((*g.state()[:2], g.md_report(0)))
            if tiles == (256, 2, 4):
                best, table = autotune.tune_tiles(g, steps=4, warmup=1)
                assert best in [t for t, _ in table] and g.tiles() == best
                assert len(table) == 6
            g.close()
        for X, V, r in out[1:]:
            np.testing.assert_array_equal(X, out[0][0])
            np.testing.assert_array_equal(V, out[0][1])
            assert r == out[0][2]
    from fib_md_amd.gpu import GPUError
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"])], precision="mixed")
    with pytest.raises(GPUError, match="tri_ns"):
        g.set_tiles(64, 4, 8)
    g.close()
