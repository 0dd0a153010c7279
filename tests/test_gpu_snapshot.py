"""Asynchronous snapshots and the split run (fib_snapshot_begin / _wait,
fib_md_run_begin / _end).

- A run split into begin / end follows fib_md_run bit for bit, also when
  narrow lists overflow inside the enqueued steps (the halt, growth and
  replay then happen in _end).
- A snapshot ordered before a run holds the state at that point, bit for
  bit equal to fib_get_state there, although the copy arrives while the
  later steps run; the two slots hold two points of the trajectory.
- While steps are in flight, calls that read or change the state are
  refused with FIB_ERR_STATE, and the run still ends normally.
"""
import numpy as np
import pytest

from fib_md_amd import configs

pytestmark = pytest.mark.gpu


def _handle(s, precision="fp64", widths=None):
    from fib_md_amd import FibGPU
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"])], precision=precision)
    if widths:
        g.set_list_width(*widths)
    g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"])
    g.md_begin([s["surface_z"]], [0.0])
    return g


def _cascade():
    s = configs.si_slab(4, 31, nz=4, T=0.0)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 3000.0, 10.0, 20.0,
                              np.random.default_rng(31))
    return s, x, v


def test_run_begin_end_equals_run_with_overflow():
    s, x, v = _cascade()
    ref = _handle(s, widths=(64, 64))
    ref.md_fire(0, 1, x, v)
    ref.md_run(300)
    Xr, Vr, _ = ref.state()
    rr = ref.md_report(0)
    ref.close()
    g = _handle(s, widths=(16, 4))
    g.md_fire(0, 1, x, v)
    for n in (100, 150, 50):
        g.md_run_begin(n)
        g.md_run_end()
    X, V, _ = g.state()
    r = g.md_report(0)
    w = g.list_width()
    g.close()
    print(f"\nsplit run: lists grew to {w}, builds {r['builds']}", flush=True)
    assert w[0] > 16 and w[1] > 4
    np.testing.assert_array_equal(X, Xr)
    np.testing.assert_array_equal(V, Vr)
    assert r == rr


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_snapshot_holds_state_at_begin(precision):
    s, x, v = _cascade()
    g = _handle(s, precision)
    g.md_fire(0, 1, x, v)
    g.md_run(40)
    X0, V0, _ = g.state()
    g.snapshot_begin(0, slot=0)
    g.md_run_begin(60)
    # slot 0 arrives while the 60 steps run
    Xs, Vs = g.snapshot_wait(0)
    g.md_run_end()
    np.testing.assert_array_equal(Xs, X0)
    np.testing.assert_array_equal(Vs, V0)
    X1, V1, _ = g.state()
    assert not np.array_equal(X1, X0)
    # two slots, two points of the trajectory, waited for in either order
    g.snapshot_begin(0, slot=1)
    g.md_run(30)
    X2, V2, _ = g.state()
    g.snapshot_begin(0, slot=0)
    g.md_run_begin(30)
    Xb, Vb = g.snapshot_wait(0)
    Xa, Va = g.snapshot_wait(1)
    g.md_run_end()
    g.close()
    np.testing.assert_array_equal(Xa, X1)
    np.testing.assert_array_equal(Va, V1)
    np.testing.assert_array_equal(Xb, X2)
    np.testing.assert_array_equal(Vb, V2)


def test_calls_refused_while_steps_in_flight():
    from fib_md_amd.gpu import GPUError
    s, x, v = _cascade()
    g = _handle(s)
    g.md_fire(0, 1, x, v)
    g.md_run_begin(50)
    for call in (lambda: g.state(), lambda: g.md_report(0), lambda: g.md_energy(),
                 lambda: g.md_fire(0, 1, x, v), lambda: g.md_run(1),
                 lambda: g.snapshot_begin(0, 0), lambda: g.compute(),
                 lambda: g.build_nlist(), lambda: g.set_tiles(128, 2, 4),
                 lambda: g.set_state(s["X"], V=s["V"], atype=s["atype"])):
        with pytest.raises(GPUError, match="in flight"):
            call()
    g.md_run_end()
    assert g.md_report(0)["steps"] == 50
    with pytest.raises(GPUError, match="no snapshot"):
        g.snapshot_wait(1)
    g.close()
