"""Autotuner of the MD-step mode's list parameters and launch tiles.

The list radius factor `listfac` (rc = listfac * a * sigma) sets both the
cell size of the list build and the skin rc - 0.999 a sigma that decides
how often an instance rebuilds (when the two largest displacements since
the build add up to more than the skin).  A larger factor means fewer
rebuilds but more candidates per build and longer lists for the pair pass.  The best value depends on the workload (instance sizes,
ion energies, precision), so it is measured: each candidate runs a short
stretch of the actual batch on the GPU and the fastest steps win.  Lists
hold every pair within rc and rc lies beyond the SW cutoff, so the mixed
path's forces (analytic pair term) are the same for every candidate up to
the summation order; the FP64 path's tabulated pair derivative spans
2.5 * listfac * a, so its interpolation error (~1e-9 relative) moves with
the factor.

The reference-mode path (fib_step) keeps the reference's factor 1.1: its
list sets are part of the parity contract.

The launch tiles (fib_set_tiles: legs per stage run of the vertex pass and
neighbour loads in flight per lane, cells per wavefront of the list build)
change no result, bit for bit, only the time; tune_tiles measures them on a
live handle, one after the other.
"""
from __future__ import annotations

import time

from .gpu import FibGPU, GPUError, default_params

CANDIDATES = (1.10, 1.15, 1.20, 1.25, 1.30)
# (legs per stage run, loads in flight per lane) of the vertex pass; cells
# per wavefront of the list build
TRI_TILES = ((256, 4), (256, 2), (128, 4), (128, 2))
NB_CELLS = (8, 4, 16)


def md_handle(instances, precision, params=None, use_graph=True):
    """A handle with the instances loaded, in MD-step mode, ions fired."""
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"]) for s in instances],
               precision=precision, params=params)
    for k, s in enumerate(instances):
        g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"], inst=k)
    g.md_begin([s["surface_z"] for s in instances], [s["bottom_z"] for s in instances],
               use_graph=use_graph)
    g.md_fire_many(list(range(len(instances))), [1] * len(instances),
                   [s["ion_x"] for s in instances], [s["ion_v"] for s in instances])
    return g


def tune_listfac(instances, precision="mixed", candidates=CANDIDATES, steps=60, warmup=20):
    """Time `steps` MD steps (after `warmup`) of the batch for each candidate
    listfac; returns (best listfac, [(listfac, ms per step or None, note)]).
    A candidate whose lists outgrow their width is skipped."""
    table = []
    for lf in candidates:
        p = default_params()
        p.listfac = float(lf)
        g = None
        try:
            g = md_handle(instances, precision, params=p)
            g.md_run(warmup)
            t0 = time.perf_counter()
            g.md_run(steps)  # returns after the stream has drained
            t = time.perf_counter() - t0
            table.append((float(lf), t * 1e3 / steps, ""))
        except GPUError as e:
            table.append((float(lf), None, str(e)))
        finally:
            if g is not None:
                g.close()
    ok = [r for r in table if r[1] is not None]
    if not ok:
        raise RuntimeError(f"no listfac candidate ran: {table}")
    best = min(ok, key=lambda r: r[1])[0]
    return best, table


def _ms_per_step(g, steps, warmup):
    g.md_run(warmup)
    t0 = time.perf_counter()
    g.md_run(steps)  # returns after the stream has drained
    return (time.perf_counter() - t0) * 1e3 / steps


def tune_tiles(g, steps=30, warmup=5, tri_tiles=TRI_TILES, nb_cells=NB_CELLS):
    """Launch tiles of a live handle in MD-step mode (ions fired, mid-run):
    first the vertex pass's (stage size, loads in flight), then the list
    build's cells per wavefront with the best of those, each timed over
    `steps` MD steps after `warmup`.  Leaves the handle in the best tiles and
    returns (best (tri_ns, tri_batch, nb_cells), [(tiles, ms per step)])."""
    table = []
    for ns, tb in tri_tiles:
        g.set_tiles(ns, tb, nb_cells[0])
        table.append(((ns, tb, nb_cells[0]), _ms_per_step(g, steps, warmup)))
    ns, tb, _ = min(table, key=lambda r: r[1])[0]
    for nc in nb_cells[1:]:
        g.set_tiles(ns, tb, nc)
        table.append(((ns, tb, nc), _ms_per_step(g, steps, warmup)))
    best = min(table, key=lambda r: r[1])[0]
    g.set_tiles(*best)
    return best, table
