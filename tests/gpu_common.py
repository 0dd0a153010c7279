"""Shared setup of paired oracle / GPU systems for the parity tests."""
from __future__ import annotations

import numpy as np

from fib_md_amd import configs

MASS = {1: 46.637063e-27, 2: 120.6049350e-27, 3: 1.157777417e-25}


def masses(at):
    return np.array([MASS[int(t)] for t in at])


def oracle_system(X, atype, lb, nsg, nlj, impact=0, V=None):
    from oracle.oracle import Oracle
    o = Oracle()
    n = len(X)
    o.set_geometry(n, nsg, nlj, lb)
    o.set_state(X=X, V=np.zeros((n, 3)) if V is None else V, atype=atype, mass=masses(atype))
    o.init_nlist()
    o.mktables()
    o.seti("impact", impact)
    o.si_nlist(X)
    return o


def gpu_system(X, atype, lb, nsg, nlj, precision="fp64", V=None, xsides=None, Xi=None):
    from fib_md_amd import FibGPU
    n = len(X)
    g = FibGPU([(n, nsg, nlj, lb)], precision=precision)
    g.set_state(X, V=V, atype=atype, xsides=xsides, Xi=Xi)
    g.build_nlist()
    return g


def crystal_with_ions(n=4, sigma=0.1e-10, seed=7, nions=3):
    """Periodic n^3 crystal plus `nions` Ga placed inside the lattice."""
    rng = np.random.default_rng(seed)
    X = configs.perturbed_crystal(n, sigma, seed)
    L = n * configs.A_SI
    ions = rng.random((nions, 3)) * L
    # keep ions away from lattice sites (> 1.2 A)
    for k in range(nions):
        for _ in range(1000):
            d = X - ions[k]
            d -= np.round(d / L) * L
            if np.sqrt((d * d).sum(1)).min() > 1.2e-10:
                break
            ions[k] = rng.random(3) * L
    Xa = np.vstack([X, ions])
    at = np.concatenate([np.ones(len(X), np.int32), np.full(nions, 3, np.int32)])
    return Xa, at, (L, L, L), len(X), nions


def with_ge(at, frac=0.3, seed=5):
    """A copy of `at` with a random fraction of the Si atoms made Ge (type 2)."""
    rng = np.random.default_rng(seed)
    at = np.array(at, np.int32)
    si = np.nonzero(at == 1)[0]
    at[rng.choice(si, int(frac * len(si)), replace=False)] = 2
    return at


def reference_slab(nc=(4, 4, 3), nions=2, seed=11):
    """The reference's geometry: pureepi lattice in a box 10x the slab in z,
    ions in storage at ionz (prms.f90:484-557, data.f90:223-260)."""
    a = configs.A_SI_REF
    X = configs.diamond(*nc, a=a)
    L = (a * nc[0], a * nc[1], a * nc[2] * 10)
    rng = np.random.default_rng(seed)
    ions = np.stack([rng.random(nions) * L[0], rng.random(nions) * L[1],
                     np.full(nions, 5.5 * L[2] / 10.0)], axis=1)
    Xa = np.vstack([X, ions])
    at = np.concatenate([np.ones(len(X), np.int32), np.full(nions, 3, np.int32)])
    return Xa, at, L, len(X), nions


def rel_err(Fg, Fo, floor_frac=1e-3):
    """max and RMS of |dF_i| / max(|F_i|, floor_frac * F_rms) and ||dF||/||F||."""
    d = np.sqrt(((Fg - Fo) ** 2).sum(1))
    m = np.sqrt((Fo ** 2).sum(1))
    frms = np.sqrt((m ** 2).mean())
    e = d / np.maximum(m, floor_frac * frms)
    return float(e.max()), float(np.sqrt((e ** 2).mean())), float(np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo))
