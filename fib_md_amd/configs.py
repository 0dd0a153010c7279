"""Synthetic configurations for tests and benchmarks (host-side numpy).

Diamond-cubic Si lattices, seeded perturbations, an amorphous (randomised)
sphere, Maxwell-Boltzmann velocities and the reference's own initial state
(pureepi, prms.f90:484-557; randvel, data.f90:169-204) are generated here.
All lengths in m, velocities in m/s.
"""
from __future__ import annotations

import math

import numpy as np

A_SI_REF = 5.431073e-10   # reference lattice constant (prms.f90:494)
A_SI = 5.431e-10          # lattice constant of the benchmark configurations
MASS_SI = 46.637063e-27
MASS_GA = 1.157777417e-25
KB = 1.381e-23
EV = 1.60217646e-19

# 8-site basis of the conventional cell in the reference's order
# (prms.f90:498-505), fractions of the lattice constant.
BASIS = np.array([
    [0.00, 0.00, 0.00], [0.00, 0.50, 0.50], [0.25, 0.25, 0.75], [0.25, 0.75, 0.25],
    [0.50, 0.00, 0.50], [0.50, 0.50, 0.00], [0.75, 0.25, 0.25], [0.75, 0.75, 0.75],
])


def diamond(n1: int, n2: int, n3: int, a: float = A_SI) -> np.ndarray:
    """Lattice sites in the reference's loop order x, y, z, basis -> (N,3)."""
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    pts = (cells[:, None, :] + BASIS[None, :, :]).reshape(-1, 3) * a
    return pts


def perturbed_crystal(n: int, sigma: float, seed: int, a: float = A_SI) -> np.ndarray:
    """n^3-cell periodic crystal with N(0, sigma) displacement per component."""
    rng = np.random.default_rng(seed)
    X = diamond(n, n, n, a)
    X = X + rng.normal(0.0, sigma, X.shape)
    L = n * a
    return np.mod(X, L)


def amorphous_sphere(X: np.ndarray, L: float, radius: float, density: float, rmin: float,
                     seed: int) -> np.ndarray:
    """Replace the atoms inside a central sphere by uniform random points at
    `density` (atoms/m^3) with rejection minimum distance `rmin`."""
    rng = np.random.default_rng(seed)
    c = np.array([L / 2, L / 2, L / 2])
    d = X - c
    inside = (d * d).sum(1) < radius * radius
    outside = X[~inside]
    n_new = int(round(density * 4.0 / 3.0 * math.pi * radius ** 3))
    # neighbours of the sphere shell constrain the rejection test too
    shell = outside[((outside - c) ** 2).sum(1) < (radius + rmin) ** 2]
    pts = []
    tries = 0
    r2min = rmin * rmin
    while len(pts) < n_new and tries < 200 * n_new:
        tries += 1
        p = c + radius * (2 * rng.random(3) - 1)
        if ((p - c) ** 2).sum() >= radius * radius:
            continue
        if len(pts) and (((np.array(pts) - p) ** 2).sum(1) < r2min).any():
            continue
        if len(shell) and (((shell - p) ** 2).sum(1) < r2min).any():
            continue
        pts.append(p)
    return np.vstack([outside, np.array(pts)]) if pts else outside


def maxwell_boltzmann(masses: np.ndarray, T: float, seed: int) -> np.ndarray:
    """Velocities at temperature T with zero net momentum."""
    rng = np.random.default_rng(seed)
    s = np.sqrt(KB * T / masses)[:, None]
    V = rng.normal(0.0, 1.0, (len(masses), 3)) * s
    p = (masses[:, None] * V).sum(0) / masses.sum()
    return V - p[None, :]


def force_microbench(n: int = 64, seed: int = 1002):
    """The force microbenchmark: n^3-cell periodic Si crystal, N(0, 0.1 A)
    noise, a central R = 20 A sphere randomised at 2.28 g/cc (rmin 2.0 A).
    Returns (X, atype, Lb)."""
    L = n * A_SI
    X = perturbed_crystal(n, 0.1e-10, seed)
    rho = 2.28e3 / (28.0855 * 1.66053907e-27)  # atoms/m^3 at 2.28 g/cc
    X = amorphous_sphere(X, L, 20e-10, rho, 2.0e-10, seed + 1)
    X = np.mod(X, L)
    return X, np.ones(len(X), np.int32), (L, L, L)


# Ga launch speed per sqrt(eV) (knock.f90:29): v = 1663.6334 sqrt(E/eV) m/s
GA_SPEED = 1663.6334
# height of the unfired ion above the surface, inside the removal height so it
# stays in the simulation (unfired ions carry no forces)
ION_PARK = 8.0e-10


def si_slab(n: int, seed: int, nz: int | None = None, vacuum: float = 40e-10, T: float = 300.0,
            a: float = A_SI, nions: int = 1):
    """Si(100) slab of n x n x nz cells for the benchmark configurations:
    periodic in x/y, free top (box z = slab + vacuum), the bottom two atomic
    layers (z = 0 and a/4) pinned, Maxwell-Boltzmann velocities at T (pinned
    atoms at rest), `nions` Ga ions parked above the surface.
    Returns dict(X, V, atype, xsides, L, nsg, nlj, surface_z, bottom_z)."""
    nz = n if nz is None else nz
    S = diamond(n, n, nz, a)
    L = (n * a, n * a, nz * a + vacuum)
    pinned = S[:, 2] < 0.3 * a
    V = maxwell_boltzmann(np.full(len(S), MASS_SI), T, seed)
    V[pinned] = 0.0
    surface = float(S[:, 2].max())
    rng = np.random.default_rng(seed + 7919)
    ions = np.column_stack([rng.random(nions) * L[0], rng.random(nions) * L[1],
                            np.full(nions, surface + ION_PARK)])
    X = np.vstack([S, ions])
    V = np.vstack([V, np.zeros((nions, 3))])
    atype = np.concatenate([np.ones(len(S), np.int32), np.full(nions, 3, np.int32)])
    xsides = np.zeros(len(X), np.int32)
    xsides[np.nonzero(pinned)[0]] = np.nonzero(pinned)[0] + 1
    return dict(X=X, V=V, atype=atype, xsides=xsides, L=L, nsg=len(S), nlj=nions,
                surface_z=surface, bottom_z=0.0)


def ion_launch(L, surface_z: float, energy_ev: float, polar_deg: float, azimuth_deg: float,
               rng, height: float = 5e-10):
    """Ion start 5 A above the surface at a uniform point of the surface cell,
    speed for `energy_ev`, direction (polar from -z, azimuth from +x)."""
    x = np.array([rng.random() * L[0], rng.random() * L[1], surface_z + height])
    th, ph = math.radians(polar_deg), math.radians(azimuth_deg)
    s = GA_SPEED * math.sqrt(energy_ev)
    v = np.array([s * math.sin(th) * math.cos(ph), s * math.sin(th) * math.sin(ph),
                  -s * math.cos(th)])
    return x, v


def ensemble(n_inst: int = 64, seed: int = 2000, nmin: float = 1e4, nmax: float = 2.5e5,
             emin: float = 1e3, emax: float = 1e4, polar_max: float = 60.0):
    """The batched-ensemble configuration: instance k (seed + k) is a cubic
    Si(100) slab of n = round((N/8)^(1/3)) cells with N log-uniform in
    [nmin, nmax], hit by one Ga ion of energy log-uniform in [emin, emax] eV,
    polar angle uniform in [0, polar_max] deg, azimuth uniform.
    Returns a list of dicts (si_slab fields plus ion x, v, energy, angles)."""
    out = []
    for k in range(n_inst):
        rng = np.random.default_rng(seed + k)
        N = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        n = max(3, int(round((N / 8.0) ** (1.0 / 3.0))))
        E = math.exp(rng.uniform(math.log(emin), math.log(emax)))
        th = rng.uniform(0.0, polar_max)
        ph = rng.uniform(0.0, 360.0)
        slab = si_slab(n, seed + k)
        x, v = ion_launch(slab["L"], slab["surface_z"], E, th, ph, rng)
        slab.update(ion_x=x, ion_v=v, energy_ev=E, polar=th, azimuth=ph, ncell=n)
        out.append(slab)
    return out


def big_impact(seed: int = 1003, n: int = 60, nz: int = 80, energy_ev: float = 30e3,
               polar_deg: float = 7.0):
    """The single large impact: an n x n x nz-cell Si(100) slab at 300 K
    (2,304,000 atoms for 60 x 60 x 80) hit by one Ga ion of `energy_ev` at
    `polar_deg` off normal, azimuth drawn from the seed."""
    slab = si_slab(n, seed, nz=nz)
    rng = np.random.default_rng(seed + 1)
    az = rng.uniform(0.0, 360.0)
    x, v = ion_launch(slab["L"], slab["surface_z"], energy_ev, polar_deg, az, rng)
    slab.update(ion_x=x, ion_v=v, energy_ev=energy_ev, polar=polar_deg, azimuth=az)
    return slab


def dose_replicas(n_rep: int = 16, n: int = 40, impacts: int = 200, seed: int = 3000,
                  energy_ev: float = 2e3):
    """The cumulative-dose configuration: `n_rep` Si(100) slabs of n^3 cells
    (512,000 atoms for n = 40), each with `impacts` Ga ions parked above the
    surface, fired one after another at normal incidence from seeded random
    (x, y) points 5 A above the surface.  Returns the slabs and, per replica,
    the (x, v) of every impact in firing order."""
    reps, shots = [], []
    for r in range(n_rep):
        slab = si_slab(n, seed + r, nions=impacts)
        rng = np.random.default_rng(seed + r + 104729)
        shots.append([ion_launch(slab["L"], slab["surface_z"], energy_ev, 0.0, 0.0, rng)
                      for _ in range(impacts)])
        reps.append(slab)
    return reps, shots
