"""Error model of the mixed-precision SW forces (CPU, numpy float32 against
float64): why the mixed mode forms the pair bracket and the per-leg values
from FP64 geometry.  On the force microbench lattice (Si crystal, 0.1 A
noise, amorphous core) the forces are computed in float64, then with FP32
arithmetic in two layouts:

- all-FP32 geometry: the bond vector rounded to FP32, then r, 1/r, the pair
  bracket -4 B r^-5 - (B r^-4 - 1)/(r-a)^2 and the leg values in FP32;
- FP64 geometry: r, 1/r, 1/(r-a), the bracket and the leg values in FP64,
  each rounded once to FP32, the exponentials and the vertex terms in FP32;
- FP64 geometry with a compensated D (the mixed mode's layout): as above,
  and D = cos + 1/3 of the vertex terms formed from the FP32 unit vectors
  plus their low parts (10-bit fields in units of ulp/1024), with the
  products' exact errors and TwoSum errors carried, each rounded once.

Near the bond length the bracket's two parts cancel (dphi/dr = 0 at the
pair minimum, r = 2^(1/6) sigma = 2.35 A), so its FP32 rounding dominates
the per-atom error of the all-FP32 layout.  Near the tetrahedral angle D
cancels, which dominates once the geometry is FP64, most of all where the
forces are small (a lattice with 0.015 A noise has the forces of a 300 K
slab).  Per-atom metric: |dF| / max(|F|, 1e-3 F_rms).  Reduced units
(sigma = eps = 1), the SW parameters of prms.f90."""
import numpy as np
import pytest

from fib_md_amd import configs

scipy_spatial = pytest.importorskip("scipy.spatial")

A, B, ALPHA, LAM, GAM = 7.049556277, 0.6022245584, 1.8, 21.0, 1.2
SIG = 2.0951e-10
CUT = 0.999 * ALPHA
XLEG = 1.76  # mixed leg gate (E < e^-30 beyond it)
f4 = np.float32


def _pairs(n, noise=None):
    if noise is None:
        X, _, Lb = configs.force_microbench(n, seed=1002)
    else:
        L = n * configs.A_SI
        X, Lb = np.mod(configs.perturbed_crystal(n, noise, 1002), L), (L, L, L)
    Xs, Ls = X / SIG, Lb[0] / SIG
    tree = scipy_spatial.cKDTree(np.mod(Xs, Ls), boxsize=Ls)
    p = tree.query_pairs(CUT, output_type="ndarray")
    i = np.concatenate([p[:, 0], p[:, 1]])
    j = np.concatenate([p[:, 1], p[:, 0]])
    d = Xs[i] - Xs[j]
    d -= Ls * np.rint(d / Ls)
    return len(Xs), i, j, d


def _ulp(h):
    ex = np.abs(h).view(np.uint32) & np.uint32(0x7f800000)
    return np.where(ex > (23 << 23), (ex - np.uint32(23 << 23)).view(f4), f4(0)).astype(f4)


def _low_parts(x64, h):
    """x - h quantized to a 10-bit field in units of ulp(h)/1024, as FP32."""
    u = _ulp(h).astype(np.float64)
    q = np.where(u > 0, np.rint((x64 - h.astype(np.float64)) * 1024.0 / np.where(u > 0, u, 1.0)), 0)
    return (np.clip(q, -511, 511) * (u / 1024.0)).astype(f4)


def _two_sum_err(a, b, s):
    bb = (s - a).astype(f4)
    return ((a - (s - bb).astype(f4)).astype(f4) + (b - bb).astype(f4)).astype(f4)


def _cos_d(ea, eb, la, lb):
    """cos and cos + 1/3 from FP32 unit vectors and their low parts, FP32
    operations with the products' exact errors (what an FMA gives) and the
    sums' TwoSum errors carried."""
    p = (ea * eb).astype(f4)
    perr = (ea.astype(np.float64) * eb.astype(np.float64) - p).astype(f4)
    rp = ((perr[:, 0] + perr[:, 1]).astype(f4) + perr[:, 2]).astype(f4)
    cross = ((ea * lb).astype(f4) + (la * eb).astype(f4)).astype(f4)
    x = ((cross[:, 0] + cross[:, 1]).astype(f4) + cross[:, 2]).astype(f4)
    s1 = (p[:, 0] + p[:, 1]).astype(f4)
    s2 = (s1 + p[:, 2]).astype(f4)
    lo = ((_two_sum_err(p[:, 0], p[:, 1], s1) + _two_sum_err(s1, p[:, 2], s2)).astype(f4)
          + (rp + x).astype(f4)).astype(f4)
    th, tl = f4(1.0 / 3.0), f4(1.0 / 3.0 - float(f4(1.0 / 3.0)))
    s3 = (s2 + th).astype(f4)
    D = (s3 + (_two_sum_err(s2, np.full_like(s2, th), s3) + (lo + tl).astype(f4)).astype(f4)).astype(f4)
    return (s2 + lo).astype(f4), D


def _forces(N, i, j, d, mode):
    """mode: "fp64", "fp32" (all-FP32 geometry), "geo64" (FP64 geometry)
    or "geo64dc" (FP64 geometry, compensated D)."""
    F = np.zeros((N, 3))
    r64 = np.sqrt((d * d).sum(1))
    d32 = d.astype(f4)
    if mode == "fp32":
        rr = (d32 * d32).sum(1, dtype=f4)
        rinv = (f4(1) / np.sqrt(rr)).astype(f4)
        rs = (rr * rinv).astype(f4)
        ira = (f4(1) / (rs - f4(ALPHA))).astype(f4)
        ex = np.exp(ira).astype(f4)
        ir4 = (rinv * rinv) ** 2
        br = (f4(-4) * f4(B) * ir4 * rinv - (f4(B) * ir4 - f4(1)) * ira * ira).astype(f4)
        fs = (-(f4(A) * ex * br) * rinv).astype(f4)
        fp = (fs[:, None] * d32).astype(np.float64)
    else:
        ira = 1 / (r64 - ALPHA)
        ir4 = r64 ** -4
        ex = np.exp(ira) if mode == "fp64" else np.exp(ira.astype(f4)).astype(np.float64)
        br = -4 * B * ir4 / r64 - (B * ir4 - 1) * ira * ira
        fs = -(A * ex * br) / r64
        fp = fs[:, None] * d if mode == "fp64" else (fs.astype(f4)[:, None] * d32).astype(np.float64)
    np.add.at(F, i, fp)
    # legs of each centre (unit vector from the neighbour to the centre)
    leg = r64 < (CUT if mode == "fp64" else XLEG)
    li, lj, ld = i[leg], j[leg], d[leg]
    o = np.argsort(li, kind="stable")
    li, lj, ld = li[o], lj[o], ld[o]
    lr = np.sqrt((ld * ld).sum(1))
    if mode == "fp32":
        l32 = ld.astype(f4)
        rr = (l32 * l32).sum(1, dtype=f4)
        rinv = (f4(1) / np.sqrt(rr)).astype(f4)
        iraL = (f4(1) / ((rr * rinv).astype(f4) - f4(ALPHA))).astype(f4)
        e = (l32 * rinv[:, None]).astype(f4)
        E = np.exp(f4(GAM) * iraL).astype(f4)
        g = (-f4(GAM) * iraL * iraL).astype(f4)
        T = f4
    else:
        iraL = 1 / (lr - ALPHA)
        e, rinv = ld / lr[:, None], 1 / lr
        g = -GAM * iraL * iraL
        if mode == "fp64":
            E, T = np.exp(GAM * iraL), np.float64
        else:
            E, T = np.exp((GAM * iraL).astype(f4)), f4
            e64 = e
            e, rinv, g, E = e.astype(f4), rinv.astype(f4), g.astype(f4), E.astype(f4)
            elo = _low_parts(e64, e)
    starts = np.searchsorted(li, np.arange(N))
    cnt = np.searchsorted(li, np.arange(N), side="right") - starts
    ia, ib = [], []
    for k1 in range(cnt.max()):
        for k2 in range(cnt.max()):
            if k1 != k2:
                m = cnt > max(k1, k2)
                ia.append(starts[m] + k1)
                ib.append(starts[m] + k2)
    ia, ib = np.concatenate(ia), np.concatenate(ib)
    ea, eb = e[ia], e[ib]
    if mode == "geo64dc":
        cs, D = _cos_d(ea, eb, elo[ia], elo[ib])
    else:
        cs = (ea * eb).sum(1, dtype=T).astype(T)
        D = (cs + T(1.0 / 3.0)).astype(T)
    Eab = (T(LAM) * E[ia] * E[ib]).astype(T)
    c1 = (Eab * D * D * g[ia]).astype(T)
    c2 = (T(2) * Eab * D * rinv[ia]).astype(T)
    Fa = (c1[:, None] * ea + c2[:, None] * (eb - cs[:, None] * ea)).astype(np.float64)
    np.add.at(F, lj[ia], Fa)
    np.add.at(F, li[ia], -Fa)
    return F


def _err(Fm, Fd):
    dF = np.sqrt(((Fm - Fd) ** 2).sum(1))
    m = np.sqrt((Fd ** 2).sum(1))
    frms = np.sqrt((m ** 2).mean())
    return dF / np.maximum(m, 1e-3 * frms), dF / frms


def test_fp64_geometry_cuts_the_mixed_force_error():
    N, i, j, d = _pairs(12)
    ref = _forces(N, i, j, d, "fp64")
    # the float64 model is a force field: forces sum to zero
    assert np.abs(ref.sum(0)).max() < 1e-9 * np.abs(ref).sum()
    e32, a32 = _err(_forces(N, i, j, d, "fp32"), ref)
    e64, a64 = _err(_forces(N, i, j, d, "geo64"), ref)
    print(f"all-FP32 geometry: median |dF|/F_rms {np.median(a32):.2e}, max e {e32.max():.2e}; "
          f"FP64 geometry: {np.median(a64):.2e}, max e {e64.max():.2e}")
    # FP32 rounding of ~16 terms of a few eps/sigma each: a few 1e-7 of F_rms
    assert 2e-7 < np.median(a32) < 8e-7
    # FP64 geometry takes the typical error down at least fourfold and keeps
    # every atom within 1e-5
    assert np.median(a64) < np.median(a32) / 4
    assert e64.max() <= 1e-5


def test_compensated_d_on_small_forces():
    # a lattice with the forces of a 300 K slab, where D = cos + 1/3 sits
    # near zero for most leg pairs
    N, i, j, d = _pairs(16, noise=0.015e-10)
    ref = _forces(N, i, j, d, "fp64")
    e64, a64 = _err(_forces(N, i, j, d, "geo64"), ref)
    edc, adc = _err(_forces(N, i, j, d, "geo64dc"), ref)
    print(f"FP64 geometry: median |dF|/F_rms {np.median(a64):.2e}, max e {e64.max():.2e}; "
          f"compensated D: {np.median(adc):.2e}, max e {edc.max():.2e}")
    # with plain FP32 D the worst atoms exceed 1e-5; the compensated D takes
    # the typical error down at least fourfold and every atom within 1e-5
    assert e64.max() > 1e-5
    assert np.median(adc) < np.median(a64) / 4
    assert edc.max() <= 1e-5
