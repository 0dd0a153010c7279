"""Properties at the benchmark configurations' full sizes (the oracle only
finishes small systems in seconds; these checks need no oracle run).

Force microbench (2,097,107 atoms: 64^3-cell crystal, 0.1 A noise, 20 A
amorphous sphere):
- cell ids of every atom equal the reference's formula (nlist.f90:29,
  INT((X - FLOOR(X iLb) Lb) / (Lb/N)) + 1, evaluated here in numpy's FP64);
- the full neighbour sets of 256 sampled atoms equal a brute-force scan of
  all atoms with the reference predicate (si.f90:156-171: NINT minimum image,
  r^2 < rc^2, no contraction);
- Newton's third law (sum of forces ~ 0 against the sum of |F|), a
  symmetric virial, and the energy per atom between the perfect crystal's
  -2 eps and -4.0 eV;
- mixed against FP64: the stated bound 1e-5 on the RMS and norm relative
  force error, fewer than 0.01 % of the atoms above it, energy within 1e-6.

Batched ensemble (64 slabs, 5.6 M atoms, 1-10 keV ions): the energy
ledger of every instance stays within 1e-3 of its ion's energy over 200
steps, in both precisions.
"""
import numpy as np
import pytest

from fib_md_amd import configs

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19


@pytest.fixture(scope="module")
def microbench():
    X, at, L = configs.force_microbench(64)
    print(f"\nmicrobench: {len(X)} atoms, box {L[0] * 1e10:.2f} A", flush=True)
    return X, at, L


def _nint(t):
    """Fortran NINT: nearest integer, halves away from zero."""
    return np.where(t >= 0, np.floor(t + 0.5), np.ceil(t - 0.5))


def test_microbench_cells_lists_forces(microbench):
    from fib_md_amd import FibGPU
    X, at, L = microbench
    n = len(X)
    res = {}
    for prec in ("fp64", "mixed"):
        g = FibGPU([(n, n, 0, L)], precision=prec)
        g.set_state(X, atype=at)
        g.build_nlist()
        r = g.force(impact=0, energy=True, virial=True)
        if prec == "fp64":
            ic = g.bin_cells()
            lists = g.nlist()
            ncell = g.cells()
        g.close()
        res[prec] = r
        print(f"{prec}: utot {r['utot']:.9e} J", flush=True)

    # cell ids, every atom
    Lv = np.asarray(L, dtype=np.float64)
    iL = 1.0 / Lv
    Nc = np.asarray(ncell)
    w = X - np.floor(X * iL) * Lv
    ref = (w / (Lv / Nc)).astype(np.int64) + 1
    np.testing.assert_array_equal(ic, ref)

    # neighbour sets of sampled atoms against brute force
    nss, nl2 = lists["nss"], lists["nlist2"]
    # owner atom of every nlist2 entry (atom i holds entries nss[i,0]..nss[i,1], 1-based)
    cnt = np.maximum(nss[:, 1] - nss[:, 0] + 1, 0)
    excl = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    pos = np.repeat(nss[:, 0] - 1, cnt) + (np.arange(cnt.sum()) - np.repeat(excl, cnt))
    owner = np.full(len(nl2), -1, np.int64)
    owner[pos] = np.repeat(np.arange(n), cnt)
    assert (owner >= 0).all()
    rc = 1.1 * 1.8 * 2.0951e-10
    rc2 = rc * rc
    rng = np.random.default_rng(17)
    sample = rng.choice(n, 256, replace=False)
    for i in sample:
        d = X[i] - X
        d = d - _nint(d * iL) * Lv
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        bf = set(np.nonzero(r2 < rc2)[0].tolist()) - {int(i)}
        a, b = nss[i]
        mine = set((nl2[a - 1:b] - 1).tolist()) if b >= a else set()
        back = set(owner[np.nonzero(nl2 == i + 1)[0]].tolist())
        got = mine | back
        assert got == bf, (int(i), sorted(got ^ bf)[:8])
        assert all(j > i for j in mine)

    F64, Fm = res["fp64"]["F"], res["mixed"]["F"]
    mag = np.sqrt((F64 ** 2).sum(1))
    assert np.abs(F64.sum(0)).max() <= 1e-9 * mag.sum()
    assert np.abs(Fm.sum(0)).max() <= 1e-7 * mag.sum()
    W = res["fp64"]["virial"]
    assert np.abs(W - W.T).max() <= 1e-8 * np.abs(W).max()
    e_atom = res["fp64"]["utot"] / n / EV
    print(f"energy per atom {e_atom:.5f} eV", flush=True)
    assert -2 * 2.1683 - 1e-3 < e_atom < -4.0
    d = np.sqrt(((Fm - F64) ** 2).sum(1))
    frms = np.sqrt((mag ** 2).mean())
    e = d / np.maximum(mag, 1e-3 * frms)
    rms, norm = float(np.sqrt((e ** 2).mean())), float(np.linalg.norm(Fm - F64) / np.linalg.norm(F64))
    frac = float((e > 1e-5).mean())
    print(f"mixed vs fp64: rms {rms:.2e}, norm {norm:.2e}, above 1e-5: {frac:.2e} of atoms, "
          f"max {e.max():.2e}", flush=True)
    assert rms < 1e-5 and norm < 1e-5 and frac < 1e-4
    # every atom within 1e-5 of the larger of its force and 1e-3 F_rms
    assert e.max() <= 1e-5
    assert abs(res["mixed"]["utot"] - res["fp64"]["utot"]) <= 1e-6 * abs(res["fp64"]["utot"])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_ensemble_full_size_ledger(precision):
    from fib_md_amd import autotune
    ens = configs.ensemble(64)
    print(f"\nensemble: {sum(len(s['X']) for s in ens)} atoms, {precision}", flush=True)
    g = autotune.md_handle(ens, precision)
    L0 = g.md_ledger()
    g.md_run(200)
    L1 = g.md_ledger()
    fired = np.array([g.md_report(k)["e_fired"] for k in range(len(ens))])
    g.close()
    drift = np.abs(L1 - L0)
    # the ions were fired inside md_handle, after its ledger start; L0 here
    # already holds them, so the drift is the integration error alone
    print(f"ledger drift over 200 steps: max {drift.max() / EV:.3f} eV "
          f"({(drift / fired).max():.2e} of the ion energy)", flush=True)
    assert np.all(drift <= 1e-3 * fired)
