"""End-to-end run of the Fortran host (fortran/mdrun_gpu) on the GPU against
the oracle: tests/golden/siga_small.in (3x3x2 cells, 2 Ga ions at 500 eV,
7 deg / 30 deg incidence), fixed seed.  The oracle replays main/tint
(main.f90:31-74, timeint.f90:36-120) with the same generators; the
data/mdrun2.xyz frames must agree: same frame count, lt cadence, headers,
labels and rank column, positions within 2e-4 A and KE within one unit of
the printed last digit (the two FP64 force sums add in different orders)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle, siga_from_values
from tests.test_fortran_host import EXPECTED, SIGA

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortran", "mdrun_gpu")
SEED = -24680
MASS = {1: 46.637063e-27, 3: 1.157777417e-25}


def oracle_frames():
    o = Oracle()
    o.readprms(siga_from_values(EXPECTED), SEED)
    o.initranions()
    o.initdata()
    o.init_nlist(); o.mktables()
    X, V, _ = o.state()
    o.si_nlist(X)
    at, _ = o.types()
    frames = []

    def frame():
        X, V, _ = o.state()
        frames.append((o.getd("time"), X.copy(), V.copy()))

    frames.append((0.0, X.copy(), V.copy()))
    o.tint_init()
    o.setd("Tau", 1e-14)
    o.ti_begin(0)
    while not o.ti_iter(0, EXPECTED["ttar1"]):
        pass
    for impact in range(1, EXPECTED["nlj"] + 1):
        o.seti("impact", impact); o.seti("atom_is_fast", 1); o.seti("atom_is_slow", 0)
        o.ionrun(1)
        o.knock(impact)
        while not o.ti_iter(1, EXPECTED["ttar1"]):
            lt, ao = o.geti("lt"), o.geti("atm_out")
            if ao > 0 and lt % ao == 0:
                frame()
        frame()
    return frames, at


def parse(path):
    lines = open(path).read().split("\n")[:-1]
    n = int(lines[0])
    out = []
    for f in range(len(lines) // (n + 2)):
        b = f * (n + 2)
        hdr = lines[b + 1]
        rows = lines[b + 2:b + 2 + n]
        out.append((hdr, rows))
    return n, out


def test_fortran_driver_matches_oracle(tmp_path, oracle_lib):
    assert os.path.exists(EXE), "build fortran/mdrun_gpu first"
    shutil.copy(SIGA, tmp_path / "siga.in")
    os.makedirs(tmp_path / "data")
    print(f"mdrun_gpu --seed {SEED} in {tmp_path}", flush=True)
    r = subprocess.run([EXE, "--seed", str(SEED)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    frames, at = oracle_frames()
    n, got = parse(tmp_path / "data" / "mdrun2.xyz")
    assert n == len(at)
    assert len(got) == len(frames), (len(got), len(frames))
    for (hdr, rows), (time, X, V) in zip(got, frames):
        assert hdr.startswith("Time (ps) = ") and hdr[12:24] == f"{time * 1e12:12.5f}"
        for i in range(n):
            row = rows[i]
            x = [float(row[12 * c:12 * c + 12]) for c in range(3)]
            assert np.allclose(x, X[i] * 1e10, rtol=0, atol=2e-4), (i, row, X[i] * 1e10)
            assert row[36:44] == f"{'Si' if at[i] == 1 else 'Ga':>8s}"
            assert row[44:52] == f"{0:8d}"
            ke = 0.5 * MASS[int(at[i])] * (V[i] ** 2).sum() * 6.252e18
            got_ke = float(row[52:66])
            assert abs(got_ke - ke) <= 1e-3 * abs(ke) + 1e-30, (i, row, ke)
    snap = open(tmp_path / "data" / "3dmol2.xyz").read().split("\n")
    assert int(snap[0]) == n


def _restart_header(path):
    """First record of a restart file (out.f90:46): int32 lt, real(8) clock,
    between 4-byte sequential-record markers."""
    raw = open(path, "rb").read(24)
    n = int(np.frombuffer(raw[0:4], np.int32)[0])
    assert n == 12, n
    lt = int(np.frombuffer(raw[4:8], np.int32)[0])
    t = float(np.frombuffer(raw[8:16], np.float64)[0])
    return lt, t


def _set(text, key, value):
    lines = text.split("\n")
    hit = [k for k, ln in enumerate(lines) if ln.split("\\")[-1].strip().startswith(key)]
    assert len(hit) == 1, (key, hit)
    lines[hit[0]] = f"{value}    \\ {key}"
    return "\n".join(lines)


def test_fortran_driver_restart(tmp_path):
    """Restart records on the res_out cadence (out.f90:34-59, 129-150) and a
    whichic-0 start from one (data.f90:71-93, timeint.f90:50-111).  The run
    writes a record every step; the last record of impact 1 is the state at
    the last step of that impact.  A run restarted from it (the record plus
    its ran1 companion file) skips equilibration, keeps the step counter and
    the simulated clock of the record, fires only ion 2, from the same spot,
    and ends impact 2 with the uninterrupted run's final frame: same clock,
    positions within 2e-4 A and KE within 1e-3 (the two runs rebuild their
    lists from different atom orders, so the FP64 force sums round
    differently).  Cadences of 1 step make both runs rebuild, freeze
    sputtered atoms and test the regime on every step, although the
    restarted run's step counter is one behind (the reference's TI loop
    counts one more step when an impact ends, timeint.f90:140-145)."""
    assert os.path.exists(EXE), "build fortran/mdrun_gpu first"
    text = open(SIGA).read()
    text = _set(text, "res_out", 1)
    for k in ("ntlist", "atlist", "dslist"):
        for s in ("_i", "_r"):
            text = _set(text, k + s, 1)
    a = tmp_path / "a"
    b = tmp_path / "b"
    for top in (a, b):
        for d in ("data", "D", "F"):
            os.makedirs(top / d)
    (a / "siga.in").write_text(text)
    print(f"mdrun_gpu --seed {SEED} (res_out 1) in {a}", flush=True)
    r = subprocess.run([EXE, "--seed", str(SEED)], cwd=a, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    names = sorted(os.listdir(a / "D"))
    assert "restart.in_old" in names and "restart.rng_old" in names
    rec1 = sorted((f for f in names if f.startswith("restart.in000001_")),
                  key=lambda f: int(f.split("_")[1]))
    assert len(rec1) > 10, names[:5]
    last = rec1[-1]
    lt_rec = int(last.split("_")[1])
    lt_h, t_rec = _restart_header(a / "D" / last)
    assert lt_h == lt_rec
    # the record's clock is the clock of its step: impact 1's final frame
    # (written after the TI loop exits, no step in between) shows it
    n, fa = parse(a / "data" / "mdrun2.xyz")
    times_a = [float(h[12:24]) for h, _ in fa]
    assert any(abs(t_rec * 1e12 - t) < 1e-5 for t in times_a), (t_rec, times_a)
    assert t_rec > EXPECTED["dti"] and t_rec <= EXPECTED["dti"] + 1.01 * EXPECTED["ts_r"]
    # earlier records carry earlier clocks
    lt_e, t_e = _restart_header(a / "D" / rec1[-11])
    assert lt_e == lt_rec - 10 and t_e < t_rec
    for stem in ("restart.in", "restart.rng"):
        shutil.copy(a / "D" / last.replace("restart.in", stem), b / "D")
    rtext = _set(_set(text, "whichic", 0), "startimpact", 1)
    (b / "siga.in").write_text(rtext)
    print(f"mdrun_gpu restarted from {last}", flush=True)
    r2 = subprocess.run([EXE, "--seed", str(SEED)], cwd=b, capture_output=True, text=True,
                        timeout=300)
    assert r2.returncode == 0, r2.stdout[-2000:] + r2.stderr[-2000:]
    out = r2.stdout
    nsg = 8 * 3 * 3 * 2
    assert "Initializing target temperature" not in out
    knocks = [ln for ln in out.splitlines() if "KNOCK" in ln]
    assert len(knocks) == 1 and knocks[0].split()[-1] == str(nsg + 2), knocks
    # progress lines of the restarted run continue the step counter and clock
    prog = [ln.split() for ln in out.splitlines()
            if len(ln.split()) == 3 and ln.split()[0].isdigit()]
    assert prog and int(prog[0][0]) > lt_rec, (prog[:3], lt_rec)
    assert float(prog[0][1]) > t_rec * 1e12
    n2, fb = parse(b / "data" / "mdrun2.xyz")
    assert n2 == n
    (ha, ra), (hb, rb) = fa[-1], fb[-1]
    assert ha[:24] == hb[:24], (ha[:24], hb[:24])
    for i in range(n):
        xa = [float(ra[i][12 * c:12 * c + 12]) for c in range(3)]
        xb = [float(rb[i][12 * c:12 * c + 12]) for c in range(3)]
        assert np.allclose(xa, xb, rtol=0, atol=2e-4), (i, ra[i], rb[i])
        assert ra[i][36:52] == rb[i][36:52]
        ka, kb = float(ra[i][52:66]), float(rb[i][52:66])
        assert abs(ka - kb) <= 1e-3 * abs(ka) + 1e-30, (i, ra[i], rb[i])


CFG = os.path.join(ROOT, "fortran", "mdcfg_gpu")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fortran_config_driver(tmp_path, prec):
    """The Fortran host's driver of the benchmark configurations' impacts
    (fortran/mdcfg_gpu, MD-step mode): two replicas of a 4x4x4-cell slab,
    two sequential 1 keV impacts each at 7 deg, 3 A Berendsen shells.  Pins:
    the energy each impact brings in (1663.6334 sqrt(eV) m/s for Ga), the
    energy ledger (constant to 1e-3 of the energy fired while the shells
    work), dt within the configs' cap, and replica 1's frames in the
    reference's data/ format (initial state and one per impact)."""
    assert os.path.exists(CFG), "build fortran/mdcfg_gpu first"
    os.makedirs(tmp_path / "data")
    args = ["--cells", "4", "--nz", "4", "--replicas", "2", "--impacts", "2", "--steps", "400",
            "--energy", "1000", "--polar", "7", "--azimuth", "30", "--shell", "3", "--seed", "11",
            "--frames"] + (["--mixed"] if prec == "mixed" else [])
    print("mdcfg_gpu " + " ".join(args), flush=True)
    r = subprocess.run([CFG, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = []
    for ln in r.stdout.splitlines():
        t = ln.split()
        if t and t[0] == "impact":
            rows.append({t[k]: float(t[k + 1]) for k in range(0, len(t) - 1, 2)})
    assert [int(x["impact"]) for x in rows] == [1, 2], r.stdout
    print(r.stdout, flush=True)
    for m, x in enumerate(rows, start=1):
        # the parameter block's Ga mass makes 1663.6334^2 eV/(2 e) 1 eV to 1e-7
        assert abs(x["fired_eV"] - 1000.0 * m) < 1e-3 * m
        assert abs(x["drift_max_eV"]) < 1e-3 * x["fired_eV"], x
        assert 0 < x["dt_s"] <= 1e-15 and x["time_s"] > 0
        assert x["yield"] >= 0
    natm = 8 * 4 * 4 * 4 + 2
    n, frames = parse(tmp_path / "data" / "mdrun2.xyz")
    assert n == natm and len(frames) == 3
    for hdr, fr in frames:
        assert hdr.startswith("Time (ps) = ")
        labels = [row[36:44].strip() for row in fr]
        assert labels.count("Si") == natm - 2 and labels.count("Ga") == 2
    assert int(open(tmp_path / "data" / "3dmol2.xyz").read().split("\n")[0]) == natm


def test_fortran_config_driver_groups_and_async_frames(tmp_path):
    """mdcfg_gpu over three replicas: the frames written from asynchronous
    snapshots (each written while the next impact runs) are byte for byte
    those of synchronous fetches (--sync-frames), and a run over replicas
    2..3 only (--first 2 --count 2) gives those replicas' yields and ledger
    drifts exactly as the run over all three."""
    assert os.path.exists(CFG), "build fortran/mdcfg_gpu first"
    base = ["--cells", "4", "--nz", "4", "--replicas", "3", "--impacts", "2", "--steps", "300",
            "--energy", "1000", "--polar", "7", "--shell", "3", "--seed", "12"]
    out = {}
    for name, extra in (("async", ["--frames"]), ("sync", ["--sync-frames"]),
                        ("group", ["--first", "2", "--count", "2"])):
        d = tmp_path / name
        os.makedirs(d / "data")
        print("mdcfg_gpu " + " ".join(base + extra), flush=True)
        r = subprocess.run([CFG, *base, *extra], cwd=d, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        per = [ln.split()[2:] for ln in r.stdout.splitlines()
               if ln.split()[:1] == ["replicas"] and ln.split()[2] != "simulated"]
        out[name] = (d, per)
    for f in ("mdrun2.xyz", "3dmol2.xyz"):
        a = open(out["async"][0] / "data" / f, "rb").read()
        b = open(out["sync"][0] / "data" / f, "rb").read()
        assert len(a) > 0 and a == b, f
    full, grp = out["async"][1], out["group"][1]
    assert len(full) == len(grp) == 2
    for a, b in zip(full, grp):
        assert len(a) == 6 and len(b) == 4
        assert a[2:] == b, (a, b)
