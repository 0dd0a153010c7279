"""The Fortran host routines (fortran/fibhost.f90) on the CPU, through their C
entry points in libfibhost.so:

- siga.in reader: list-directed reads, one record per value, comment tails
  and blank records skipped (prms.f90:221-277), on tests/golden/siga_small.in;
- ran1, pureepi + si.dat round trip, initranions + fwhm.dat round trip, initial
  ions and velocities, knock: bit-identical with the oracle (whose generators
  are pinned in test_oracle_init_integrate.py);
- data/mdrun2.xyz and data/3dmol2.xyz: the reference's format strings
  (out.f90:74-96), checked field by field against independently formatted
  expectations;
- restart records (out.f90:34-59): the unformatted sequential layout of the
  reference's three records, parsed byte by byte, read back as a whichic-0
  start (data.f90:71-93), and the ran1 sequence carried on across it.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle, siga_from_values

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "fortran")
SIGA = os.path.join(ROOT, "tests", "golden", "siga_small.in")

EXPECTED = dict(
    amorcrys=0, whichic=4, startimpact=0, eV=500.0, nc=(3, 3, 2), npc=(1, 1, 1), nlj=2, nt=100,
    ts_i=2.0e-16, ts_r=1.0e-15, ttar1=300.0, ke_limit_1=600.0, ke_limit_2=10.0, tau_i=1.0e-13,
    tau_r=1.0e-13, teps=30.0, uo=1500.0, ntlist_i=10, ntlist_r=10, atlist_i=10, atlist_r=10,
    dslist_i=40, dslist_r=40, atm_out_i=25, atm_out_r=25, res_out=-5000, eng_out=-10,
    tmp_out=-100, nrand=200, outz=(1.3, 3.5, 7.5, 9.8), atomz=(4.5, 6.5), ionz=5.5,
    knockz=1.05, sidewidth=1.5, dti=6.2e-14, fwhm=2.0, phiz=7.0, phixy=30.0)
ORDER = ["amorcrys", "whichic", "startimpact", "eV", "nc", "npc", "nlj", "nt", "ts_i", "ts_r",
         "ttar1", "ke_limit_1", "ke_limit_2", "tau_i", "tau_r", "teps", "uo", "ntlist_i",
         "ntlist_r", "atlist_i", "atlist_r", "dslist_i", "dslist_r", "atm_out_i", "atm_out_r",
         "res_out", "eng_out", "tmp_out", "nrand", "outz", "atomz", "ionz", "knockz",
         "sidewidth", "dti", "fwhm", "phiz", "phixy"]


@pytest.fixture(scope="module")
def fh():
    subprocess.run(["make", "-s", "-C", FDIR, "libfibhost.so"], check=True)
    L = C.CDLL(os.path.join(FDIR, "libfibhost.so"))
    L.fh_ran1.restype = C.c_double
    L.fh_temperature.restype = C.c_double
    L.fh_set_seed.argtypes = [C.c_int]
    L.fh_setup.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                           C.c_void_p]
    L.fh_writeatoms.argtypes = [C.c_int64, C.c_double, C.c_double]
    L.fh_open_outputs.argtypes = [C.c_int]
    L.fh_close_outputs.argtypes = [C.c_int]
    L.fh_write_restart.argtypes = [C.c_int, C.c_double, C.c_int]
    L.fh_read_restart.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int)]
    return L


def flat_expected():
    out = []
    for k in ORDER:
        v = EXPECTED[k]
        out.extend(v if isinstance(v, tuple) else [v])
    return out


def test_siga_reader(fh):
    v = np.zeros(46)
    assert fh.fh_read_siga(SIGA.encode(), v.ctypes.data_as(C.c_void_p)) == 0
    np.testing.assert_array_equal(v, np.array(flat_expected(), dtype=float))


def test_ran1_matches_oracle(fh, oracle_lib):
    o = Oracle()
    o.seti("ranseed", -4321)
    fh.fh_set_seed(-4321)
    a = [fh.fh_ran1() for _ in range(5000)]
    b = [o.ran1() for _ in range(5000)]
    assert a == b
    assert fh.fh_get_seed() == o.geti("ranseed")


def _setup(fh, tmp_path, seed=-9876):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        natm, nsg = C.c_int(), C.c_int()
        lb = np.zeros(3)
        assert fh.fh_setup(SIGA.encode(), seed, C.byref(natm), C.byref(nsg),
                           lb.ctypes.data_as(C.c_void_p)) == 0
    finally:
        os.chdir(cwd)
    n = natm.value
    X = np.zeros((n, 3), order="F"); V = np.zeros((n, 3), order="F")
    Xi = np.zeros((n, 3), order="F")
    at = np.zeros(n, np.int32); xs = np.zeros(n, np.int32)
    xr = np.zeros((EXPECTED["nrand"], 2), order="F")
    fh.fh_get_state(*(a.ctypes.data_as(C.c_void_p) for a in (X, V, at, xs, Xi, xr)))
    return n, nsg.value, lb, X, V, at, xs, Xi, xr


def test_initial_state_matches_oracle(fh, tmp_path, oracle_lib):
    n, nsg, lb, X, V, at, xs, Xi, xr = _setup(fh, tmp_path)
    o = Oracle()
    o.readprms(siga_from_values(EXPECTED), -9876)
    o.initranions()
    o.initdata()
    Xo, Vo, _ = o.state()
    ato, _ = o.types()
    assert (n, nsg) == (o.natm, o.geti("nsg")) == (8 * 18 + 2, 8 * 18)
    assert tuple(lb) == (o.getd("lb1"), o.getd("lb2"), o.getd("lb3"))
    np.testing.assert_array_equal(X, Xo)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(Xi, o.xi())
    np.testing.assert_array_equal(at, ato)
    np.testing.assert_array_equal(xs, o.xsides())
    np.testing.assert_array_equal(xr, o.xrand())
    assert os.path.exists(tmp_path / "si.dat") and os.path.exists(tmp_path / "fwhm.dat")
    # si.dat is E20.10 text: 3 x 20 columns + I10
    first = open(tmp_path / "si.dat").readline().rstrip("\n")
    assert len(first) == 70 and first.endswith("         1")


def test_knock_matches_oracle(fh, tmp_path, oracle_lib):
    _setup(fh, tmp_path, seed=-555)
    o = Oracle()
    o.readprms(siga_from_values(EXPECTED), -555)
    o.initranions()
    o.initdata()
    xion, vion = np.zeros(3), np.zeros(3)
    fh.fh_knock(xion.ctypes.data_as(C.c_void_p), vion.ctypes.data_as(C.c_void_p))
    o.seti("impact", 1)
    o.knock(1)
    Xo, Vo, _ = o.state()
    nsg = o.geti("nsg")
    np.testing.assert_array_equal(xion, Xo[nsg])
    np.testing.assert_array_equal(vion, Vo[nsg])


def test_writers_layout(fh, tmp_path, oracle_lib):
    n, nsg, lb, X, V, at, xs, Xi, xr = _setup(fh, tmp_path, seed=-77)
    os.makedirs(tmp_path / "data", exist_ok=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        assert fh.fh_open_outputs(25) == 0
        fh.fh_writeatoms(0, 0.0, 5.25)
        fh.fh_writeatoms(25, 5.0e-15, 6.5)
        fh.fh_close_outputs(25)
    finally:
        os.chdir(cwd)
    lines = open(tmp_path / "data" / "mdrun2.xyz").read().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    assert len(lines) == 2 * (n + 2)
    hdr = "Time (ps) = " + f"{0.0:12.5f}" + "       x (A)       y (A)       z (A)" + \
        "    type   proc#" + "         KE (eV)"
    assert lines[0] == f"{n:9d}"
    assert lines[1] == hdr
    assert lines[n + 3] == "Time (ps) = " + f"{5.0e-15 * 1e12:12.5f}" + hdr[24:]
    m = {1: 46.637063e-27, 3: 1.157777417e-25}
    for i in (0, 1, nsg - 1, nsg, n - 1):
        lab = "Si" if at[i] == 1 else "Ga"
        ke = 0.5 * m[at[i]] * (V[i] ** 2).sum() * 6.252e18
        exp = "".join(f"{x * 1e10:12.4f}" for x in X[i]) + f"{lab:>8s}" + f"{0:8d}" + f"{ke:14.3E}"
        assert lines[2 + i] == exp, (i, lines[2 + i], exp)
    snap = open(tmp_path / "data" / "3dmol2.xyz").read().split("\n")[:-1]
    assert len(snap) == n + 2
    assert snap[0] == f"{n:9d}"
    assert snap[1] == "        type       x (A)       y (A)       z (A)Time (ps) = " + \
        f"{5.0e-15 * 1e12:12.5f}"
    for i in (0, n - 1):
        lab = "Si" if at[i] == 1 else "Ga"
        assert snap[2 + i] == f"{lab:>8s}" + "".join(f"{x * 1e10:10.4f}" for x in X[i])


def _records(path):
    """The records of an unformatted sequential file: 4-byte length markers
    before and after each record (the layout of gfortran and flang)."""
    b = open(path, "rb").read()
    out, p = [], 0
    while p < len(b):
        n = int(np.frombuffer(b[p:p + 4], np.int32)[0])
        out.append(b[p + 4:p + 4 + n])
        assert int(np.frombuffer(b[p + 4 + n:p + 8 + n], np.int32)[0]) == n
        p += 8 + n
    return out


def test_restart_records(fh, tmp_path):
    n, nsg, lb, X, V, at, xs, Xi, xr = _setup(fh, tmp_path)
    rng = np.random.default_rng(3)
    X2 = np.asfortranarray(X + rng.normal(0, 1e-11, X.shape))
    V2 = np.asfortranarray(rng.normal(0, 500.0, V.shape))
    fh.fh_set_xv(X2.ctypes.data_as(C.c_void_p), V2.ctypes.data_as(C.c_void_p))
    for _ in range(7):
        fh.fh_ran1()
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        os.makedirs("D", exist_ok=True)
        fh.fh_write_restart(1234, 2.5e-13, 1)
        after = [fh.fh_ran1() for _ in range(5)]
        for name in ("D/restart.in_old", "D/restart.in000001_000001234"):
            r = _records(name)
            assert len(r) == 3
            assert np.frombuffer(r[0][:4], np.int32)[0] == 1234
            assert np.frombuffer(r[0][4:], np.float64)[0] == 2.5e-13
            assert len(r[1]) == n * (7 * 8 + 4)
            body = np.frombuffer(r[1][:n * 56], np.float64)
            np.testing.assert_array_equal(body[:3 * n].reshape(3, n).T, X2)
            np.testing.assert_array_equal(body[3 * n:6 * n].reshape(3, n).T, V2)
            m = body[6 * n:]
            assert np.all(m[:nsg] == 46.637063e-27) and np.all(m[nsg:] == 1.157777417e-25)
            np.testing.assert_array_equal(np.frombuffer(r[1][n * 56:], np.int32), at)
            assert np.frombuffer(r[2], np.int32)[0] == 1
        # a whichic-0 start from the record: state back, clock and counters
        # as recorded, unfired ions at rest, ion sites at their positions, and
        # the random sequence carried on
        nt0, rt, ions = C.c_int(), C.c_double(), C.c_int()
        assert fh.fh_read_restart(C.byref(nt0), C.byref(rt), C.byref(ions)) == 0
    finally:
        os.chdir(cwd)
    assert (nt0.value, rt.value, ions.value) == (1234, 2.5e-13, 1)
    X3 = np.zeros((n, 3), order="F"); V3 = np.zeros((n, 3), order="F")
    Xi3 = np.zeros((n, 3), order="F")
    at3 = np.zeros(n, np.int32); xs3 = np.zeros(n, np.int32)
    xr3 = np.zeros((EXPECTED["nrand"], 2), order="F")
    fh.fh_get_state(*(a.ctypes.data_as(C.c_void_p) for a in (X3, V3, at3, xs3, Xi3, xr3)))
    np.testing.assert_array_equal(X3, X2)
    np.testing.assert_array_equal(V3[:nsg + 1], V2[:nsg + 1])
    assert np.all(V3[nsg + 1:] == 0)
    np.testing.assert_array_equal(Xi3[nsg:], X2[nsg:])
    np.testing.assert_array_equal(Xi3[:nsg], Xi[:nsg])
    np.testing.assert_array_equal(at3, at)
    np.testing.assert_array_equal(xr3, xr)
    assert [fh.fh_ran1() for _ in range(5)] == after
