"""ctypes wrapper of the CPU oracle (fib_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the product package.  It is the
checker the HIP library is compared against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfiboracle.so")


class Siga(C.Structure):
    """The 42 siga.in values in read order (prms.f90:223-275)."""

    _fields_ = [
        ("amorcrys", C.c_int32), ("whichic", C.c_int32), ("startimpact", C.c_int32),
        ("eV", C.c_double),
        ("nc", C.c_int32 * 3), ("npc", C.c_int32 * 3),
        ("nlj", C.c_int32), ("nt", C.c_int32),
        ("ts_i", C.c_double), ("ts_r", C.c_double),
        ("ttar1", C.c_double), ("ke_limit_1", C.c_double), ("ke_limit_2", C.c_double),
        ("tau_i", C.c_double), ("tau_r", C.c_double), ("teps", C.c_double),
        ("uo", C.c_double),
        ("ntlist_i", C.c_int32), ("ntlist_r", C.c_int32),
        ("atlist_i", C.c_int32), ("atlist_r", C.c_int32),
        ("dslist_i", C.c_int32), ("dslist_r", C.c_int32),
        ("atm_out_i", C.c_int32), ("atm_out_r", C.c_int32),
        ("res_out", C.c_int32), ("eng_out", C.c_int32), ("tmp_out", C.c_int32),
        ("nrand", C.c_int32),
        ("outz", C.c_double * 4), ("atomz", C.c_double * 2),
        ("ionz", C.c_double), ("knockz", C.c_double), ("sidewidth", C.c_double),
        ("dti", C.c_double), ("fwhm", C.c_double),
        ("phiz", C.c_double), ("phixy", C.c_double),
    ]


class MdParams(C.Structure):
    """Config-mode step parameters (orc_md_params)."""

    _fields_ = [(k, C.c_double) for k in ("dt_max", "dx_max", "remove_height", "shell_width",
                                         "ttar", "tau", "displaced_r")]


MD_REPORT = ("time", "dt", "lam", "ke", "steps", "builds", "removed", "ions_removed",
             "displaced", "e_bath", "e_removed", "e_fired", "ion_z", "vmax", "epot", "live")


def build() -> str:
    """Compile the oracle library in place (gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
        P = C.POINTER
        sig = {
            "orc_new": ([], vp), "orc_free": ([vp], None),
            "orc_constants": ([vp], None),
            "orc_readprms": ([vp, P(Siga), i32], C.c_int),
            "orc_set_geometry": ([vp, i32, i32, i32, vp], C.c_int),
            "orc_ionrun": ([vp, i32], None),
            "orc_set_run": ([vp, P(Siga)], None),
            "orc_set_state": ([vp, vp, vp, vp, vp], None),
            "orc_get_state": ([vp, vp, vp, vp], None),
            "orc_set_force": ([vp, vp], None),
            "orc_set_xsides": ([vp, vp], None), "orc_set_xi": ([vp, vp], None),
            "orc_get_xi": ([vp, vp], None), "orc_get_xsides": ([vp, vp], None),
            "orc_get_types": ([vp, vp, vp], None), "orc_get_xrand": ([vp, vp], None),
            "orc_get_int": ([vp, C.c_char_p], i32), "orc_get_dbl": ([vp, C.c_char_p], f64),
            "orc_set_int": ([vp, C.c_char_p, i32], None),
            "orc_set_dbl": ([vp, C.c_char_p, f64], None),
            "orc_ran1": ([vp], f64), "orc_initranions": ([vp], None),
            "orc_initdata": ([vp], C.c_int), "orc_knock": ([vp, i32], None),
            "orc_init_nlist": ([vp], None), "orc_mktables": ([vp], None),
            "orc_atomcell": ([vp, vp, vp], None), "orc_si_nlist": ([vp, vp], C.c_int),
            "orc_nlist2_len": ([vp], i32), "orc_nlistlj_len": ([vp], i32),
            "orc_nbrs": ([vp], i32),
            "orc_get_nlists": ([vp, vp, vp, vp, vp, vp], None),
            "orc_get_full_counts": ([vp, vp, vp], None),
            "orc_get_full_lists": ([vp, vp, vp], None),
            "orc_itable2bd": ([vp, f64], f64), "orc_table_value": ([vp, i32], f64),
            "orc_si_force": ([vp, vp, vp, vp], None),
            "orc_si_potential": ([vp, vp, vp, vp], None),
            "orc_temperature": ([vp, vp], None),
            "orc_adjusttemp": ([vp, f64, f64], None),
            "orc_controlsidetemp": ([vp, f64], None),
            "orc_freezsides": ([vp], None), "orc_rebox": ([vp], None),
            "orc_frzsptrdatms": ([vp], None), "orc_calc_atm": ([vp], None),
            "orc_onestep": ([vp, i32, f64, i32], C.c_int),
            "orc_tint_init": ([vp], None), "orc_ti_begin": ([vp, i32], None),
            "orc_ti_iter": ([vp, i32, f64], C.c_int),
            "orc_md_begin": ([vp, P(MdParams), f64, f64], C.c_int),
            "orc_md_fire": ([vp, i32, vp, vp], C.c_int),
            "orc_md_step": ([vp], C.c_int),
            "orc_md_report": ([vp, vp], None),
            "orc_md_dead": ([vp, vp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def col(x: np.ndarray) -> np.ndarray:
    """(N,3) array -> Fortran-ordered float64 copy (reference's X(Natm,3))."""
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


class Oracle:
    """One instance of the reference's module state."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.orc_new()
        self.L.orc_constants(self.h)

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    # ---- setup -----------------------------------------------------------
    def readprms(self, siga: Siga, ranseed: int) -> None:
        rc = self.L.orc_readprms(self.h, C.byref(siga), int(ranseed))
        if rc != 0:
            raise ValueError("orc_readprms failed")

    def set_run(self, siga: Siga) -> None:
        self.L.orc_set_run(self.h, C.byref(siga))

    def set_geometry(self, natm, nsg, nlj, lb) -> None:
        lbv = np.asarray(lb, dtype=np.float64)
        self.L.orc_set_geometry(self.h, natm, nsg, nlj, _p(lbv))

    def geti(self, name):
        return int(self.L.orc_get_int(self.h, name.encode()))

    def getd(self, name):
        return float(self.L.orc_get_dbl(self.h, name.encode()))

    def seti(self, name, v):
        self.L.orc_set_int(self.h, name.encode(), int(v))

    def setd(self, name, v):
        self.L.orc_set_dbl(self.h, name.encode(), float(v))

    @property
    def natm(self):
        return self.geti("natm")

    def set_state(self, X=None, V=None, atype=None, mass=None):
        X = col(X) if X is not None else None
        V = col(V) if V is not None else None
        at = np.ascontiguousarray(atype, dtype=np.int32) if atype is not None else None
        ms = np.ascontiguousarray(mass, dtype=np.float64) if mass is not None else None
        self.L.orc_set_state(self.h, _p(X), _p(V), _p(at), _p(ms))

    def state(self):
        n = self.natm
        X = np.zeros((n, 3), order="F"); V = np.zeros((n, 3), order="F")
        F = np.zeros((n, 3), order="F")
        self.L.orc_get_state(self.h, _p(X), _p(V), _p(F))
        return X, V, F

    def types(self):
        n = self.natm
        at = np.zeros(n, np.int32); ms = np.zeros(n)
        self.L.orc_get_types(self.h, _p(at), _p(ms))
        return at, ms

    def set_force(self, F):
        F = col(F)
        self.L.orc_set_force(self.h, _p(F))

    def set_xsides(self, xs):
        xs = np.ascontiguousarray(xs, dtype=np.int32)
        self.L.orc_set_xsides(self.h, _p(xs))

    def xsides(self):
        xs = np.zeros(self.natm, np.int32)
        self.L.orc_get_xsides(self.h, _p(xs))
        return xs

    def set_xi(self, Xi):
        Xi = col(Xi)
        self.L.orc_set_xi(self.h, _p(Xi))

    def xi(self):
        Xi = np.zeros((self.natm, 3), order="F")
        self.L.orc_get_xi(self.h, _p(Xi))
        return Xi

    def xrand(self):
        n = self.geti("nrand")
        xr = np.zeros((n, 2), order="F")
        self.L.orc_get_xrand(self.h, _p(xr))
        return xr

    # ---- RNG / init ----------------------------------------------------------
    def ran1(self) -> float:
        return float(self.L.orc_ran1(self.h))

    def initranions(self):
        self.L.orc_initranions(self.h)

    def initdata(self):
        if self.L.orc_initdata(self.h) != 0:
            raise ValueError("orc_initdata failed")

    def knock(self, knc):
        self.L.orc_knock(self.h, int(knc))

    # ---- lists ---------------------------------------------------------------
    def init_nlist(self):
        self.L.orc_init_nlist(self.h)

    def mktables(self):
        self.L.orc_mktables(self.h)

    def atomcell(self, X):
        X = col(X)
        ic = np.zeros((self.natm, 3), np.int32, order="F")
        self.L.orc_atomcell(self.h, _p(X), _p(ic))
        return ic

    def si_nlist(self, X):
        X = col(X)
        return int(self.L.orc_si_nlist(self.h, _p(X)))

    def nlists(self):
        n = self.natm
        nss = np.zeros((n, 2), np.int32, order="F")
        nsslj = np.zeros((n, 2), np.int32, order="F")
        n2 = np.zeros(max(1, self.L.orc_nlist2_len(self.h)), np.int32)
        nlj = np.zeros(max(1, self.L.orc_nlistlj_len(self.h)), np.int32)
        n3 = np.zeros((max(1, self.L.orc_nbrs(self.h)), 3), np.int32)
        self.L.orc_get_nlists(self.h, _p(nss), _p(n2), _p(nsslj), _p(nlj), _p(n3))
        return dict(nss=nss, nlist2=n2, nsslj=nsslj, nlistlj=nlj,
                    nlist3=n3[: self.L.orc_nbrs(self.h)])

    def half_sets(self):
        """Per-atom sets of 1-based half-list neighbours (j > i)."""
        d = self.nlists()
        out2, outlj = [], []
        for i in range(self.natm):
            a, b = d["nss"][i]
            out2.append(set(int(v) for v in d["nlist2"][a - 1:b]) if b >= a else set())
            a, b = d["nsslj"][i]
            outlj.append(set(int(v) for v in d["nlistlj"][a - 1:b]) if b >= a else set())
        return out2, outlj

    def full_lists(self):
        n = self.natm
        c2 = np.zeros(n, np.int32); clj = np.zeros(n, np.int32)
        self.L.orc_get_full_counts(self.h, _p(c2), _p(clj))
        f2 = np.zeros(max(1, int(c2.sum())), np.int32)
        flj = np.zeros(max(1, int(clj.sum())), np.int32)
        self.L.orc_get_full_lists(self.h, _p(f2), _p(flj))
        return c2, f2[: int(c2.sum())], clj, flj[: int(clj.sum())]

    # ---- forces --------------------------------------------------------------
    def itable2bd(self, ue):
        return float(self.L.orc_itable2bd(self.h, float(ue)))

    def table(self, i1):
        return float(self.L.orc_table_value(self.h, int(i1)))

    def si_force(self, X, virial=False):
        X = col(X)
        F = np.zeros((self.natm, 3), order="F")
        W = np.zeros(9) if virial else None
        self.L.orc_si_force(self.h, _p(X), _p(F), _p(W))
        return (F, W.reshape(3, 3)) if virial else F

    def si_potential(self, X):
        X = col(X)
        u = np.zeros(self.natm)
        ut = np.zeros(1)
        self.L.orc_si_potential(self.h, _p(X), _p(ut), _p(u))
        return float(ut[0]), u

    # ---- integration -----------------------------------------------------------
    def temperature(self):
        t = np.zeros(1)
        self.L.orc_temperature(self.h, _p(t))
        return float(t[0])

    def onestep(self, whichctrl, ttar, lt):
        return int(self.L.orc_onestep(self.h, int(whichctrl), float(ttar), int(lt)))

    def tint_init(self):
        self.L.orc_tint_init(self.h)

    def ti_begin(self, whichctrl):
        self.L.orc_ti_begin(self.h, int(whichctrl))

    def ti_iter(self, whichctrl, ttar):
        return int(self.L.orc_ti_iter(self.h, int(whichctrl), float(ttar)))

    def ionrun(self, tic):
        self.L.orc_ionrun(self.h, int(tic))

    def calc_atm(self):
        self.L.orc_calc_atm(self.h)

    def controlsidetemp(self, ttar):
        self.L.orc_controlsidetemp(self.h, float(ttar))

    def frzsptrdatms(self):
        self.L.orc_frzsptrdatms(self.h)

    # ---- config mode (the benchmark configurations' integrator) ----------------
    def md_begin(self, surface_z, bottom_z, dt_max=1e-15, dx_max=0.05e-10, remove_height=10e-10,
                 shell_width=0.0, ttar=300.0, tau=1e-13, displaced_r=1.0e-10):
        p = MdParams(dt_max, dx_max, remove_height, shell_width, ttar, tau, displaced_r)
        if self.L.orc_md_begin(self.h, C.byref(p), float(surface_z), float(bottom_z)) != 0:
            raise ValueError("orc_md_begin failed")

    def md_fire(self, ion, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if self.L.orc_md_fire(self.h, int(ion), _p(x), _p(v)) != 0:
            raise ValueError("orc_md_fire failed")

    def md_step(self, n=1):
        for _ in range(n):
            self.L.orc_md_step(self.h)

    def md_report(self) -> dict:
        out = np.zeros(16)
        self.L.orc_md_report(self.h, _p(out))
        return dict(zip(MD_REPORT, (float(v) for v in out)))

    def md_dead(self):
        d = np.zeros(self.natm, np.int32)
        self.L.orc_md_dead(self.h, _p(d))
        return d

    def adjusttemp(self, temp, ttar):
        self.L.orc_adjusttemp(self.h, float(temp), float(ttar))

    def freezsides(self):
        self.L.orc_freezsides(self.h)

    def rebox(self):
        self.L.orc_rebox(self.h)


def siga_from_values(v: dict) -> Siga:
    s = Siga()
    for k, val in v.items():
        if k in ("nc", "npc", "outz", "atomz"):
            arr = getattr(s, k)
            for i, x in enumerate(val):
                arr[i] = x
        else:
            setattr(s, k, val)
    return s


# Reference siga.in (siga.in:1-53), as the values the reference reads.
REFERENCE_SIGA = dict(
    amorcrys=0, whichic=4, startimpact=0, eV=2000.0, nc=(9, 9, 3), npc=(2, 2, 1),
    nlj=1, nt=2000, ts_i=2.0e-17, ts_r=2.0e-16, ttar1=300.0, ke_limit_1=600.0,
    ke_limit_2=10.0, tau_i=1.0e-13, tau_r=1.0e-13, teps=30.0, uo=150.0,
    ntlist_i=50, ntlist_r=20, atlist_i=50, atlist_r=20, dslist_i=500, dslist_r=500,
    atm_out_i=50, atm_out_r=500, res_out=-5000, eng_out=-10, tmp_out=-100,
    nrand=10000, outz=(1.3, 3.5, 7.5, 9.8), atomz=(4.5, 6.5), ionz=5.5, knockz=1.25,
    sidewidth=1.5, dti=1e-12, fwhm=2.0, phiz=0.0, phixy=0.0,
)
