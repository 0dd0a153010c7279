"""numpy-level wrapper of the C-ABI (libfibgpu.so).

Arrays follow the reference's conventions: per-atom vectors are (natm, 3) in
SI units, atom types 1 = Si, 2 = Ge, 3 = Ga, list indices 1-based.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class GPUError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {L.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def default_params() -> L.fib_params:
    p = L.fib_params()
    L.lib().fib_default_params(C.byref(p))
    return p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(x) -> np.ndarray:
    """(N,3) -> contiguous Fortran-order float64 (the reference's X(Natm,3))."""
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


@dataclass
class StepReport:
    lt: int
    time: float
    temp: float
    ts: float
    steps_done: int
    ended: bool
    output_due: bool
    regime: int
    nlist_builds: int
    atoms_due: bool


class FibGPU:
    """One handle of the library: one or more independent instances."""

    def __init__(self, instances, precision: str = "fp64", params: L.fib_params | None = None):
        """instances: list of (natm, nsg, nlj, lb[3])."""
        self.L = L.lib()
        self.params = params if params is not None else default_params()
        n = len(instances)
        arr = (L.fib_instance_desc * n)()
        self.inst = []
        for k, (natm, nsg, nlj, lb) in enumerate(instances):
            arr[k].natm, arr[k].nsg, arr[k].nlj = int(natm), int(nsg), int(nlj)
            for a in range(3):
                arr[k].lb[a] = float(lb[a])
            self.inst.append((int(natm), int(nsg), int(nlj), tuple(float(v) for v in lb)))
        prec = {"fp64": L.FIB_PREC_FP64, "mixed": L.FIB_PREC_MIXED}[precision]
        h = C.c_void_p()
        rc = self.L.fib_create(C.byref(self.params), n, arr, prec, C.byref(h))
        if rc != 0:
            raise GPUError("fib_create", rc, "no gfx950 device or bad arguments")
        self.h = h
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            self.L.fib_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, name, rc):
        if rc != 0:
            msg = self.L.fib_last_error(self.h)
            raise GPUError(name, rc, msg.decode() if msg else "")

    # ---- state -----------------------------------------------------------
    def natm(self, inst=0):
        return self.inst[inst][0]

    def cells(self, inst=0):
        nc = np.zeros(3, np.int32)
        self._chk("fib_get_cells", self.L.fib_get_cells(self.h, inst, _p(nc)))
        return nc

    def set_state(self, X, V=None, atype=None, xsides=None, Xi=None, inst=0):
        n = self.natm(inst)
        X = _f(X)
        assert X.shape == (n, 3)
        V = _f(V) if V is not None else None
        at = np.ascontiguousarray(atype, dtype=np.int32)
        assert at.shape == (n,)
        xs = np.ascontiguousarray(xsides, dtype=np.int32) if xsides is not None else None
        Xi = _f(Xi) if Xi is not None else None
        self._chk("fib_set_state",
                  self.L.fib_set_state(self.h, inst, _p(X), _p(V), _p(at), _p(xs), _p(Xi)))

    def state(self, inst=0):
        n = self.natm(inst)
        X = np.zeros((n, 3), order="F"); V = np.zeros((n, 3), order="F")
        F = np.zeros((n, 3), order="F")
        self._chk("fib_get_state", self.L.fib_get_state(self.h, inst, _p(X), _p(V), _p(F)))
        return X, V, F

    def set_impact(self, impact, inst=0):
        self._chk("fib_set_impact", self.L.fib_set_impact(self.h, inst, int(impact)))

    # ---- lists -------------------------------------------------------------
    def bin_cells(self, inst=0):
        ic = np.zeros((self.natm(inst), 3), np.int32, order="F")
        self._chk("fib_bin_cells", self.L.fib_bin_cells(self.h, inst, _p(ic)))
        return ic

    def build_nlist(self):
        self._chk("fib_build_nlist", self.L.fib_build_nlist(self.h))

    def nlist(self, inst=0):
        """Half lists in the reference's CSR layout: dict(nss, nlist2, nsslj, nlistlj)."""
        n = self.natm(inst)
        l2 = C.c_int32(); llj = C.c_int32()
        self._chk("fib_get_nlist", self.L.fib_get_nlist(self.h, inst, None, None, C.byref(l2),
                                                        None, None, C.byref(llj)))
        nss = np.zeros((n, 2), np.int32, order="F")
        nsslj = np.zeros((n, 2), np.int32, order="F")
        n2 = np.zeros(max(1, l2.value), np.int32)
        nlj = np.zeros(max(1, llj.value), np.int32)
        self._chk("fib_get_nlist", self.L.fib_get_nlist(self.h, inst, _p(nss), _p(n2), C.byref(l2),
                                                        _p(nsslj), _p(nlj), C.byref(llj)))
        return dict(nss=nss, nlist2=n2[: l2.value], nsslj=nsslj, nlistlj=nlj[: llj.value])

    def half_sets(self, inst=0):
        d = self.nlist(inst)
        out2, outlj = [], []
        for i in range(self.natm(inst)):
            a, b = d["nss"][i]
            out2.append(set(int(v) for v in d["nlist2"][a - 1:b]) if b >= a else set())
            a, b = d["nsslj"][i]
            outlj.append(set(int(v) for v in d["nlistlj"][a - 1:b]) if b >= a else set())
        return out2, outlj

    # ---- forces ------------------------------------------------------------
    def compute(self, energy=False, virial=False):
        flags = L.FIB_EVAL_FORCE | (L.FIB_EVAL_ENERGY if energy else 0) | \
            (L.FIB_EVAL_VIRIAL if virial else 0)
        self._chk("fib_compute", self.L.fib_compute(self.h, flags))

    def forces(self, inst=0, energy=False, virial=False):
        n = self.natm(inst)
        F = np.zeros((n, 3), order="F")
        u = np.zeros(n) if energy else None
        ut = np.zeros(1)
        W = np.zeros(9)
        self._chk("fib_get_forces", self.L.fib_get_forces(self.h, inst, _p(F), _p(u), _p(ut), _p(W)))
        out = {"F": F, "utot": float(ut[0]), "virial": W.reshape(3, 3)}
        if energy:
            out["u"] = u
        return out

    def force(self, impact=0, inst=0, energy=False, virial=False):
        self.set_impact(impact, inst)
        self.compute(energy=energy, virial=virial)
        return self.forces(inst, energy=energy, virial=virial)

    # ---- reference time integration ------------------------------------------
    def set_run(self, rp: L.fib_run_params):
        self._chk("fib_set_run", self.L.fib_set_run(self.h, 0, C.byref(rp)))

    def tint_init(self, lt0=0, time0=0.0):
        self._chk("fib_tint_init", self.L.fib_tint_init(self.h, int(lt0), float(time0)))

    def ti_begin(self, whichctrl, impact=0):
        self._chk("fib_ti_begin", self.L.fib_ti_begin(self.h, int(whichctrl), int(impact)))

    def knock(self, knc, xion, vion):
        x = np.ascontiguousarray(xion, dtype=np.float64)
        v = np.ascontiguousarray(vion, dtype=np.float64)
        self._chk("fib_knock", self.L.fib_knock(self.h, int(knc), _p(x), _p(v)))

    def step(self, whichctrl, max_steps=0, stop_on_output=False) -> StepReport:
        ctl = L.fib_step_ctl(int(whichctrl), int(max_steps), 1 if stop_on_output else 0, 0)
        rep = L.fib_step_report()
        self._chk("fib_step", self.L.fib_step(self.h, C.byref(ctl), C.byref(rep)))
        return StepReport(rep.lt, rep.time, rep.temp, rep.ts, rep.steps_done, bool(rep.ended),
                          bool(rep.output_due), rep.regime, rep.nlist_builds, bool(rep.atoms_due))

    def temperature(self, inst=0):
        t = np.zeros(1)
        self._chk("fib_temperature", self.L.fib_temperature(self.h, inst, _p(t)))
        return float(t[0])

    def timing(self, enable=None, only=None):
        """timing(enable=True/False[, only=names]) starts/stops the device
        kernel timers (all, or the named ones); timing() returns
        {timer name: {ms, calls, avg_ms}} since the start."""
        if enable is not None:
            mask = 0
            if enable:
                names = [self.L.fib_timer_name(k) for k in range(L.FIB_NTIMERS)]
                for k, nm in enumerate(names):
                    if nm is not None and (only is None or nm.decode() in only):
                        mask |= 1 << k
            self._chk("fib_timing_enable", self.L.fib_timing_enable(self.h, mask))
            return None
        t = L.fib_timing()
        self._chk("fib_get_timing", self.L.fib_get_timing(self.h, C.byref(t)))
        out = {}
        for k in range(L.FIB_NTIMERS):
            name = self.L.fib_timer_name(k)
            if name is None:
                break
            n = int(t.calls[k])
            if n:
                out[name.decode()] = {"ms": float(t.ms[k]), "calls": n, "avg_ms": float(t.ms[k]) / n}
        return out

    # ---- MD-step mode (batched impacts, adaptive dt) ------------------------
    def md_begin(self, surface_z, bottom_z, dt_max=1e-15, dx_max=0.05e-10, remove_height=10e-10,
                 shell_width=0.0, ttar=300.0, tau=1e-13, displaced_r=1.5e-10, use_graph=True):
        p = L.fib_md_params(dt_max, dx_max, remove_height, shell_width, ttar, tau, displaced_r,
                            1 if use_graph else 0, 0)
        sz = np.ascontiguousarray(surface_z, dtype=np.float64)
        bz = np.ascontiguousarray(bottom_z, dtype=np.float64)
        assert sz.shape == (len(self.inst),) and bz.shape == (len(self.inst),)
        self._chk("fib_md_begin", self.L.fib_md_begin(self.h, C.byref(p), _p(sz), _p(bz)))

    def md_fire(self, inst, ion, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk("fib_md_fire", self.L.fib_md_fire(self.h, int(inst), int(ion), _p(x), _p(v)))

    def md_fire_many(self, insts, ions, xs, vs):
        """Fire ions[q] of instance insts[q] from xs[q] with velocity vs[q],
        all at once (one list rebuild of the fired instances)."""
        inst = np.ascontiguousarray(insts, dtype=np.int32)
        ion = np.ascontiguousarray(ions, dtype=np.int32)
        x = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(vs, dtype=np.float64).reshape(-1, 3)
        assert len(inst) == len(ion) == len(x) == len(v)
        self._chk("fib_md_fire_many", self.L.fib_md_fire_many(self.h, len(inst), _p(inst), _p(ion),
                                                              _p(x), _p(v)))

    def md_use_graph(self, on):
        self._chk("fib_md_use_graph", self.L.fib_md_use_graph(self.h, 1 if on else 0))

    def set_list_width(self, wsw, wion):
        self._chk("fib_set_list_width", self.L.fib_set_list_width(self.h, int(wsw), int(wion)))

    def set_tiles(self, tri_ns=0, tri_batch=0, nb_cells=0):
        """Launch tiles (0: default); results do not depend on them."""
        self._chk("fib_set_tiles", self.L.fib_set_tiles(self.h, int(tri_ns), int(tri_batch),
                                                        int(nb_cells)))

    def tiles(self):
        t = np.zeros(3, np.int32)
        self._chk("fib_get_tiles", self.L.fib_get_tiles(self.h, _p(t)))
        return tuple(int(v) for v in t)

    def list_width(self):
        a, b = C.c_int32(), C.c_int32()
        self._chk("fib_get_list_width", self.L.fib_get_list_width(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def md_run(self, nsteps):
        self._chk("fib_md_run", self.L.fib_md_run(self.h, int(nsteps)))

    def md_run_begin(self, nsteps):
        """Enqueue nsteps MD steps and return at once (md_run_end waits)."""
        self._chk("fib_md_run_begin", self.L.fib_md_run_begin(self.h, int(nsteps)))

    def md_run_end(self):
        self._chk("fib_md_run_end", self.L.fib_md_run_end(self.h))

    def snapshot_begin(self, inst=0, slot=0):
        """Order an asynchronous copy of instance inst's X and V into `slot`."""
        self._chk("fib_snapshot_begin", self.L.fib_snapshot_begin(self.h, int(inst), int(slot)))
        self._snap_n = getattr(self, "_snap_n", {})
        self._snap_n[slot] = self.natm(inst)

    def snapshot_wait(self, slot=0):
        """X, V of the snapshot in `slot`, once its copy has arrived."""
        n = getattr(self, "_snap_n", {}).get(slot, 0)
        X = np.zeros((n, 3), order="F"); V = np.zeros((n, 3), order="F")
        self._chk("fib_snapshot_wait", self.L.fib_snapshot_wait(self.h, int(slot), _p(X), _p(V)))
        return X, V

    def md_report(self, inst=0):
        r = L.fib_md_report()
        self._chk("fib_md_get_report", self.L.fib_md_get_report(self.h, int(inst), C.byref(r)))
        return {k: getattr(r, k) for k, _ in L.fib_md_report._fields_}

    def md_energy(self):
        """Potential energy (J) of each instance's current state."""
        e = np.zeros(len(self.inst), dtype=np.float64)
        self._chk("fib_md_energy", self.L.fib_md_energy(self.h, _p(e)))
        return e

    def md_ledger(self):
        """Energy ledger per instance (J): kinetic + potential energy plus
        what the shells took and removed atoms carried out, minus what fired
        ions brought in.  Constant up to the integration error."""
        ep = self.md_energy()
        out = []
        for k in range(len(self.inst)):
            r = self.md_report(k)
            out.append(r["ke"] + ep[k] + r["e_bath"] + r["e_removed"] - r["e_fired"])
        return np.array(out)

    def synchronize(self):
        self._chk("fib_synchronize", self.L.fib_synchronize(self.h))


def run_params_from_siga(v: dict) -> L.fib_run_params:
    """fib_run_params from the siga.in values (prms.f90:223-275 names)."""
    rp = L.fib_run_params()
    rp.ts_i, rp.ts_r = v["ts_i"], v["ts_r"]
    rp.tau_i, rp.tau_r = v["tau_i"], v["tau_r"]
    rp.ttar, rp.teps, rp.ke_limit_1 = v["ttar1"], v["teps"], v["ke_limit_1"]
    rp.dti, rp.sidewidth = v["dti"], v["sidewidth"]
    for k in range(4):
        rp.outz[k] = v["outz"][k]
    rp.atomz[0], rp.atomz[1] = v["atomz"]
    rp.ionz = v["ionz"]
    rp.ntlist_i, rp.ntlist_r = v["ntlist_i"], v["ntlist_r"]
    rp.atlist_i, rp.atlist_r = v["atlist_i"], v["atlist_r"]
    rp.dslist_i, rp.dslist_r = v["dslist_i"], v["dslist_r"]
    rp.atm_out_i, rp.atm_out_r, rp.res_out = v["atm_out_i"], v["atm_out_r"], v["res_out"]
    return rp
