"""ctypes binding of include/fibgpu.h (libfibgpu.so)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfibgpu.so")
CSRC = os.path.join(HERE, "csrc")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fibgpu.h")

FIB_OK, FIB_ERR_ARG, FIB_ERR_HIP, FIB_ERR_OVERFLOW = 0, 1, 2, 3
FIB_ERR_TABLE, FIB_ERR_NONFINITE, FIB_ERR_STATE, FIB_ERR_NODEVICE = 4, 5, 6, 7
FIB_PREC_FP64, FIB_PREC_MIXED = 0, 1
FIB_NOALLOC_MAX_ATOMS = 16384  # include/fibgpu.h
FIB_EVAL_FORCE, FIB_EVAL_ENERGY, FIB_EVAL_VIRIAL = 1, 2, 4
FIB_TI_EQUILIBRATE, FIB_TI_IMPACT, FIB_TI_FREE = 0, 1, 2

STATUS_NAMES = {0: "FIB_OK", 1: "FIB_ERR_ARG", 2: "FIB_ERR_HIP", 3: "FIB_ERR_OVERFLOW",
                4: "FIB_ERR_TABLE", 5: "FIB_ERR_NONFINITE", 6: "FIB_ERR_STATE",
                7: "FIB_ERR_NODEVICE"}


class fib_params(C.Structure):
    _fields_ = [
        ("sw_A", C.c_double), ("sw_B", C.c_double), ("sw_p", C.c_double), ("sw_q", C.c_double),
        ("sw_a", C.c_double), ("sw_lambda_si", C.c_double), ("sw_lambda_ge", C.c_double),
        ("sw_gamma", C.c_double),
        ("eps_si", C.c_double), ("eps_ge", C.c_double), ("sigma_si", C.c_double),
        ("sigma_ge", C.c_double),
        ("mass_si", C.c_double), ("mass_ge", C.c_double), ("mass_ga", C.c_double),
        ("z_si", C.c_double), ("z_ge", C.c_double), ("z_ga", C.c_double),
        ("mol_c", C.c_double * 3), ("mol_b", C.c_double * 3),
        ("bohr", C.c_double), ("kB", C.c_double), ("eps0", C.c_double), ("ec", C.c_double),
        ("listfac", C.c_double), ("ntable", C.c_int32),
        ("mixed_fp64_moliere_r", C.c_double),
    ]


class fib_instance_desc(C.Structure):
    _fields_ = [("natm", C.c_int32), ("nsg", C.c_int32), ("nlj", C.c_int32),
                ("reserved", C.c_int32), ("lb", C.c_double * 3)]


class fib_run_params(C.Structure):
    _fields_ = [
        ("ts_i", C.c_double), ("ts_r", C.c_double), ("tau_i", C.c_double),
        ("tau_r", C.c_double), ("ttar", C.c_double), ("teps", C.c_double),
        ("ke_limit_1", C.c_double), ("dti", C.c_double), ("sidewidth", C.c_double),
        ("outz", C.c_double * 4), ("atomz", C.c_double * 2), ("ionz", C.c_double),
        ("ntlist_i", C.c_int32), ("ntlist_r", C.c_int32), ("atlist_i", C.c_int32),
        ("atlist_r", C.c_int32), ("dslist_i", C.c_int32), ("dslist_r", C.c_int32),
        ("atm_out_i", C.c_int32), ("atm_out_r", C.c_int32), ("res_out", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fib_step_ctl(C.Structure):
    _fields_ = [("whichctrl", C.c_int32), ("max_steps", C.c_int32),
                ("stop_on_output", C.c_int32), ("reserved", C.c_int32)]


class fib_step_report(C.Structure):
    _fields_ = [("lt", C.c_int64), ("time", C.c_double), ("temp", C.c_double),
                ("ts", C.c_double), ("steps_done", C.c_int32), ("ended", C.c_int32),
                ("output_due", C.c_int32), ("regime", C.c_int32),
                ("nlist_builds", C.c_int32), ("atoms_due", C.c_int32)]


FIB_NTIMERS = 16


class fib_md_params(C.Structure):
    _fields_ = [("dt_max", C.c_double), ("dx_max", C.c_double), ("remove_height", C.c_double),
                ("shell_width", C.c_double), ("ttar", C.c_double), ("tau", C.c_double),
                ("displaced_r", C.c_double), ("use_graph", C.c_int32), ("reserved", C.c_int32)]


class fib_md_report(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("ion_z", C.c_double),
                ("ke", C.c_double), ("steps", C.c_int64), ("removed", C.c_int32),
                ("ions_removed", C.c_int32), ("displaced", C.c_int32), ("builds", C.c_int32),
                ("e_bath", C.c_double), ("e_removed", C.c_double), ("e_fired", C.c_double),
                ("dt_min", C.c_double)]


class fib_timing(C.Structure):
    _fields_ = [("ms", C.c_double * FIB_NTIMERS), ("calls", C.c_int32 * FIB_NTIMERS)]


vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
P = C.POINTER
SIGNATURES = {
    "fib_default_params": ([P(fib_params)], None),
    "fib_create": ([P(fib_params), i32, P(fib_instance_desc), i32, P(vp)], i32),
    "fib_destroy": ([vp], i32),
    "fib_last_error": ([vp], C.c_char_p),
    "fib_get_cells": ([vp, i32, vp], i32),
    "fib_set_state": ([vp, i32, vp, vp, vp, vp, vp], i32),
    "fib_get_state": ([vp, i32, vp, vp, vp], i32),
    "fib_set_impact": ([vp, i32, i32], i32),
    "fib_bin_cells": ([vp, i32, vp], i32),
    "fib_build_nlist": ([vp], i32),
    "fib_get_nlist": ([vp, i32, vp, vp, vp, vp, vp, vp], i32),
    "fib_compute": ([vp, i32], i32),
    "fib_get_forces": ([vp, i32, vp, vp, vp, vp], i32),
    "fib_force": ([vp, i32, i32, vp, vp, vp, vp], i32),
    "fib_set_run": ([vp, i32, P(fib_run_params)], i32),
    "fib_tint_init": ([vp, i64, f64], i32),
    "fib_ti_begin": ([vp, i32, i32], i32),
    "fib_knock": ([vp, i32, vp, vp], i32),
    "fib_step": ([vp, P(fib_step_ctl), P(fib_step_report)], i32),
    "fib_temperature": ([vp, i32, vp], i32),
    "fib_timing_enable": ([vp, i32], i32),
    "fib_get_timing": ([vp, P(fib_timing)], i32),
    "fib_timer_name": ([i32], C.c_char_p),
    "fib_md_begin": ([vp, P(fib_md_params), vp, vp], i32),
    "fib_md_fire": ([vp, i32, i32, vp, vp], i32),
    "fib_md_fire_many": ([vp, i32, vp, vp, vp, vp], i32),
    "fib_md_use_graph": ([vp, i32], i32),
    "fib_force_noalloc": ([P(fib_params), i32, i32, i32, i32, vp, i32, vp, vp, vp, vp, vp], i32),
    "fib_set_list_width": ([vp, i32, i32], i32),
    "fib_get_list_width": ([vp, vp, vp], i32),
    "fib_md_run": ([vp, i32], i32),
    "fib_md_run_begin": ([vp, i32], i32),
    "fib_set_tiles": ([vp, i32, i32, i32], i32),
    "fib_get_tiles": ([vp, vp], i32),
    "fib_md_run_end": ([vp], i32),
    "fib_snapshot_begin": ([vp, i32, i32], i32),
    "fib_snapshot_wait": ([vp, i32, vp, vp], i32),
    "fib_md_get_report": ([vp, i32, P(fib_md_report)], i32),
    "fib_md_energy": ([vp, vp], i32),
    "fib_synchronize": ([vp], i32),
}


def build(arch: str = "gfx950") -> str:
    """Compile libfibgpu.so in place with hipcc (cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"ARCH={arch}", "OUT=" + LIBPATH], check=True)
    return LIBPATH


_L = None


def lib() -> C.CDLL:
    """Load libfibgpu.so; raises if it has not been built (no fallback)."""
    global _L
    if _L is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with fib_md_amd._lib.build() "
                "(there is no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _L = L
    return _L
