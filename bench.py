#!/usr/bin/env python3
"""Throughput of the FIB_MD force-and-integrate path on one MI355X.

Headline workload (the cumulative-dose configuration): 16 replica Si(100)
slabs of 40x40x40 cells (8,192,000 Si + 16 Ga) in one batched handle, 3 A
Berendsen side and bottom shells at 300 K, each replica's first 2 keV
normal-incidence Ga ion fired at once.  One "step" is one MD step of every
replica: half kick, drift, pins, removal, gated list rebuilds, SW + Moliere
forces, second half kick, per-replica adaptive dt.  Reported value:
atom-steps/s of the mixed-precision path (FP32 terms from FP64 pair and
leg geometry with cos + 1/3 compensated, compensated FP32 force sums, FP64
integration and energies); the FP64 path runs beside it.
Roofline: 240 algorithmic bytes per atom-step (read X, V, F, type and 16
list indices, write X, V, F) over the step's HIP-event time.

Extras in the same JSON line: the force microbenchmark (64^3-cell crystal,
one list build + force/energy/virial, FP64 vs mixed force error), the single
500 eV impact (alone and as 1024 replicas), the 64-instance ensemble, the
30 keV impact on 2.3 M atoms and a two-impact dose series, with mixed-vs-FP64
force error and energy drift per configuration; a CPU baseline of the
oracle's MD step on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --profile fp64|mixed|md-fp64|md-mixed|md-replicas|dose-mixed|dose-fp64
                       (few steps + a 256 MiB copy, for rocprofv3)
       python bench.py --summarize-stats DIR OUT.txt CMD [LAST_STEPS]
                       (rocprofv3 --stats csv -> text; LAST_STEPS: the kernel trace's
                       dispatches of the last MD steps only)
       python bench.py --summarize-pmc FETCH_DIR WRITE_DIR OUT.txt CMD [ATOMS]
                       (rocprofv3 --pmc csv -> text + profiles/pmc_traffic.json)
       python bench.py --summarize-step-pmc FETCH_DIR WRITE_DIR OUT.txt CMD
                       (rocprofv3 --pmc csv of a dose profile -> HBM bytes per MD step)
       python bench.py --summarize-counters OUT.txt CMD DIR [DIR ...]
                       (rocprofv3 --pmc SQ/GRBM counters -> text)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per atom of one evaluation: build 88 B (read X 24 B, write
# 16 neighbour indices 64 B); force + energy 124 B (read X 24 B, type 4 B,
# list 64 B; write F 24 B, u 8 B).  Per kernel: k_nbuild carries the build's
# 88 B; of the force, k_legs reads X, type and list (92 B) and k_force3 writes
# F and u (32 B); k_tri only moves intermediate data (0 B).
KERNEL_BYTES = {"k_nbuild": 88.0, "k_legs": 92.0, "k_force3": 32.0, "k_tri": 0.0}
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")
# Algorithmic bytes per atom-step of the MD step: integrate 124 B (read X, V,
# F 72 B and type 4 B, write X, V 48 B) + force 116 B (read X 24 B, type 4 B,
# 16 list indices 64 B, write F 24 B).  Per kernel: k_pos (k_md_pos) carries
# the integrate's 124 B, k_legs reads X, type and list (92 B), k_force3
# writes F (24 B); the other kernels move intermediate data only.
STEP_BYTES = 240.0
KERNEL_BYTES_MD = {"k_pos": 124.0, "k_legs": 92.0, "k_force3": 24.0}
STEP_PMC_FILE = os.path.join(ROOT, "profiles", "md_dose_step_traffic.json")
DOSE_REPLICAS, DOSE_CELLS = 16, 40
# untimed steps after the first impact is fired, before the headline's
# warm-up and timed steps: the slabs start on perfect lattice sites, and it
# takes ~20 fs (about 250 steps at the ion's small dt) until thermal motion
# brings second neighbours within the three-body range and the list
# rebuilds reach their rate; from then on a step costs the same for the
# rest of the impact (about 20 % more than the first steps)
DOSE_SETTLE = 250
PROFILE_STEPS = 30


COPY_KIB = 256 * 1024  # the plain copy of --profile mode: 256 MiB read, 256 MiB written
EV = 1.602176634e-19  # J


def short_name(n: str) -> str:
    import re
    n = re.sub(r"^void ", "", n)
    n = re.sub(r"\(.*$", "", n)
    n = n.replace("fibk::", "")
    return n


def _last_steps_stats(d, nsteps):
    """Per-kernel rows like run_kernel_stats.csv, from the kernel trace's
    dispatches of the last `nsteps` MD steps (from the nsteps-th last
    k_md_pos dispatch on)."""
    import csv
    from collections import defaultdict
    tr = list(csv.DictReader(open(os.path.join(d, "run_kernel_trace.csv"))))
    pos = sorted(int(r["Dispatch_Id"]) for r in tr if short_name(r["Kernel_Name"]) == "k_md_pos")
    first = pos[-nsteps]
    dur = defaultdict(list)
    for r in tr:
        if int(r["Dispatch_Id"]) >= first:
            dur[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    tot = sum(sum(v) for v in dur.values())
    rows = [{"Name": k, "Calls": len(v), "AverageNs": sum(v) / len(v), "MinNs": min(v),
             "MaxNs": max(v), "Percentage": 100.0 * sum(v) / tot} for k, v in dur.items()]
    return sorted(rows, key=lambda r: -r["Percentage"])


def summarize_stats(d, out, cmd, last_steps=None):
    """Kernel table of a rocprofv3 --kernel-trace --stats --output-format csv
    run; with last_steps, only the dispatches of the last MD steps (from the
    kernel trace)."""
    import csv
    if last_steps:
        rows = _last_steps_stats(d, int(last_steps))
    else:
        rows = list(csv.DictReader(open(os.path.join(d, "run_kernel_stats.csv"))))
    lines = [f"# rocprofv3 kernel statistics", f"# command: {cmd}", "",
             f"{'kernel':44s} {'calls':>6s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'pct':>6s}"]
    for r in rows:
        lines.append(f"{short_name(r['Name'])[:44]:44s} {int(r['Calls']):6d} "
                     f"{float(r['AverageNs']) / 1e3:10.1f} {float(r['MinNs']) / 1e3:10.1f} "
                     f"{float(r['MaxNs']) / 1e3:10.1f} {float(r['Percentage']):6.2f}")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_pmc(dfetch, dwrite, out, cmd, atoms=None):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of two rocprofv3 --pmc runs,
    scaled by what the counters read for the 256 MiB copy.  Bytes per atom
    are per atom of the microbench, or of `atoms` when given (then the
    microbench traffic file that `value`'s roofline reads is left alone)."""
    import csv
    from collections import defaultdict

    def per_kernel(d, counter):
        rows = list(csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        disp = defaultdict(float)
        name = {}
        for r in rows:
            if r["Counter_Name"] != counter:
                continue
            k = int(r["Dispatch_Id"])
            disp[k] += float(r["Counter_Value"])
            name[k] = short_name(r["Kernel_Name"])
        agg = defaultdict(list)
        for k, v in disp.items():
            agg[name[k]].append(v)
        copy = max(v for k, v in disp.items() if name[k] == "__amd_rocclr_copyBuffer")
        return {k: sum(v) / len(v) for k, v in agg.items()}, copy

    fetch, fcopy = per_kernel(dfetch, "FETCH_SIZE")
    write, wcopy = per_kernel(dwrite, "WRITE_SIZE")
    ff, wf = COPY_KIB / fcopy, COPY_KIB / wcopy
    n = int(atoms) if atoms is not None else len(make_workload(64)[0])
    res = {"command": cmd, "atoms": n, "copy_fetch_kib": fcopy, "copy_write_kib": wcopy,
           "fetch_factor": ff, "write_factor": wf, "kernels": {}}
    lines = ["# HBM traffic per dispatch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB)",
             f"# command: {cmd}",
             f"# calibration: 256 MiB copy reads FETCH_SIZE {fcopy:.0f} KiB (factor {ff:.3f}), "
             f"WRITE_SIZE {wcopy:.0f} KiB (factor {wf:.3f})", "",
             f"{'kernel':36s} {'fetch_KiB':>12s} {'write_KiB':>12s} {'traffic_MB':>12s} {'B/atom':>8s}"]
    for k in sorted(fetch, key=lambda k: -(fetch[k] + write.get(k, 0))):
        if k not in write:
            continue
        tb = (fetch[k] * ff + write[k] * wf) * 1024.0
        res["kernels"][k] = {"fetch_kib": fetch[k], "write_kib": write[k], "traffic_bytes": tb,
                             "bytes_per_atom": tb / n}
        lines.append(f"{k[:36]:36s} {fetch[k]:12.0f} {write[k]:12.0f} {tb / 1e6:12.1f} "
                     f"{tb / n:8.1f}")
    open(out, "w").write("\n".join(lines) + "\n")
    if atoms is None:
        with open(PMC_FILE, "w") as f:
            json.dump(res, f, indent=1)
    print("\n".join(lines))


def summarize_counters(out, cmd, *dirs):
    """Per-kernel averages per dispatch of the SQ/GRBM counters of one or more
    rocprofv3 --pmc runs (each counter summed over the hardware blocks that
    report it), with VALU instructions per wavefront and per atom."""
    import csv
    from collections import defaultdict
    per = defaultdict(lambda: defaultdict(list))
    for d in dirs:
        disp = defaultdict(float)
        name = {}
        for r in csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))):
            key = (int(r["Dispatch_Id"]), r["Counter_Name"])
            disp[key] += float(r["Counter_Value"])
            name[key[0]] = short_name(r["Kernel_Name"])
        for (k, c), v in disp.items():
            per[name[k]][c].append(v)
    avg = {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in per.items()}
    natm = len(make_workload(64)[0])
    counters = sorted({c for cs in avg.values() for c in cs})
    lines = ["# rocprofv3 --pmc counters, average per dispatch (summed over the reporting blocks)",
             f"# command: {cmd}", f"# atoms: {natm}", ""]
    hot = [k for k in avg if any(t in k for t in ("k_nbuild", "k_tri", "k_legs", "k_force3"))]
    for k in sorted(hot):
        cs = avg[k]
        lines.append(k)
        for c in counters:
            if c in cs:
                lines.append(f"  {c:28s} {cs[c]:16.0f}")
        if "SQ_INSTS_VALU" in cs and "SQ_WAVES" in cs and cs["SQ_WAVES"] > 0:
            lines.append(f"  {'VALU instr per wavefront':28s} {cs['SQ_INSTS_VALU'] / cs['SQ_WAVES']:16.1f}")
            lines.append(f"  {'VALU instr per atom':28s} {cs['SQ_INSTS_VALU'] / natm:16.1f}")
        if "SQ_INSTS_VALU" in cs and cs.get("GRBM_GUI_ACTIVE", 0) > 0:
            # GRBM_GUI_ACTIVE sums the busy cycles of the 8 XCDs; a CU issues
            # at most one wave64 VALU instruction per cycle (4 SIMDs, 4 cycles each)
            cu_cycles = 256 * cs["GRBM_GUI_ACTIVE"] / 8
            lines.append(f"  {'VALU instr per CU cycle':28s} {cs['SQ_INSTS_VALU'] / cu_cycles:16.3f}")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_step_pmc(dfetch, dwrite, out, cmd):
    """HBM bytes per MD step of a dose profile (--profile dose-*): every
    dispatch of its last PROFILE_STEPS steps (from that step's k_md_pos on);
    their FETCH_SIZE + WRITE_SIZE (scaled by the 256 MiB copy's readings)
    summed and divided by the number of steps."""
    import csv
    from collections import defaultdict

    def per_dispatch(d, counter):
        rows = list(csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        disp, name = defaultdict(float), {}
        for r in rows:
            if r["Counter_Name"] != counter:
                continue
            k = int(r["Dispatch_Id"])
            disp[k] += float(r["Counter_Value"])
            name[k] = short_name(r["Kernel_Name"])
        return disp, name

    fd, fn = per_dispatch(dfetch, "FETCH_SIZE")
    wd, wn = per_dispatch(dwrite, "WRITE_SIZE")
    fcopy = max(v for k, v in fd.items() if fn[k] == "__amd_rocclr_copyBuffer")
    wcopy = max(v for k, v in wd.items() if wn[k] == "__amd_rocclr_copyBuffer")
    ff, wf = COPY_KIB / fcopy, COPY_KIB / wcopy

    def steps_part(disp, name, fac):
        # the profiled steps: the last PROFILE_STEPS of the run
        pos = sorted(k for k in disp if name[k] == "k_md_pos")
        nsteps = min(PROFILE_STEPS, len(pos))
        first = pos[-nsteps]
        per = defaultdict(float)
        for k, v in disp.items():
            if k >= first:
                per[name[k]] += v * fac * 1024.0
        return {k: v / nsteps for k, v in per.items()}, nsteps

    fper, nsteps = steps_part(fd, fn, ff)
    wper, _ = steps_part(wd, wn, wf)
    atoms = DOSE_REPLICAS * (8 * DOSE_CELLS ** 3 + 1)
    kern = {}
    for k in set(fper) | set(wper):
        b = fper.get(k, 0.0) + wper.get(k, 0.0)
        kern[k] = {"fetch_bytes": fper.get(k, 0.0), "write_bytes": wper.get(k, 0.0),
                   "bytes_per_step": b, "bytes_per_atom_step": b / atoms}
    tot = sum(v["bytes_per_step"] for v in kern.values())
    res = {"command": cmd, "atoms": atoms, "steps": nsteps, "fetch_factor": ff,
           "write_factor": wf, "traffic_per_step_bytes": tot,
           "traffic_per_atom_step": tot / atoms, "kernels": kern}
    lines = ["# HBM traffic per MD step of the cumulative-dose batch, rocprofv3 --pmc "
             "FETCH_SIZE and WRITE_SIZE", f"# command: {cmd}",
             f"# calibration: 256 MiB copy reads FETCH_SIZE {fcopy:.0f} KiB (factor {ff:.3f}), "
             f"WRITE_SIZE {wcopy:.0f} KiB (factor {wf:.3f})",
             f"# the last {nsteps} steps of {atoms} atoms (after {DOSE_SETTLE} steps of settling); "
             "every dispatch from the first of them's k_md_pos on",
             "", f"{'kernel':36s} {'MB/step':>10s} {'B/atom-step':>12s}"]
    for k in sorted(kern, key=lambda k: -kern[k]["bytes_per_step"]):
        v = kern[k]
        lines.append(f"{k[:36]:36s} {v['bytes_per_step'] / 1e6:10.1f} {v['bytes_per_atom_step']:12.1f}")
    lines.append(f"{'total':36s} {tot / 1e6:10.1f} {tot / atoms:12.1f}")
    open(out, "w").write("\n".join(lines) + "\n")
    with open(STEP_PMC_FILE, "w") as f:
        json.dump(res, f, indent=1)
    print("\n".join(lines))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(ncell: int, seed: int = 1002):
    from fib_md_amd import configs
    X, at, L = configs.force_microbench(ncell, seed=seed)
    return np.asfortranarray(X), at, L


def run_gpu(X, at, L, precision, steps, warmup):
    import torch
    from fib_md_amd import FibGPU
    n = len(X)
    g = FibGPU([(n, n, 0, L)], precision=precision)
    g.set_state(X, atype=at)
    for _ in range(warmup):
        g.build_nlist()
        g.compute(energy=True, virial=True)
    g.synchronize()
    torch.cuda.synchronize()
    # kernel timers (HIP events on the library's stream) over the timed steps,
    # on the build and force kernels only
    g.timing(enable=True, only=("k_nbuild", "k_legs", "k_tri", "k_force3"))
    t0 = time.perf_counter()
    for _ in range(steps):
        g.build_nlist()
        g.compute(energy=True, virial=True)
    g.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tm = g.timing()
    for v in tm.values():
        v["calls_per_step"] = v["calls"] / steps
    res = g.forces(energy=False, virial=True)
    g.timing(enable=False)
    out = dict(seconds=t1 - t0, ms_per_step=(t1 - t0) * 1e3 / steps,
               rate=n * steps / (t1 - t0), timing=tm, F=res["F"], utot=res["utot"],
               virial=res["virial"])
    g.close()
    return out


STEPS_PER_IMPACT = 2000  # the single-impact configuration's step count
STEP_KERNELS = ("k_pos", "k_bin", "sort", "k_nbuild", "list_aux", "k_legs", "k_tri",
                "k_extra", "k_force3", "reduce")


def _md_handle(instances, precision, use_graph=True, params=None):
    from fib_md_amd.autotune import md_handle
    return md_handle(instances, precision, params=params, use_graph=use_graph)


def _kernel_times(g, steps):
    """Per-kernel HIP-event times over `steps` steps run as ordinary launches
    (the trajectory is the graph's, bit for bit): {name: avg ms per launch,
    launches per step}."""
    g.md_use_graph(False)
    g.timing(enable=True, only=STEP_KERNELS)
    g.md_run(steps)
    tm = g.timing()
    g.timing(enable=False)
    g.md_use_graph(True)
    return {k: {"avg_ms": v["avg_ms"], "per_step": v["calls"] / steps} for k, v in tm.items()}


def md_rate(instances, precision, steps, warmup=5, use_graph=True, params=None,
            keep_reports=False, ledger=False, snapshot=False):
    """Timed MD steps of a batch: atom-steps/s over all instances (with
    keep_reports, also the per-instance cascade outputs at the end; with
    ledger, the energy ledger's drift over warmup + timed steps; with
    snapshot, the final state of every instance)."""
    import torch
    g = _md_handle(instances, precision, use_graph, params)
    L0 = g.md_ledger() if ledger else None
    g.md_run(warmup)
    b0 = [g.md_report(k)["builds"] for k in range(len(instances))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.md_run(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    n = sum(len(s["X"]) for s in instances)
    reps = [g.md_report(k) for k in range(len(instances))]
    # list builds per instance during the timed steps (each instance counts
    # its own gated rebuilds)
    nb = [r["builds"] - b for r, b in zip(reps, b0)]
    out = dict(atom_steps_per_s=n * steps / t, ms_per_step=t * 1e3 / steps, atoms=n,
               builds_per_instance_timed=float(np.mean(nb)),
               removed=sum(r["removed"] for r in reps),
               dt_min_s=min(r["dt_min"] for r in reps), dt_end_s=min(r["dt"] for r in reps),
               sim_time_s=max(r["time"] for r in reps))
    if ledger:
        d = (g.md_ledger() - L0) / EV
        out["ledger_drift_eV"] = d.tolist()
        out["fired_eV"] = [r["e_fired"] / EV for r in reps]
    if snapshot:
        out["snapshot"] = [(g.state(k)[0], s["atype"], s["L"], s["nsg"], s["nlj"], 1)
                           for k, s in enumerate(instances)]
    g.close()
    if keep_reports:
        out["per"] = _per_instance(instances, reps)
    return out


def force_error(snapshot):
    """Forces of the same configurations (one batch of instances: X, atype,
    Lb, nsg, nlj, fired ions) in FP64 and mixed precision: max / RMS per-atom
    relative error (floor 1e-3 F_rms) and the norm ratio over all atoms."""
    from fib_md_amd import FibGPU
    F = {}
    for prec in ("fp64", "mixed"):
        g = FibGPU([(len(X), nsg, nlj, L) for X, _, L, nsg, nlj, _ in snapshot], precision=prec)
        for k, (X, at, L, nsg, nlj, imp) in enumerate(snapshot):
            g.set_state(X, atype=at, inst=k)
            g.set_impact(imp, inst=k)
        g.build_nlist()
        g.compute()
        F[prec] = np.vstack([g.forces(k)["F"] for k in range(len(snapshot))])
        g.close()
    emax, erms, enorm, spread = rel_err(F["mixed"], F["fp64"], detail=True)
    return {"max_rel": emax, "rms_rel": erms, "norm_rel": enorm, "floor": "1e-3 F_rms", **spread}


def drift_summary(rf, rm):
    """Energy-ledger drift (eV) of the mixed and the FP64 run of the same
    impacts, and mixed - FP64 instance by instance."""
    a, b = np.array(rm["ledger_drift_eV"]), np.array(rf["ledger_drift_eV"])
    return {"mixed_mean_eV": float(a.mean()), "mixed_maxabs_eV": float(np.abs(a).max()),
            "fp64_mean_eV": float(b.mean()), "fp64_maxabs_eV": float(np.abs(b).max()),
            "mixed_minus_fp64_mean_eV": float((a - b).mean()),
            "mixed_minus_fp64_maxabs_eV": float(np.abs(a - b).max()),
            "fired_eV_mean": float(np.mean(rm["fired_eV"]))}
def _per_instance(instances, reps):
    """Cascade outputs per instance: sputter yield (Si removed above the
    surface), Ga projected range (surface - ion z, A; ions removed are left
    out), displaced Si (> 1.5 A from the initial site)."""
    return {
        "yield": np.array([r["removed"] for r in reps], float),
        "range_A": np.array([(s["surface_z"] - r["ion_z"]) * 1e10 if r["ion_z"] < 1e299
                             else np.nan for s, r in zip(instances, reps)]),
        "displaced": np.array([r["displaced"] for r in reps], float),
        "ions_removed": int(sum(r["ions_removed"] for r in reps)),
    }


def _compare(per, steps, n_inst):
    """Means, standard errors and the Welch z of mixed against FP64 per
    output, and the paired differences (the same impacts in both)."""
    out = {"steps": steps, "instances": n_inst,
           "ions_removed": {p: per[p]["ions_removed"] for p in per}}
    for key in ("yield", "range_A", "displaced"):
        st = {}
        for p in ("fp64", "mixed"):
            a = per[p][key][np.isfinite(per[p][key])]
            st[p] = float(a.mean())
            st[p + "_sem"] = float(a.std(ddof=1) / np.sqrt(len(a)))
        se = float(np.hypot(st["fp64_sem"], st["mixed_sem"]))
        st["z_mixed_vs_fp64"] = (st["mixed"] - st["fp64"]) / se if se > 0 else 0.0
        d = per["mixed"][key] - per["fp64"][key]
        d = d[np.isfinite(d)]
        sd = float(d.std(ddof=1) / np.sqrt(len(d))) if len(d) > 1 else 0.0
        st["paired_mean_diff"] = float(d.mean()) if len(d) else 0.0
        st["paired_z"] = st["paired_mean_diff"] / sd if sd > 0 else 0.0
        out[key] = st
    return out


def _dose_handle(precision, n_rep=DOSE_REPLICAS, ncell=DOSE_CELLS, impacts=1, listfac=None):
    """The cumulative-dose batch: `n_rep` replicas of ncell^3 cells with
    `impacts` parked ions each, 3 A Berendsen side and bottom shells at 300 K
    (tau 0.1 ps), sputtered atoms removed, Ga retained."""
    from fib_md_amd import FibGPU, configs
    from fib_md_amd.gpu import default_params
    reps, shots = configs.dose_replicas(n_rep, ncell, impacts)
    p = default_params()
    if listfac is not None:
        p.listfac = float(listfac)
    g = FibGPU([(len(s["X"]), s["nsg"], s["nlj"], s["L"]) for s in reps], precision=precision,
               params=p)
    for k, s in enumerate(reps):
        g.set_state(s["X"], V=s["V"], atype=s["atype"], xsides=s["xsides"], Xi=s["X"], inst=k)
    g.md_begin([s["surface_z"] for s in reps], [s["bottom_z"] for s in reps],
               shell_width=3e-10, ttar=300.0, tau=1e-13)
    return g, reps, shots


def _fire_all(g, shots, m):
    """Impact m (0-based) of every replica at once (one gated rebuild)."""
    n = len(shots)
    g.md_fire_many(list(range(n)), [m + 1] * n, [shots[k][m][0] for k in range(n)],
                   [shots[k][m][1] for k in range(n)])


LISTFAC_CANDIDATES = (1.10, 1.15, 1.20, 1.25)


def tune_dose_listfac(precision, steps=40, warmup=10):
    """The project's autotuner on the headline batch itself: the list radius
    factor (cell size and skin; physics unchanged, every pair within the
    cutoff is listed either way) that runs the batch's steps fastest, each
    candidate timed over `steps` steps after its first impact.  Untimed
    set-up of the headline run."""
    import torch
    table = []
    for lf in LISTFAC_CANDIDATES:
        g, reps, shots = _dose_handle(precision, listfac=lf)
        _fire_all(g, shots, 0)
        g.md_run(DOSE_SETTLE + warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.md_run(steps)
        torch.cuda.synchronize()
        table.append((lf, (time.perf_counter() - t0) * 1e3 / steps))
        g.close()
    best = min(table, key=lambda r: r[1])[0]
    return best, {f"{a:.2f}": b for a, b in table}


def tune_dose_tiles(precision, listfac, steps=30, warmup=5):
    """The autotuner's launch tiles on the headline batch at its list factor,
    after the settling steps (untimed set-up of the headline run)."""
    from fib_md_amd import autotune
    g, reps, shots = _dose_handle(precision, listfac=listfac)
    _fire_all(g, shots, 0)
    g.md_run(DOSE_SETTLE)
    best, table = autotune.tune_tiles(g, steps=steps, warmup=warmup)
    g.close()
    return best, {"/".join(str(v) for v in t): ms for t, ms in table}


def dose_rate(precision, steps, warmup, kernel_steps=20, listfac=None, tiles=None):
    """The headline: the cumulative-dose batch during its first impact,
    DOSE_SETTLE + `warmup` untimed steps after the fire, then `steps` timed
    steps (wall clock around them, and HIP events around each step's graph
    launch on the library's stream); per-kernel HIP-event times over
    `kernel_steps` further steps."""
    import torch
    g, reps, shots = _dose_handle(precision, listfac=listfac)
    if tiles is not None:
        g.set_tiles(*tiles)
    _fire_all(g, shots, 0)
    g.md_run(DOSE_SETTLE + warmup)
    b0 = [g.md_report(k)["builds"] for k in range(len(reps))]
    torch.cuda.synchronize()
    g.timing(enable=True, only=("md_step",))
    t0 = time.perf_counter()
    g.md_run(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ev = g.timing()["md_step"]
    g.timing(enable=False)
    rr = [g.md_report(k) for k in range(len(reps))]
    kern = _kernel_times(g, kernel_steps) if kernel_steps > 0 else {}
    g.close()
    n = sum(len(s["X"]) for s in reps)
    return dict(atoms=n, replicas=len(reps), steps=steps, warmup=warmup,
                listfac=listfac if listfac is not None else 1.1,
                tiles=list(tiles) if tiles is not None else [256, 4, 8],
                atom_steps_per_s=n * steps / t, ms_per_step=t * 1e3 / steps,
                step_ms_hip_events=ev["avg_ms"], step_events=ev["calls"],
                builds_per_replica_timed=float(np.mean([r["builds"] - b for r, b in zip(rr, b0)])),
                dt_min_s=min(r["dt_min"] for r in rr), kernels=kern)


def dose_run(n_rep, ncell, impacts, steps_per_impact, precision):
    """Cumulative dose: `n_rep` replicas of ncell^3 cells, `impacts` sequential
    2 keV normal Ga impacts each (all replicas fire together), the Berendsen
    side and bottom 3 A shells at 300 K throughout, sputtered atoms removed,
    Ga retained.  Sputter yield per replica after each impact, the FP64
    energy ledger after each impact (kinetic + potential + energy the shells
    took + energy removed atoms carried - energy the ions brought, minus its
    value before the first impact: the integration drift), and the rate.
    The timed region holds the firing and the steps; reports and the
    ledger's energy evaluations lie outside it."""
    import torch
    g, reps, shots = _dose_handle(precision, n_rep, ncell, impacts)
    L0 = g.md_ledger()
    yields, ke, drift, fired, bath = [], [], [], [], []
    t = 0.0
    for m in range(impacts):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _fire_all(g, shots, m)
        g.md_run(steps_per_impact)
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
        rr = [g.md_report(k) for k in range(n_rep)]
        yields.append([r["removed"] for r in rr])
        ke.append(float(np.mean([r["ke"] for r in rr])))
        drift.append((g.md_ledger() - L0) / EV)
        fired.append(float(np.mean([r["e_fired"] for r in rr])) / EV)
        bath.append(float(np.mean([r["e_bath"] for r in rr])) / EV)
    g.close()
    n = sum(len(s["X"]) for s in reps)
    y = np.array(yields, float)
    return dict(replicas=n_rep, atoms=n, impacts=impacts, steps_per_impact=steps_per_impact,
                atom_steps_per_s=n * impacts * steps_per_impact / t,
                ms_per_step=t * 1e3 / (impacts * steps_per_impact),
                cumulative_yield_mean=y.mean(1).tolist(),
                cumulative_yield_sem=(y.std(1, ddof=1) / np.sqrt(n_rep)).tolist(),
                mean_ke_J=ke, ledger_drift_eV_mean=[float(d.mean()) for d in drift],
                ledger_drift_eV_maxabs=[float(np.abs(d).max()) for d in drift],
                ledger_drift_eV_per_replica=[d.tolist() for d in drift],
                fired_eV_mean=fired, shells_took_eV_mean=bath)


def cascade_stats(instances, steps):
    """Per-instance cascade outputs after `steps` MD steps of the batch, FP64
    and mixed, compared (cascades are chaotic: the statistics, not single
    runs, must agree); the ledger drift of both runs and the force error of
    mixed against FP64 on the FP64 run's final configurations."""
    per, runs = {}, {}
    for prec in ("fp64", "mixed"):
        r = md_rate(instances, prec, steps, warmup=0, keep_reports=True, ledger=True,
                    snapshot=(prec == "fp64"))
        per[prec] = r.pop("per")
        runs[prec] = r
    out = _compare(per, steps, len(instances))
    out["energy_drift_vs_fp64"] = drift_summary(runs["fp64"], runs["mixed"])
    out["force_error_mixed_vs_fp64"] = force_error(runs["fp64"].pop("snapshot"))
    return out


def single_impact_replicas(n):
    """n replicas of the single-impact configuration (8^3-cell Si(100) slab,
    500 eV Ga at normal incidence), replica k with seed 1001 + k for its
    velocities and impact point."""
    from fib_md_amd import configs
    reps = []
    for k in range(n):
        sk = configs.si_slab(8, 1001 + k)
        x, v = configs.ion_launch(sk["L"], sk["surface_z"], 500.0, 0.0, 0.0,
                                  np.random.default_rng(1001 + k))
        sk.update(ion_x=x, ion_v=v)
        reps.append(sk)
    return reps


def md_bench(args):
    """MD-step throughput beside the headline: the single 500 eV impact on an
    8^3-cell slab (alone, and as a batch of replicas with their own seeds),
    the batched ensemble (64 slabs of 10k-250k atoms, 1-10 keV, 0-60 deg),
    mixed-batched against the plain FP64 path (one instance at a time), the
    30 keV impact on 2.3 M atoms and a shortened dose series; mixed-vs-FP64
    force error and energy drift per configuration."""
    from fib_md_amd import configs
    out = {}
    # single impact: 8x8x8 cells, 500 eV normal incidence, 300 K, seed 1001
    s = configs.si_slab(8, 1001)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 500.0, 0.0, 0.0, np.random.default_rng(1001))
    s.update(ion_x=x, ion_v=v)
    runs = {}
    for prec in ("fp64", "mixed"):
        # the whole impact: all 2000 steps of the configuration are timed
        r = md_rate([s], prec, STEPS_PER_IMPACT, ledger=True, snapshot=(prec == "fp64"))
        r["impacts_per_hour"] = 3600.0 / (r["ms_per_step"] * 1e-3 * STEPS_PER_IMPACT)
        runs[prec] = r
        log(f"  single impact {prec}: {r['ms_per_step']:.3f} ms/step, "
            f"{r['atom_steps_per_s']:.4g} atom-steps/s, ledger drift "
            f"{r['ledger_drift_eV'][0]:.3g} eV of {r['fired_eV'][0]:.0f} eV")
    fe = force_error(runs["fp64"].pop("snapshot"))
    for prec in ("fp64", "mixed"):
        out[f"single_impact_4097_{prec}"] = runs[prec]
    out["single_impact_4097_energy_drift_vs_fp64"] = drift_summary(runs["fp64"], runs["mixed"])
    out["single_impact_4097_force_error_mixed_vs_fp64"] = fe
    log(f"  single impact force error mixed vs fp64 after 2005 steps: max {fe['max_rel']:.3g}, "
        f"rms {fe['rms_rel']:.3g}")
    # the same configuration as a batch of replicas, each with its own seed
    # (velocities and impact point), all 2000 steps timed: impacts/hour of
    # the batched path, and the configuration's cascade statistics in both
    # precisions
    if args.replicas > 0:
        reps = single_impact_replicas(args.replicas)
        per = {}
        for prec in ("fp64", "mixed"):
            r = md_rate(reps, prec, STEPS_PER_IMPACT, keep_reports=True)
            per[prec] = r.pop("per")
            r["replicas"] = len(reps)
            r["impacts_per_hour"] = 3600.0 * len(reps) / (r["ms_per_step"] * 1e-3 * STEPS_PER_IMPACT)
            out[f"single_impact_4097_x{len(reps)}_{prec}"] = r
            log(f"  single impact x{len(reps)} {prec}: {r['ms_per_step']:.3f} ms/step, "
                f"{r['atom_steps_per_s']:.4g} atom-steps/s, {r['impacts_per_hour']:.4g} impacts/hour")
        cs = _compare(per, STEPS_PER_IMPACT + 5, len(reps))
        out[f"single_impact_4097_x{len(reps)}_stats"] = cs
        # the list radius factor chosen by the autotuner for this batch
        from fib_md_amd import autotune
        from fib_md_amd.gpu import default_params
        lf, table = autotune.tune_listfac(reps, "mixed")
        pt = default_params()
        pt.listfac = lf
        r = md_rate(reps, "mixed", STEPS_PER_IMPACT, params=pt)
        r.update(replicas=len(reps), listfac=lf,
                 tuning_ms_per_step={f"{a:.2f}": b for a, b, _ in table})
        r["impacts_per_hour"] = 3600.0 * len(reps) / (r["ms_per_step"] * 1e-3 * STEPS_PER_IMPACT)
        out[f"single_impact_4097_x{len(reps)}_mixed_autotuned"] = r
        log(f"  single impact x{len(reps)} mixed, autotuned listfac {lf:.2f}: "
            f"{r['ms_per_step']:.3f} ms/step, {r['impacts_per_hour']:.4g} impacts/hour")
        for key in ("yield", "range_A", "displaced"):
            st = cs[key]
            log(f"  x{len(reps)} {key:9s} fp64 {st['fp64']:.3f} +- {st['fp64_sem']:.3f}, mixed "
                f"{st['mixed']:.3f} +- {st['mixed_sem']:.3f}, z {st['z_mixed_vs_fp64']:.2f}")
    ens = configs.ensemble(args.ensemble)
    nat = sum(len(e["X"]) for e in ens)
    log(f"  ensemble: {len(ens)} instances, {nat} atoms")
    mb = md_rate(ens, "mixed", args.md_steps)
    fb = md_rate(ens, "fp64", args.md_steps)
    # the list radius factor (cell size and skin) chosen by the autotuner
    from fib_md_amd import autotune
    from fib_md_amd.gpu import default_params
    lf, table = autotune.tune_listfac(ens, "mixed")
    pt = default_params()
    pt.listfac = lf
    mt = md_rate(ens, "mixed", args.md_steps, params=pt)
    mt.update(listfac=lf, tuning_ms_per_step={f"{a:.2f}": b for a, b, _ in table})
    log(f"  ensemble mixed batched, autotuned listfac {lf:.2f}: {mt['atom_steps_per_s']:.4g} "
        f"atom-steps/s (default 1.10: {mb['atom_steps_per_s']:.4g})")
    # plain FP64 path: each instance on its own, one after the other, with
    # the step replayed as a graph and with ordinary launches
    half = max(1, args.md_steps // 2)
    seq = {}
    for name, graph in (("fp64_one_at_a_time_graph", True), ("fp64_one_at_a_time_launches", False)):
        tot_t, tot_as = 0.0, 0.0
        for e in ens:
            r = md_rate([e], "fp64", half, warmup=2, use_graph=graph)
            tot_t += r["ms_per_step"] * 1e-3 * half
            tot_as += r["atoms"] * half
        seq[name] = dict(atom_steps_per_s=tot_as / tot_t, atoms=nat,
                         impacts_per_hour=3600.0 * len(ens) / (tot_t / half * STEPS_PER_IMPACT))
    for r in (mb, fb, mt):
        r["impacts_per_hour"] = 3600.0 * len(ens) / (r["ms_per_step"] * 1e-3 * STEPS_PER_IMPACT)
    out["ensemble"] = {"instances": len(ens), "atoms": nat, "steps_timed": args.md_steps,
                       "mixed_batched": mb, "mixed_batched_autotuned": mt, "fp64_batched": fb,
                       **seq,
                       "mixed_batched_autotuned_over_fp64_one_at_a_time_graph":
                           mt["atom_steps_per_s"] / seq["fp64_one_at_a_time_graph"]["atom_steps_per_s"],
                       "mixed_batched_over_fp64_one_at_a_time_graph":
                           mb["atom_steps_per_s"] / seq["fp64_one_at_a_time_graph"]["atom_steps_per_s"],
                       "mixed_batched_over_fp64_one_at_a_time_launches":
                           mb["atom_steps_per_s"] / seq["fp64_one_at_a_time_launches"]["atom_steps_per_s"]}
    log(f"  ensemble mixed batched {mb['atom_steps_per_s']:.4g}, fp64 batched "
        f"{fb['atom_steps_per_s']:.4g}, fp64 one at a time (graph) "
        f"{seq['fp64_one_at_a_time_graph']['atom_steps_per_s']:.4g}, (launches) "
        f"{seq['fp64_one_at_a_time_launches']['atom_steps_per_s']:.4g} atom-steps/s")
    if args.stats_steps > 0:
        cs = cascade_stats(ens, args.stats_steps)
        out["ensemble"]["cascade_stats"] = cs
        for key in ("yield", "range_A", "displaced"):
            st = cs[key]
            log(f"  {key:9s} fp64 {st['fp64']:.3f} +- {st['fp64_sem']:.3f}, mixed "
                f"{st['mixed']:.3f} +- {st['mixed_sem']:.3f}, z {st['z_mixed_vs_fp64']:.2f}")
        ed, fe = cs["energy_drift_vs_fp64"], cs["force_error_mixed_vs_fp64"]
        log(f"  ensemble drift mixed {ed['mixed_mean_eV']:.3g} eV, fp64 {ed['fp64_mean_eV']:.3g} eV;"
            f" force error max {fe['max_rel']:.3g}, rms {fe['rms_rel']:.3g}")
    # the single large impact: 60 x 60 x 80 cells (2,304,000 atoms), 30 keV at
    # 7 deg; impacts/hour at 5000 steps per impact
    if args.big_steps > 0:
        s3 = configs.big_impact()
        runs = {}
        for prec in ("mixed", "fp64"):
            r = md_rate([s3], prec, args.big_steps, ledger=True, snapshot=(prec == "fp64"))
            r["impacts_per_hour"] = 3600.0 / (r["ms_per_step"] * 1e-3 * 5000)
            runs[prec] = r
            log(f"  30 keV impact, 2.3 M atoms, {prec}: {r['ms_per_step']:.3f} ms/step, "
                f"{r['atom_steps_per_s']:.4g} atom-steps/s, smallest dt {r['dt_min_s'] * 1e15:.4f} fs, "
                f"ledger drift {r['ledger_drift_eV'][0]:.3g} eV")
        fe = force_error(runs["fp64"].pop("snapshot"))
        for prec in ("mixed", "fp64"):
            out[f"big_impact_2304001_{prec}"] = runs[prec]
        out["big_impact_2304001_energy_drift_vs_fp64"] = drift_summary(runs["fp64"], runs["mixed"])
        out["big_impact_2304001_force_error_mixed_vs_fp64"] = fe
        log(f"  30 keV force error mixed vs fp64: max {fe['max_rel']:.3g}, rms {fe['rms_rel']:.3g}")
    # cumulative dose, shortened: 16 replicas of 40^3 cells, a few impacts
    if args.dose_impacts > 0:
        out["dose"] = {}
        for prec in ("mixed", "fp64"):
            r = dose_run(DOSE_REPLICAS, DOSE_CELLS, args.dose_impacts, args.dose_steps, prec)
            out["dose"][prec] = r
            log(f"  dose {prec}: {r['atoms']} atoms, {r['atom_steps_per_s']:.4g} atom-steps/s, "
                f"cumulative yield {['%.2f' % y for y in r['cumulative_yield_mean']]}, ledger drift "
                f"{['%.3g' % d for d in r['ledger_drift_eV_mean']]} eV of "
                f"{['%.0f' % e for e in r['fired_eV_mean']]} eV fired")
        # energy drift of the mixed run against the FP64 run, replica by replica
        dm = np.array(out["dose"]["mixed"]["ledger_drift_eV_per_replica"])
        df = np.array(out["dose"]["fp64"]["ledger_drift_eV_per_replica"])
        out["dose"]["ledger_drift_mixed_minus_fp64_eV"] = {
            "mean": (dm - df).mean(1).tolist(), "maxabs": np.abs(dm - df).max(1).tolist()}
    return out


def rel_err(Fg, Fo, floor_frac=1e-3, detail=False):
    d = np.sqrt(((Fg - Fo) ** 2).sum(1))
    m = np.sqrt((Fo ** 2).sum(1))
    frms = np.sqrt((m ** 2).mean())
    e = d / np.maximum(m, floor_frac * frms)
    out = (float(e.max()), float(np.sqrt((e ** 2).mean())),
           float(np.linalg.norm(Fg - Fo) / np.linalg.norm(Fo)))
    if detail:
        # how the per-atom error is spread: atoms above 1e-5, the 99.99th
        # percentile, and |F| of the worst atom against F_rms
        k = int(np.argmax(e))
        return out + ({"atoms_above_1e-5": int((e > 1e-5).sum()),
                       "p9999_rel": float(np.quantile(e, 0.9999)),
                       "worst_atom_F_over_Frms": float(m[k] / frms),
                       "worst_atom_abs_err_over_Frms": float(d[k] / frms)},)
    return out


CPU_PROCS = 16  # the host cores of one GPU's share of the machine
MASS = {1: 46.637063e-27, 2: 120.6049350e-27, 3: 1.157777417e-25}


def _cpu_md_rate(ncell, steps, seed, settle=0):
    """One process: the oracle's MD step (orc_md_step, the configurations'
    integrator around the reference's si_force and si_nlist) on one replica
    of the dose configuration's kind, an ncell^3-cell Si(100) slab with 3 A
    shells at 300 K, its 2 keV normal impact fired and `settle` untimed steps
    taken (the lattice thermalized, as in the GPU headline).  (atoms,
    seconds of the `steps` timed steps)."""
    from fib_md_amd import configs
    from oracle.oracle import Oracle
    s = configs.si_slab(ncell, seed, nions=1)
    x, v = configs.ion_launch(s["L"], s["surface_z"], 2000.0, 0.0, 0.0,
                              np.random.default_rng(seed + 104729))
    n = len(s["X"])
    o = Oracle()
    o.set_geometry(n, s["nsg"], s["nlj"], s["L"])
    o.set_state(X=s["X"], V=s["V"], atype=s["atype"],
                mass=np.array([MASS[int(t)] for t in s["atype"]]))
    o.init_nlist()
    o.mktables()
    o.si_nlist(s["X"])
    o.set_xsides(s["xsides"])
    o.set_xi(s["X"])
    o.md_begin(s["surface_z"], 0.0, shell_width=3e-10, ttar=300.0, tau=1e-13)
    o.md_fire(1, x, v)
    if settle:
        o.md_step(settle)
    t0 = time.perf_counter()
    o.md_step(steps)
    return n, time.perf_counter() - t0


def _cpu_eval_rate(ncell, reps):
    """One process: the oracle's si_nlist + si_force + si_potential on the
    ncell^3-cell microbench sample, median of `reps` evaluations."""
    from oracle.oracle import Oracle
    X, at, L = make_workload(ncell, seed=1002)
    n = len(X)
    o = Oracle()
    o.set_geometry(n, n, 0, L)
    o.set_state(X=X, V=np.zeros((n, 3)), atype=at, mass=np.full(n, MASS[1]))
    o.init_nlist()
    o.mktables()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.si_nlist(X)
        o.si_force(X)
        o.si_potential(X)
        ts.append(time.perf_counter() - t0)
    return n, float(np.median(ts))


def _parallel_workers(argv_list):
    """Run the CPU workers at once (one per core), each printing a JSON line
    of atoms and seconds; returns their results."""
    cmd = [sys.executable, os.path.abspath(__file__)]
    ps = [subprocess.Popen(cmd + a, stdout=subprocess.PIPE, text=True) for a in argv_list]
    out = []
    for p in ps:
        o, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"cpu worker failed ({p.returncode})")
        out.append(json.loads(o.strip().splitlines()[-1]))
    return out


def cpu_baseline(ncell=12, steps=60, procs=CPU_PROCS, settle=None):
    """The oracle's MD step (a scalar C restatement) on a bounded sample of
    the headline workload: `procs` replicas (seeds 3000 + r) of an
    ncell^3-cell slab, each in its own process on its own core (as the
    reference's MPI ranks would use the host), `steps` timed steps after the
    impact and the same untimed settling steps as the GPU headline;
    atom-steps/s summed over the processes."""
    settle = DOSE_SETTLE if settle is None else settle
    from oracle.oracle import build
    build()
    procs = max(1, min(procs, os.cpu_count() or 1))
    res = _parallel_workers([["--cpu-md-worker", str(ncell), str(steps), str(3000 + r),
                              str(settle)] for r in range(procs)])
    rates = [w["atoms"] * steps / w["seconds"] for w in res]
    n = res[0]["atoms"]
    return dict(value=float(sum(rates)), unit="atom-steps/s", cores=procs, kind="port",
                per_core_mean=float(np.mean(rates)),
                sample=f"{procs} replicas (seeds 3000+r) of a {ncell}^3-cell Si(100) slab "
                       f"({n} atoms, 3 A shells at 300 K), a 2 keV normal Ga impact fired, "
                       f"{settle} untimed steps, then {steps} timed MD steps, the oracle's "
                       f"orc_md_step around the "
                       f"reference's si_force/si_nlist; one process per core at once, "
                       f"rates summed ({np.mean([w['seconds'] for w in res]):.1f} s per process)")


def cpu_eval_baseline(ncell=24, reps=3, procs=CPU_PROCS):
    """The microbench's CPU line: the oracle's si_nlist + si_force +
    si_potential on a 24^3-cell sample, `procs` processes at once."""
    procs = max(1, min(procs, os.cpu_count() or 1))
    res = _parallel_workers([["--cpu-worker", str(ncell), str(reps)] for _ in range(procs)])
    rates = [w["atoms"] / w["seconds"] for w in res]
    return dict(value=float(sum(rates)), unit="atom-evals/s", cores=procs, kind="port",
                sample=f"{ncell}^3-cell microbench sample ({res[0]['atoms']} atoms), "
                       f"si_nlist+si_force+si_potential, median of {reps}; {procs} processes")


def microbench(args):
    """The force-kernel microbenchmark (64^3 cells, 2.1 M atoms): FP64 and
    mixed list build + force/energy/virial, force error of mixed vs FP64,
    roofline of its longest kernel."""
    X, at, L = make_workload(args.ncell)
    n = len(X)
    res = {}
    for prec in ("fp64", "mixed"):
        res[prec] = run_gpu(X, at, L, prec, max(args.steps, 10), args.warmup)
        log(f"  microbench {prec}: {res[prec]['ms_per_step']:.3f} ms/step, "
            f"{res[prec]['rate']:.4g} atom-evals/s")
    emax, erms, enorm, espread = rel_err(res["mixed"]["F"], res["fp64"]["F"], detail=True)
    tm = res["mixed"]["timing"]
    dom = max((k for k in tm if KERNEL_BYTES.get(k, 0) > 0), key=lambda k: tm[k]["avg_ms"])
    dur = tm[dom]["avg_ms"]
    ach = KERNEL_BYTES[dom] * n / (dur * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PMC_FILE):
        pmc = json.load(open(PMC_FILE))
        for name, k in pmc.get("kernels", {}).items():
            if name.split("<")[0] == dom and pmc.get("atoms") == n:
                traffic = k["traffic_bytes"]
    return {"atoms": n, "box_A": L[0] * 1e10,
            "mixed": {"value": res["mixed"]["rate"], "ms_per_step": res["mixed"]["ms_per_step"],
                      "kernels_avg_ms": {k: v["avg_ms"] for k, v in tm.items()}},
            "fp64": {"value": res["fp64"]["rate"], "ms_per_step": res["fp64"]["ms_per_step"],
                     "kernels_avg_ms": {k: v["avg_ms"] for k, v in res["fp64"]["timing"].items()}},
            "unit": "atom-evals/s", "mixed_over_fp64": res["mixed"]["rate"] / res["fp64"]["rate"],
            "force_error_mixed_vs_fp64": {"max_rel": emax, "rms_rel": erms, "norm_rel": enorm,
                                          "floor": "1e-3 F_rms", **espread},
            "energy_rel_diff": abs(res["mixed"]["utot"] - res["fp64"]["utot"]) /
            abs(res["fp64"]["utot"]),
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                         "bytes_per_atom": KERNEL_BYTES[dom], "avg_launch_ms": dur}}


def step_roofline(h):
    """Roofline of the headline step: 240 algorithmic bytes per atom-step over
    the step's HIP-event time (a captured graph of the step's kernels, one
    launch per step), the measured HBM traffic per step when a dose profile
    of this workload was summarized, and each kernel's share."""
    n = h["atoms"]
    ms = h["step_ms_hip_events"]
    ach = STEP_BYTES * n / (ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(STEP_PMC_FILE):
        pmc = json.load(open(STEP_PMC_FILE))
        if pmc.get("atoms") == n:
            traffic = pmc["traffic_per_step_bytes"]
    kern = {}
    for k, v in h["kernels"].items():
        b = KERNEL_BYTES_MD.get(k, 0.0)
        t = v["avg_ms"] * v["per_step"]
        kern[k] = {"ms_per_step": t, "launches_per_step": v["per_step"],
                   "bytes_per_atom": b,
                   "achieved_GBs": b * n / (v["avg_ms"] * 1e-3) / 1e9 if b > 0 else 0.0}
    dom = max((k for k in kern if kern[k]["bytes_per_atom"] > 0),
              key=lambda k: kern[k]["ms_per_step"], default=None)
    out = {"bound": "hbm", "kernel": "md_step (one captured graph per step)", "achieved": ach,
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
           "bytes_per_atom_step": STEP_BYTES, "avg_launch_ms": ms, "kernels": kern}
    if dom is not None:
        d = kern[dom]
        out["dominant_kernel"] = {"kernel": dom, "achieved": d["achieved_GBs"],
                                  "frac": d["achieved_GBs"] / HBM_PEAK_GBS,
                                  "bytes_per_atom": d["bytes_per_atom"],
                                  "avg_launch_ms": h["kernels"][dom]["avg_ms"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-md", action="store_true", help="headline only (no extras)")
    ap.add_argument("--no-tune", action="store_true",
                    help="headline with the reference's list factor 1.1 (no autotuning)")
    ap.add_argument("--md-steps", type=int, default=100)
    ap.add_argument("--ensemble", type=int, default=64)
    ap.add_argument("--replicas", type=int, default=1024,
                    help="replicas of the single-impact configuration run as one batch")
    ap.add_argument("--stats-steps", type=int, default=2000,
                    help="MD steps of the cascade-statistics run (0: skip)")
    ap.add_argument("--big-steps", type=int, default=5000,
                    help="timed MD steps of the 2.3 M-atom 30 keV impact (0: skip; "
                         "the configuration's impact is 5000 steps)")
    ap.add_argument("--dose-impacts", type=int, default=2,
                    help="sequential impacts per replica of the shortened dose run (0: skip)")
    ap.add_argument("--dose-steps", type=int, default=1500,
                    help="MD steps per impact of the dose run")
    ap.add_argument("--profile", choices=["fp64", "mixed", "md-fp64", "md-mixed", "md-replicas",
                                          "dose-mixed", "dose-fp64"], default=None)
    ap.add_argument("--summarize-stats", nargs="+", metavar="DIR OUT CMD [LAST_STEPS]")
    ap.add_argument("--listfac", type=float, default=None,
                    help="list radius factor of the dose profile (default: the reference's)")
    ap.add_argument("--tiles", type=int, nargs=3, default=None, metavar=("TRI_NS", "TRI_BATCH", "NB_CELLS"),
                    help="launch tiles of the dose profile (default: the library's)")
    ap.add_argument("--cpu-worker", nargs=2, metavar=("NCELL", "REPS"), help=argparse.SUPPRESS)
    ap.add_argument("--cpu-md-worker", nargs=4, metavar=("NCELL", "STEPS", "SEED", "SETTLE"),
                    help=argparse.SUPPRESS)
    ap.add_argument("--summarize-pmc", nargs="+", metavar="FETCH_DIR WRITE_DIR OUT CMD [ATOMS]")
    ap.add_argument("--summarize-step-pmc", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "OUT", "CMD"))
    ap.add_argument("--summarize-counters", nargs="+", metavar="OUT CMD DIR",
                    help="OUT CMD DIR [DIR ...]: SQ/GRBM counter table of rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.cpu_worker:
        # one CPU baseline process (no GPU use): atoms and median seconds
        n, t = _cpu_eval_rate(int(args.cpu_worker[0]), int(args.cpu_worker[1]))
        print(json.dumps({"atoms": n, "seconds": t}), flush=True)
        return
    if args.cpu_md_worker:
        a = args.cpu_md_worker
        n, t = _cpu_md_rate(int(a[0]), int(a[1]), int(a[2]), int(a[3]))
        print(json.dumps({"atoms": n, "seconds": t}), flush=True)
        return
    if args.summarize_stats:
        summarize_stats(*args.summarize_stats)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc[:5])
        return
    if args.summarize_step_pmc:
        summarize_step_pmc(*args.summarize_step_pmc)
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is designed, not built or measured.", flush=True)
        sys.exit(2)
    import torch
    torch.cuda.init()
    from fib_md_amd import _lib
    _lib.lib()
    if args.profile:
        # a plain copy of known size calibrates the FETCH/WRITE counters
        a = torch.empty(256 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.profile in ("dose-mixed", "dose-fp64"):
        # the headline batch at the headline's list factor: first impact,
        # the settling steps, then PROFILE_STEPS steps (for rocprofv3; the
        # summaries take the last PROFILE_STEPS steps)
        g, reps, shots = _dose_handle(args.profile[5:], listfac=args.listfac)
        if args.tiles:
            g.set_tiles(*args.tiles)
        _fire_all(g, shots, 0)
        # in chunks, with a line each: under counter collection every
        # dispatch is serialized and the settling steps take minutes
        for k in range(0, DOSE_SETTLE, 25):
            g.md_run(min(25, DOSE_SETTLE - k))
            log(f"  settling: {k + 25} steps")
        g.md_run(PROFILE_STEPS)
        g.close()
        log(f"dose profile {args.profile}: {DOSE_SETTLE} + {PROFILE_STEPS} steps")
        return
    if args.profile in ("md-fp64", "md-mixed"):
        # the batched ensemble, a few MD steps (for rocprofv3)
        from fib_md_amd import configs
        ens = configs.ensemble(args.ensemble)
        r = md_rate(ens, args.profile[3:], 60, warmup=2)
        log(json.dumps(r))
        return
    if args.profile == "md-replicas":
        # replicas of the single-impact configuration, mixed, a few MD steps
        r = md_rate(single_impact_replicas(args.replicas), "mixed", 60, warmup=2)
        log(json.dumps(r))
        return
    if args.profile:
        X, at, L = make_workload(args.ncell)
        r = run_gpu(X, at, L, args.profile, 10, 2)
        log(json.dumps({k: v for k, v in r.items() if k in ("ms_per_step", "timing")}))
        return
    # ---- headline: the cumulative-dose batch, mixed and FP64 --------------
    head, tuned, default_lf = {}, {}, {}
    for prec in ("mixed", "fp64"):
        lf, table = (None, {}) if args.no_tune else tune_dose_listfac(prec)
        tiles, ttable = (None, {}) if args.no_tune else tune_dose_tiles(prec, lf)
        tuned[prec] = {"listfac": lf if lf is not None else 1.1, "ms_per_step": table,
                       "tiles": list(tiles) if tiles is not None else [256, 4, 8],
                       "tiles_ms_per_step": ttable}
        log(f"dose batch {prec}: list factor {tuned[prec]['listfac']:.2f} (autotuner: {table}); "
            f"tiles {tuned[prec]['tiles']} ({ttable}); "
            f"{args.warmup} warmup + {args.steps} timed steps ...")
        head[prec] = dose_rate(prec, args.steps, args.warmup, listfac=lf, tiles=tiles)
        if lf is not None and abs(lf - 1.1) > 1e-9:
            d = dose_rate(prec, args.steps, args.warmup, kernel_steps=0)
            default_lf[prec] = {"atom_steps_per_s": d["atom_steps_per_s"],
                                "ms_per_step": d["ms_per_step"]}
        else:
            default_lf[prec] = {"atom_steps_per_s": head[prec]["atom_steps_per_s"],
                                "ms_per_step": head[prec]["ms_per_step"]}
        h = head[prec]
        log(f"  {prec}: {h['ms_per_step']:.3f} ms/step (HIP events {h['step_ms_hip_events']:.3f}), "
            f"{h['atom_steps_per_s']:.4g} atom-steps/s, {h['builds_per_replica_timed']:.1f} "
            f"builds per replica")
        for k, v in sorted(h["kernels"].items(), key=lambda kv: -kv[1]["avg_ms"] * kv[1]["per_step"]):
            log(f"    {k:10s} {v['avg_ms'] * v['per_step']:.4f} ms/step ({v['per_step']:.0f} launches)")
    m = head["mixed"]
    roof = step_roofline(m)
    log(f"  roofline: {roof['achieved']:.0f} GB/s of {STEP_BYTES:.0f} B/atom-step = "
        f"{100 * roof['frac']:.2f} % of HBM peak")
    extras = {}
    if not args.no_md:
        log("force microbench ...")
        extras["microbench"] = microbench(args)
        log("MD configurations ...")
        extras["md_steps"] = md_bench(args)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle MD step) ...")
        cpu = cpu_baseline()
        log(f"  cpu: {cpu['value']:.4g} atom-steps/s on {cpu['cores']} cores")
        if not args.no_md:
            extras["microbench"]["cpu_baseline"] = cpu_eval_baseline()
    n = m["atoms"]
    line = {
        "metric": "atom-steps/s on 1 MI355X, mixed vs FP64 (SW+Moliere); Ga impacts/hour",
        "value": m["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 terms (from f64 pair/leg geometry, compensated cos+1/3) + compensated f32 force sums; f64 integration and energies (mixed)",
        "data": "synthetic",
        "config": {"workload": f"cumulative dose: {DOSE_REPLICAS} replica Si(100) slabs of "
                               f"{DOSE_CELLS}^3 cells in one batch, first 2 keV normal Ga impact "
                               f"of every replica, 3 A Berendsen shells at 300 K; "
                               f"{DOSE_SETTLE} steps after the fire (thermalized lattice), "
                               f"{args.warmup} warmup + {args.steps} timed MD steps; list "
                               f"factor {m['listfac']:.2f} and tiles {m['tiles']} from the "
                               f"autotuner",
                   "atoms": n, "replicas": DOSE_REPLICAS, "parallelism": "single GPU"},
        "fp64": {"value": head["fp64"]["atom_steps_per_s"],
                 "ms_per_step": head["fp64"]["ms_per_step"],
                 "step_ms_hip_events": head["fp64"]["step_ms_hip_events"],
                 "listfac": head["fp64"]["listfac"], "tiles": head["fp64"]["tiles"]},
        "autotuner": {"knobs": "list radius factor (cell size and skin), then launch tiles "
                               "(vertex-pass stage legs / loads in flight / list-build cells "
                               "per wavefront)", **tuned},
        "reference_listfac_1.10": default_lf,
        "mixed_over_fp64": m["atom_steps_per_s"] / head["fp64"]["atom_steps_per_s"],
        "step_ms_hip_events": m["step_ms_hip_events"],
        "builds_per_replica_timed": m["builds_per_replica_timed"],
        "roofline": roof,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
