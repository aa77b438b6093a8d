#!/usr/bin/env python3
"""Benchmark of the SHLL time-step loop on one MI355X.

Headline workload (the batched FP32-compensated form the throughput target is set on):
  4096 independent second-order shock tubes (seed 20261202), N log-uniform in
  [256, 65536] cells, rho_L/rho_R in [2, 20], T_L/T_R in [0.5, 2], diaphragm in
  [0.3, 0.7], u = 0, outflow walls, each run to t = 0.2 (its own step count from the
  reference's float time loop), FP32 fluxes with a compensated (hi, lo) FP32 state
  updated by one Kahan step per cell and variable (SHLL_C32K), one launch per k steps
  over an (instance, tile) work list, HIP graphs.
A "step" is one time step of every unfinished tube.  The default run times the whole
job (every tube to t = 0.2).  value = total cell-updates / wall time of the timed region.

Also reported (extra): the same batch with the exactly conservative TwoSum update
(SHLL_C32), the single-instance configurations (1D first order 1024 cells, 1D second
order 65536 cells and 2D first order 4096^2 in FP64 and both FP32-compensated forms, and
64 uneven 2D domains), each run in full with its error against the FP64 kernel.

cpu_baseline: the oracle's reference-precision port of the reference C loop (unfused
passes, gcc -O3) on one host core, on a stratified sample of the same tubes.

cpu_baseline_omp: the same port with one OpenMP worksharing loop per pass (the base-omp
pattern) on up to 16 host threads.

--workload domains times the 64 uneven 2D first-order domains instead (same JSON line,
2D roofline), for profiling the 2D kernel.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
                       [--workload tubes|domains] [--form c32k|c32] [--launch C,k,waves,graph]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shll-sve-cfd_amd"))
# launch shapes chosen by the library's autotuner on MI355X (tune/make_tune_cache.py);
# --autotune re-tunes live instead
TUNE_CACHE = os.path.join(ROOT, "shll-sve-cfd_amd", "tune", "gfx950.txt")

METRIC = "cell-updates/s + % HBM roofline, FP32-comp vs FP64, 1D/2D SHLL, 1×MI355X"
PEAK_HBM_GBS = 8000.0
SEED = 20261202
NTUBES = 4096


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_update(dim, prec):
    """Algorithmic bytes per cell-update: read and write the conserved state once
    (2 * nvar * bytes per value; FP32-compensated stores hi and lo)."""
    import shll_amd as S
    nv = 3 if dim == 1 else 4
    s = {S.F64: 8, S.REF32: 4, S.C32: 8, S.C32K: 8}[prec]
    return 2 * nv * s


def traffic_per_update(shape, form="c32k", summary=None):
    """Calibrated HBM bytes per cell-update of the batch step kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy), when the
    summary was taken with the same launch shape (cells/lane, steps/launch, waves/block);
    else None."""
    import re
    path = os.path.join(ROOT, "profiles", summary or "k_step1d_%s_o2_tubes4096_hbm_traffic.md" % form)
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_cell_update:\s*([0-9.eE+-]+)\s*\(launch shape (\d+),(\d+),(\d+)\)",
                  open(path).read())
    if not m or tuple(int(x) for x in m.groups()[1:]) != tuple(shape[:3]):
        return None
    return float(m.group(1))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def calibration_copy(nbytes):
    """One plain device-to-device copy of nbytes (HIP runtime blit kernel), so a
    rocprofv3 --pmc run sees FETCH_SIZE / WRITE_SIZE of a transfer of known size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed for the calibration copy")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError("calibration copy failed (%d)" % rc)
    log("calibration copy: %d bytes device-to-device" % nbytes)


def copy_ceiling(nbytes=1 << 30, reps=10):
    """Measured HBM ceiling: median rate of reps device-to-device copies of nbytes (the HIP
    runtime's copy kernel; each moves 2 x nbytes), timed with HIP events, in GB/s."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed for the copy ceiling")
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    rates = []
    for r in range(reps + 1):
        hip.hipEventRecord(e0, None)
        rc = hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        if rc:
            break
        if r > 0:  # the first copy warms up
            rates.append(2 * nbytes / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    if rc or not rates:
        raise RuntimeError("copy ceiling: hipMemcpyAsync failed (%d)" % rc)
    return sorted(rates)[len(rates) // 2]


def run_batch(S, pbs, prims, steps, launch, K, W, progress=0, trial=2048):
    b = S.Batch(pbs, prims, steps)
    if launch:
        b.set_launch(*launch)
    else:
        b.autotune(trial)
    if W > 0:
        b.advance(W)
        b.reset()
    b.set_timing(True)
    if progress > 0:
        # pieces of whole launches; event time and launch counts are summed over the pieces
        k = b.get_launch()[1]
        piece = max(k, progress // k * k)
        done, ev_ms, launches = 0, 0.0, 0
        t0 = time.perf_counter()
        while done < K:
            n = min(piece, K - done)
            b.advance(n)
            done += n
            _, nl, tot = b.last_timing()
            ev_ms += tot
            launches += nl
            log("progress: %d / %d steps" % (done, K))
        wall = time.perf_counter() - t0
        return b, wall, ev_ms / max(1, launches), launches, ev_ms
    t0 = time.perf_counter()
    b.advance(K)
    wall = time.perf_counter() - t0
    avg_ms, launches, ev_ms = b.last_timing()
    return b, wall, avg_ms, launches, ev_ms


def updates_done(pbs, steps, K):
    return sum(p.nx * p.ny * min(K, s) for p, s in zip(pbs, steps))


def cpu_baseline(S, pbs, prims, steps, nthreads=1):
    """Oracle REF32 (the reference's arithmetic and pass structure) on every 256th tube of
    the batch for up to 2000 steps each: serially (the base-c loop), or with one OpenMP
    worksharing loop per pass on nthreads threads (the base-omp pattern)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    idx = list(range(0, len(pbs), 256))
    total, secs = 0, 0.0
    for k in idx:
        pb = pbs[k]
        n = min(steps[k], 2000)
        op = oracle.problem(1, 2, pb.nx)
        _, _, el = oracle.run(op, oracle.REF32, prims[k], n, nthreads=nthreads, timed=True)
        total += pb.nx * n
        secs += el
    how = "1 thread" if nthreads == 1 else "%d OpenMP threads, one worksharing loop per pass" % nthreads
    return {"value": total / secs, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "sample": "%d tubes (every 256th of the batch, N=%d..%d), min(steps,2000) steps each, "
                      "%.3g cell-updates in %.1f s; oracle REF32 = reference float arithmetic and "
                      "unfused pass structure, gcc -O3, %s; host %s" % (
                          len(idx), min(pbs[k].nx for k in idx), max(pbs[k].nx for k in idx), total, secs,
                          how, cpu_model())}


def omp_threads():
    """Threads for the OpenMP baseline: the CPUs this process may use, at most 16 (a
    one-GPU share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_rate(pb, prim, steps):
    """cell-updates/s of the oracle's REF32 port (reference arithmetic and pass structure,
    gcc -O3, one core) on one problem for a given number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    op = oracle.problem(pb.dim, pb.order, pb.nx, pb.ny, bc=pb.bc, lx=pb.lx, ly=pb.ly)
    _, _, el = oracle.run(op, oracle.REF32, prim, steps, nthreads=1, timed=True)
    return {"value": pb.nx * pb.ny * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same problem, %.1f s" % (steps, el)}


def reference_cpu():
    """The reference solvers themselves (compiled from their own sources into oracle/_ref by
    oracle/Makefile, when that build exists), timed on one core at their compiled-in sizes:
    wall time of the stage loop run by oracle/ref_harness.c for a fixed number of steps."""
    import tempfile
    runs = [("base_shll", "base-c/base_shll.c: 1D first order, 256 cells", 256, 20000),
            ("base_shll_2d", "base-c/base_shll_2d.c: 2D first order, 256^2", 256 * 256, 200),
            ("2nd_order_base_shll", "base-c/2nd_order_base_shll.c: 2D second order, 256^2", 256 * 256, 100),
            ("omp_2nd_order_base_shll", "base-omp/2nd_order_base_shll.c (MC limiter, OpenMP off): 2D second "
             "order, 1024^2", 1024 * 1024, 10)]
    out = {}
    for exe, what, cells, steps in runs:
        path = os.path.join(ROOT, "oracle", "_ref", exe + ".run")
        if not os.path.exists(path):
            continue
        with tempfile.TemporaryDirectory() as td:
            t0 = time.perf_counter()
            subprocess.run([path, "steps", os.path.join(td, "s"), "-", "0,%d" % steps], cwd=td, check=True,
                           capture_output=True, timeout=300)
            dt = time.perf_counter() - t0
        out[exe] = {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                    "sample": "%s, %d steps in %.2f s (process start, set-up and two array dumps included)"
                              % (what, steps, dt)}
    return out


def extra_single(S, cpu=True):
    """The single-instance configurations, each in full, with errors against FP64 and (cpu)
    the oracle's one-core rate on a few steps of the same problem."""
    out = {}
    # 1D first order, 1024 cells, t = 0.2, FP64
    pb = S.problem(1, 1, 1024, precision=S.F64)
    ic = S.ic_tube(1024)
    with S.Solver(pb, ic) as s:
        s.autotune(400)
        s.run_to_time(0.2)
        s2 = S.Solver(pb, ic)
        s2.set_launch(*s.get_launch())
        t0 = time.perf_counter()
        n = s2.run_to_time(0.2)
        dt = time.perf_counter() - t0
        s2.close()
    out["c0_1d_o1_f64_N1024"] = {"steps": n, "seconds": dt, "cell_updates_per_s": 1024 * n / dt}
    if cpu:
        out["c0_1d_o1_f64_N1024"]["cpu_baseline"] = cpu_rate(pb, ic, n)
    # 1D second order, 65536 cells, t = 0.2, FP64 and both FP32-compensated forms
    n1 = 65536
    ic = S.ic_tube(n1)
    sol = {}
    for name, prec in (("f64", S.F64), ("c32", S.C32), ("c32k", S.C32K)):
        pb = S.problem(1, 2, n1, precision=prec)
        s = S.Solver(pb, ic)
        s.autotune(8000)
        shape = s.get_launch()
        s.step(4000)
        s.close()
        s = S.Solver(pb, ic)
        s.set_launch(*shape)
        t0 = time.perf_counter()
        n = s.run_to_time(0.2)
        dt = time.perf_counter() - t0
        sol[name] = s
        out["c1_1d_o2_%s_N65536" % name] = {"steps": n, "seconds": dt, "cell_updates_per_s": n1 * n / dt,
                                           "launch": list(s.get_launch())}
    for name in ("c32", "c32k"):
        out["c1_1d_o2_%s_N65536" % name]["error_vs_f64"] = sol[name].compare(sol["f64"])
    if cpu:
        out["c1_1d_o2_f64_N65536"]["cpu_baseline"] = cpu_rate(S.problem(1, 2, n1), ic, 400)
    for s in sol.values():
        s.close()
    # 2D first order, 4096^2, 2000 steps, FP64 and FP32-compensated
    n2 = 4096
    ic = S.ic_circle(n2, n2)
    sol = {}
    for name, prec in (("f64", S.F64), ("c32", S.C32), ("c32k", S.C32K)):
        s = S.Solver(S.problem(2, 1, n2, n2, precision=prec), ic)
        s.step(20)
        s.close()
        s = S.Solver(S.problem(2, 1, n2, n2, precision=prec), ic)
        sums0 = s.conserved_sums()
        t0 = time.perf_counter()
        s.step(2000)
        dt = time.perf_counter() - t0
        sums1 = s.conserved_sums()
        sol[name] = s
        # reflective walls: mass and energy are conserved (the momenta change by the wall impulse)
        out["c2_2d_o1_%s_4096sq" % name] = {"steps": 2000, "seconds": dt,
                                           "cell_updates_per_s": n2 * n2 * 2000 / dt,
                                           "launch": list(s.get_launch()),
                                           "mass_drift": abs(sums1[0] - sums0[0]) / abs(sums0[0]),
                                           "energy_drift": abs(sums1[3] - sums0[3]) / abs(sums0[3])}
    for name in ("c32", "c32k"):
        out["c2_2d_o1_%s_4096sq" % name]["error_vs_f64"] = sol[name].compare(sol["f64"])
    if cpu:
        out["c2_2d_o1_f64_4096sq"]["cpu_baseline"] = cpu_rate(S.problem(2, 1, n2, n2), ic, 2)
    for s in sol.values():
        s.close()
    # 64 uneven 2D domains, first order, 1000 steps, both FP32-compensated
    # forms; the FP64 batch of the same domains judges accuracy
    pbs, prims = S.domain_batch(4242, 64, precision=S.C32K)
    cells = sum(p.nx * p.ny for p in pbs)
    ref = S.Batch([S.problem(2, 1, p.nx, p.ny, precision=S.F64, lx=p.lx, ly=p.ly) for p in pbs], prims,
                  [1000] * 64)
    t1 = time.perf_counter()
    ref.run()
    dt64 = time.perf_counter() - t1
    bad64 = int((ref.check_state() > 0).sum())
    for name, prec in (("c32k", S.C32K), ("c32", S.C32)):
        pbp = [S.problem(2, 1, p.nx, p.ny, precision=prec, lx=p.lx, ly=p.ly) for p in pbs]
        b = S.Batch(pbp, prims, [1000] * 64)
        b.advance(10)
        b.reset()
        t0 = time.perf_counter()
        b.run()
        dt = time.perf_counter() - t0
        errs = [b.compare(k, ref)["l1_rho"] / (pbs[k].lx * pbs[k].ly) for k in range(64)]
        launch = list(b.get_launch())
        nonfinite = bad64 + int((b.check_state() > 0).sum())
        b.close()
        out["c4_2d_o1_%s_64domains" % name] = {"steps": 1000, "cells": cells, "seconds": dt,
                                               "launch": launch,
                                               "cell_updates_per_s": cells * 1000 / dt,
                                               "f64_batch_cell_updates_per_s": cells * 1000 / dt64,
                                               "l1_rho_vs_f64_per_unit_area_max": max(errs),
                                               "l1_rho_vs_f64_per_unit_area_median": sorted(errs)[32],
                                               "nonfinite_domains": nonfinite}
    ref.close()
    return out


def make_judge(S, pbs, steps, prims, K):
    """The FP64 kernel run as a batch on every 16th tube in size order (tubes that finish
    within K steps), the accuracy judge of the FP32-compensated batches."""
    order = sorted(range(len(pbs)), key=lambda k: pbs[k].nx)
    sample = [k for k in order[::16] if steps[k] <= K]
    if not sample:
        return None
    ref = S.Batch([S.problem(1, 2, pbs[k].nx, precision=S.F64) for k in sample], [prims[k] for k in sample],
                  [steps[k] for k in sample])
    ref.run()
    return ref, sample, ref.check_state()


def tube_errors(S, b, pbs, judge):
    """Accuracy of an FP32-compensated batch against the FP64 judge.  The reference scheme's
    fixed time step DT = 0.125*DX assumes wave speeds near R+1; hot, strongly expanding
    tubes exceed (|u|+a)*DT/DX = 0.5 and the forward-Euler second-order scheme then goes
    non-finite in FP64 as well, so those tubes are counted and left out of the error
    statistics.  Tubes on the edge of that instability can blow up in one precision and
    not the other: the ones non-finite only in FP32-compensated are counted separately and
    also left out."""
    import numpy as np
    if judge is None:
        return None
    ref, sample, bad64 = judge
    bad32 = b.check_state()
    rows = []
    c32_only = 0
    for q, k in enumerate(sample):
        if bad64[q] > 0:
            continue
        if bad32[k] > 0:
            c32_only += 1
            continue
        e = b.compare(k, ref, q)
        rows.append((pbs[k].nx, e["l1_rho"], e["linf_rho"]))
    l1 = np.array([r[1] for r in rows])
    n = np.array([r[0] for r in rows])
    bands = {}
    for lo, hi in ((256, 1024), (1024, 4096), (4096, 16384), (16384, 65537)):
        m = (n >= lo) & (n < hi)
        if m.any():
            bands["N_%d_%d" % (lo, hi - 1 if hi == 65537 else hi)] = {
                "tubes": int(m.sum()), "l1_rho_median": float(np.median(l1[m])), "l1_rho_max": float(l1[m].max()),
                "frac_le_1e-6": float((l1[m] <= 1e-6).mean())}
    return {
        "judge": "FP64 kernel, every 16th tube in size order (%d tubes)" % len(sample),
        "tubes_nonfinite_fp64_in_sample": int((bad64 > 0).sum()),
        "tubes_nonfinite_c32_only_in_sample": c32_only,
        "tubes_nonfinite_c32_all": int((bad32 > 0).sum()),
        "compared_tubes": len(rows),
        "l1_rho_median": float(np.median(l1)) if len(l1) else None,
        "l1_rho_max": float(l1.max()) if len(l1) else None,
        "frac_l1_rho_le_1e-6": float((l1 <= 1e-6).mean()) if len(l1) else None,
        "by_size": bands,
    }


DOMAIN_SEED = 4242
NDOMAINS = 64
DOMAIN_STEPS = 1000


def cpu_baseline_2d(pbs, prims, nthreads=1, steps=40):
    """Oracle REF32 first-order 2D (the reference's arithmetic and pass structure) on every
    8th domain of the batch for a few steps each."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    idx = list(range(0, len(pbs), 8))
    total, secs = 0, 0.0
    for k in idx:
        pb = pbs[k]
        op = oracle.problem(2, 1, pb.nx, pb.ny, lx=pb.lx, ly=pb.ly)
        _, _, el = oracle.run(op, oracle.REF32, prims[k], steps, nthreads=nthreads, timed=True)
        total += pb.nx * pb.ny * steps
        secs += el
    how = "1 thread" if nthreads == 1 else "%d OpenMP threads, one worksharing loop per pass" % nthreads
    return {"value": total / secs, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "sample": "%d domains (every 8th of the batch), %d steps each, %.3g cell-updates in %.1f s; oracle "
                      "REF32 = reference float arithmetic and unfused pass structure, gcc -O3, %s; host %s" % (
                          len(idx), steps, total, secs, how, cpu_model())}


def main_domains(S, args):
    """64 uneven first-order 2D domains (side lengths log-uniform in [256, 2048], 1-4
    circles each), 1000 steps, FP32-compensated, one launch per k steps over the
    (domain, tile) work list; the FP64 batch of the same domains judges accuracy."""
    form = S.PREC_NAMES[args.form]
    pbs, prims = S.domain_batch(DOMAIN_SEED, NDOMAINS, precision=form)
    steps = [DOMAIN_STEPS] * NDOMAINS
    K = args.steps if args.steps > 0 else DOMAIN_STEPS
    launch = tuple(int(x) for x in args.launch.split(",")) if args.launch else None
    b, wall, avg_ms, launches, ev_ms = run_batch(S, pbs, prims, steps, launch, K, args.warmup, args.progress, 256)
    upd = updates_done(pbs, steps, K)
    lau = list(b.get_launch())
    B = bytes_per_update(2, form)
    achieved = B * upd / (ev_ms * 1e-3) / 1e9 if ev_ms > 0 else 0.0
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    acc = None
    if not args.no_accuracy:
        ref = S.Batch([S.problem(2, 1, p.nx, p.ny, precision=S.F64, lx=p.lx, ly=p.ly) for p in pbs], prims,
                      [K] * NDOMAINS)
        t0 = time.perf_counter()
        ref.run()
        dt64 = time.perf_counter() - t0
        errs = [b.compare(k, ref)["l1_rho"] / (pbs[k].lx * pbs[k].ly) for k in range(NDOMAINS)]
        acc = {"judge": "FP64 batch of the same domains (its launch shape %s, %.3e cell-updates/s)" % (
                   list(ref.get_launch()), upd / dt64),
               "l1_rho_per_unit_area_max": max(errs), "l1_rho_per_unit_area_median": sorted(errs)[NDOMAINS // 2],
               "nonfinite_domains": int((b.check_state() > 0).sum() + (ref.check_state() > 0).sum())}
        ref.close()
    b.close()
    tpu = traffic_per_update(lau, args.form, "k_step2d_tb_%s_domains64_hbm_traffic.md" % args.form)
    tr = tpu * upd / launches if tpu is not None and launches > 0 else None
    names = {S.C32: "exact TwoSum face update", S.C32K: "Kahan-compensated update"}
    result = {
        "metric": METRIC, "value": upd / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched 2D first-order SHLL domains, FP32 fluxes + (hi,lo) FP32 state, " + names[form],
                   "domains": NDOMAINS, "seed": DOMAIN_SEED, "side_range": [256, 2048],
                   "cells": int(sum(p.nx * p.ny for p in pbs)), "cell_updates_timed": int(upd),
                   "launches": launches, "parallelism": "1 GPU",
                   "launch_shape": {"steps_per_launch": lau[1], "waves_per_block": lau[2], "graph_steps": lau[3]}},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": tr,
            "kernel": "shll::k_step2d_tb<%d, %d, %d> (FP32-compensated, first order, %d steps per launch)" % (
                form, lau[1], lau[2], lau[1]) if lau[1] > 1 else "shll::k_step2d<%d, 1, %d, 0>" % (form, lau[2]),
            "algorithmic_bytes_per_cell_update": B, "avg_launch_ms": avg_ms,
            "measured_copy_GBps": copy_ceiling(),
            "note": "achieved = 64 B x cell-updates / HIP-event time of the launches (k steps per launch read "
                    "and write the state once). traffic = calibrated rocprofv3 FETCH_SIZE+WRITE_SIZE per "
                    "cell-update (profiles/k_step2d_tb_%s_domains64_hbm_traffic.md) x cell-updates per launch"
                    % args.form},
        "accuracy": acc,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_2d(pbs, prims)
        result["cpu_baseline_omp"] = cpu_baseline_2d(pbs, prims, nthreads=omp_threads())
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="time steps to time (0: the whole job)")
    ap.add_argument("--warmup", type=int, default=128,
                    help="untimed warm-up time steps (then reset); the default is one 128-step graph chunk")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 judge of the batch")
    ap.add_argument("--progress", type=int, default=0,
                    help="advance the timed run in pieces of this many steps and log each (profiling runs)")
    ap.add_argument("--launch", default="", help="C,k,waves,graph (default: autotune)")
    ap.add_argument("--calib-copy", type=int, default=0, help="bytes of a device copy to run first (profiling)")
    ap.add_argument("--autotune", action="store_true", help="re-run the autotuner instead of its recorded choices")
    ap.add_argument("--form", default="c32k", choices=["c32k", "c32"],
                    help="compensated update of the timed batch: Kahan (c32k) or exact TwoSum (c32)")
    ap.add_argument("--workload", default="tubes", choices=["tubes", "domains"],
                    help="tubes: the 4096-tube second-order batch (headline); domains: 64 uneven 2D "
                         "first-order domains, 1000 steps")
    args = ap.parse_args()
    if not args.autotune:
        os.environ.setdefault("SHLL_TUNE_CACHE", TUNE_CACHE)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import shll_amd as S

    if args.calib_copy > 0:
        calibration_copy(args.calib_copy)
    if args.workload == "domains":
        main_domains(S, args)
        return
    form = S.PREC_NAMES[args.form]
    other = S.C32 if form == S.C32K else S.C32K
    names = {S.C32: "FP32 fluxes + (hi,lo) FP32 state, exact TwoSum face update",
             S.C32K: "FP32 fluxes + (hi,lo) FP32 state, Kahan-compensated update"}
    t0 = time.perf_counter()
    pbs, prims, steps = S.tube_batch(SEED, NTUBES, precision=form)
    log("generated %d tubes in %.2fs, %d cells, max steps %d" % (len(pbs), time.perf_counter() - t0,
                                                                 sum(p.nx for p in pbs), max(steps)))
    K = args.steps if args.steps > 0 else max(steps)
    launch = tuple(int(x) for x in args.launch.split(",")) if args.launch else None
    b, wall, avg_ms, launches, ev_ms = run_batch(S, pbs, prims, steps, launch, K, args.warmup, args.progress)
    upd = updates_done(pbs, steps, K)
    value = upd / wall
    B = bytes_per_update(1, form)
    achieved = B * upd / (ev_ms * 1e-3) / 1e9 if ev_ms > 0 else 0.0
    lau = list(b.get_launch())
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    judge = None if args.no_accuracy else make_judge(S, pbs, steps, prims, K)
    errs = tube_errors(S, b, pbs, judge)
    b.close()
    alt = f64b = None
    if not args.no_extra:
        # the same batch with the other compensated update, its own tuned shape
        pbo = [S.problem(1, 2, p.nx, precision=other) for p in pbs]
        bo, wo, ao, lo_, eo = run_batch(S, pbo, prims, steps, None, K, args.warmup, args.progress)
        alt = {"form": names[other], "precision": "c32" if other == S.C32 else "c32k",
               "cell_updates_per_s": upd / wo, "avg_launch_ms": ao, "launches": lo_,
               "launch_shape": list(bo.get_launch()), "accuracy": tube_errors(S, bo, pbo, judge)}
        log("other form: %.3e cell-updates/s" % (upd / wo))
        bo.close()
        # the same batch in FP64 (state and fluxes): the precision that meets any error bar
        # against the FP64 oracle (it is bitwise that oracle), at its own tuned shape
        pbf = [S.problem(1, 2, p.nx, precision=S.F64) for p in pbs]
        bf, wf, af, lf, _ = run_batch(S, pbf, prims, steps, None, K, args.warmup, args.progress)
        f64b = {"form": "FP64 state and fluxes, reference association (bitwise the FP64 CPU oracle)",
                "cell_updates_per_s": upd / wf, "avg_launch_ms": af, "launches": lf,
                "launch_shape": list(bf.get_launch())}
        log("FP64 batch: %.3e cell-updates/s" % (upd / wf))
        bf.close()
    if judge is not None:
        judge[0].close()
    tpu = traffic_per_update(lau, args.form)
    tr = tpu * upd / launches if tpu is not None and launches > 0 else None
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched 1D second-order SHLL shock tubes, " + names[form],
            "tubes": NTUBES, "seed": SEED, "cells": int(sum(p.nx for p in pbs)),
            "N_range": [256, 65536], "t_end": 0.2, "max_steps": int(max(steps)),
            "cell_updates_timed": int(upd), "launches": launches, "parallelism": "1 GPU",
            "launch_shape": {"cells_per_lane": lau[0], "steps_per_launch": lau[1], "waves_per_block": lau[2],
                             "graph_steps": lau[3]},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": tr,
            "kernel": "shll::k_step1d<%d, 2, %d, %d, 0> (FP32-compensated %s, second order minmod, %d cells/lane, %d waves/block)" % (
                form, lau[0], lau[2], "Kahan" if form == S.C32K else "TwoSum", lau[0], lau[2]),
            "algorithmic_bytes_per_cell_update": B,
            "avg_launch_ms": avg_ms,
            "measured_copy_GBps": copy_ceiling(),
            "note": "achieved = 48 B x cell-updates / HIP-event time of the launches; temporal blocking "
                    "(k steps per launch from registers) lets it exceed the HBM peak. traffic = HBM bytes per "
                    "launch of the timed run: calibrated rocprofv3 FETCH_SIZE+WRITE_SIZE per cell-update "
                    "(profiles/k_step1d_%s_o2_tubes4096_hbm_traffic.md) x cell-updates per launch" % args.form,
        },
        "accuracy": errs,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(S, pbs, prims, steps)
        result["cpu_baseline_omp"] = cpu_baseline(S, pbs, prims, steps, nthreads=omp_threads())
    else:
        result["cpu_baseline"] = None
    if not args.no_extra:
        result["extra"] = {"c3_other_update": alt, "c3_f64_batch": f64b}
        result["extra"].update(extra_single(S, cpu=not args.no_cpu))
        if not args.no_cpu:
            result["extra"]["cpu_reference_solvers"] = reference_cpu()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
