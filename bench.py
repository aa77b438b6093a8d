#!/usr/bin/env python3
"""Benchmark of the SHLL time-step loop on one MI355X.

Headline workload: 4096 independent second-order shock tubes (seed 20261202), N
log-uniform in [256, 65536] cells, rho_L/rho_R in [2, 20], T_L/T_R in [0.5, 2], diaphragm
in [0.3, 0.7], u = 0, outflow walls, each run to t = 0.2 (its own step count from the
reference's float time loop), one launch per k steps over an (instance, tile) work list,
HIP graphs.  A "step" is one time step of every unfinished tube; --steps K times exactly
K of them (the default, 0, times the whole job).  value = cell-updates done / wall time of
the timed region, inputs resident in HBM.

Timed form (default --form f64f): the fastest form that meets the accuracy target
L1(rho) <= 1e-6 against the FP64 kernel -- FP64 state and fluxes in the compensated
kernels' face-term arrangement (SHLL_F64F).  The FP32-compensated forms (--form c32k, c32;
1.25 and 0.95 times its rate) miss that target on second-order tubes of more than ~1000 cells:
the scheme amplifies round-off about a thousandfold, so FP32 flux rounding grows to
L1 ~1e-5.  Every form's full run, its accuracy and its rate are in the line (accuracy,
extra).

accuracy: the bitwise FP64 kernel runs the same batch first (after the timed batch is
uploaded, before its warm-up steps and timed window, so the window starts on a GPU at its
working clocks, as any window of a long job does; a 20-step window on a GPU idle for 50 ms
or more runs ~17% slower while the clocks ramp up); the
timed batch is then run to the end outside the timed region (whole_job: its rate over the
window plus the rest) and every tube is compared with the FP64 run.  A second FP64 run from
initial temperatures perturbed by one part in 1e15 tells which tubes FP64 itself cannot
reproduce to 1e-6: the reference's fixed DT = 0.125*DX lets hot, strongly expanding tubes
exceed the scheme's stability limit, and there round-off grows without bound; those
tubes, and the ones that go non-finite in FP64, are counted and reported apart.

roofline: the timed kernel reads and writes HBM once per k steps, so its bound is VALU
instruction issue.  frac is the measured HBM traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE,
calibrated by a copy of known size; profiles/) over the event time and the 8 TB/s peak;
valu.frac is the kernel's VALU instructions per launch (rocprofv3 SQ_INSTS_VALU) over the
event time and the issue ceiling of its own instruction mix (per-class rates measured in
profiles/valu_issue_ceiling.md);
algorithmic_GBps is 48 B per cell-update (read and write the state once) over the event
time, which temporal blocking lets exceed the peak.

cpu_baseline: the oracle's reference-precision port of the reference loop (unfused
passes, gcc -O3) on one host core, on a stratified sample of the same tubes (the
reference has no 1D second-order solver).  cpu_baseline_omp: the same with one OpenMP
worksharing loop per pass on every host core.  cpu_reference: the reference's own solvers
compiled from their sources (oracle/_ref), timed over their stage loop only: the base-omp
second-order solver with OpenMP on 1 core and on every core, and the base-c solvers at the
configurations' sizes.

--workload domains times the 64 uneven 2D first-order domains instead (same JSON line,
2D roofline), for profiling the 2D kernel; --workload omp2d times the reference OpenMP
solver's own problem (base-omp/2nd_order_base_shll.c: 2D second order, MC limiter, four
contacts, 1024^2, t = 0.3), by default in that file's own arithmetic (REF32), beside that file
timed on the host; --workload sod65536 times one 65536-cell second-order Sod tube to t = 0.2
(a single instance: latency-bound), --workload circle4096 one 4096^2 first-order 2D domain.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form f64f|c32k|c32|f64|ref32]
                       [--no-cpu] [--no-extra] [--no-accuracy] [--workload tubes|domains|omp2d|sod65536|circle4096]
                       [--launch C,k,waves,graph]
"""
import argparse
import json
import os
import platform
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
# the CPUs this process may use, before any OpenMP runtime binds its threads (the oracle's
# OpenMP runs bind the calling thread to the first place, and child processes inherit that)
try:
    ALLOWED_CPUS = sorted(os.sched_getaffinity(0))
except AttributeError:
    ALLOWED_CPUS = list(range(os.cpu_count() or 1))
sys.path.insert(0, os.path.join(ROOT, "shll-sve-cfd_amd"))
# launch shapes chosen by the library's autotuner on MI355X (tune/make_tune_cache.py);
# --autotune re-tunes live instead
TUNE_CACHE = os.path.join(ROOT, "shll-sve-cfd_amd", "tune", "gfx950.txt")
PROFILES = os.path.join(ROOT, "profiles")

METRIC = "cell-updates/s + % HBM roofline, FP32-comp vs FP64, 1D/2D SHLL, 1×MI355X"
PEAK_HBM_GBS = 8000.0
SEED = 20261202
NTUBES = 4096
L1_BAR = 1e-6
PERTURB = 1e-15


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# form -> (library precision name, description, algorithmic bytes per 1D cell-update, dtype)
FORMS = {
    "f64f": ("f64f", "FP64 state and fluxes, compensated kernels' face-term arrangement (SHLL_F64F)", 48, "f64"),
    "c32k": ("c32k", "FP32 fluxes + (hi,lo) FP32 state, Kahan-compensated update (SHLL_C32K)", 48, "f32"),
    "c32": ("c32", "FP32 fluxes + (hi,lo) FP32 state, exact TwoSum face update (SHLL_C32)", 48, "f32"),
    "f64": ("f64", "FP64 state and fluxes, reference association (SHLL_F64, bitwise the FP64 oracle)", 48, "f64"),
    "ref32": ("ref32", "float with the reference's double promotions (SHLL_REF32, bit for bit the compiled reference)",
              24, "f32"),
}


def bytes_per_update(dim, prec):
    """Algorithmic bytes per cell-update: read and write the conserved state once
    (2 * nvar * bytes per value; FP32-compensated stores hi and lo)."""
    import shll_amd as S
    nv = 3 if dim == 1 else 4
    s = {S.F64: 8, S.REF32: 4, S.C32: 8, S.C32K: 8, S.F64F: 8}[prec]
    return 2 * nv * s


def rows_per_item(lau):
    """2D: owned rows per work item of a launch shape (its first value; 0 or 1: the
    default for its steps per launch)."""
    return lau[0] if lau[0] > 1 else {1: 16, 2: 32}.get(lau[1], 64)


def profile_figure(summary, key, shape):
    """A per-cell-update figure ("<key>: <value> (launch shape C,k,w)") from a committed
    rocprofv3 summary, when it was taken at the same launch shape; else None."""
    path = os.path.join(PROFILES, summary)
    if not os.path.exists(path):
        return None
    m = re.search(r"%s:\s*([0-9.eE+-]+)\s*\(launch shape (\d+),(\d+),(\d+)\)" % re.escape(key), open(path).read())
    if not m or tuple(int(x) for x in m.groups()[1:]) != tuple(shape[:3]):
        return None
    return float(m.group(1))


def simd_count():
    """SIMDs of the GPU the issue ceilings were measured on (profiles/valu_issue_ceiling.md)."""
    path = os.path.join(PROFILES, "valu_issue_ceiling.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"simds:\s*([0-9]+)", open(path).read())
    return int(m.group(1)) if m else None


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def host_cores():
    """The CPUs this process may use (the box's share of the host)."""
    return len(ALLOWED_CPUS)


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
    rc = 0
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


def make_batch(S, pbs, prims, steps, launch, trial=2048):
    """Build a batch (its inputs uploaded to HBM) and set its launch shape: the given one,
    else the autotuner's (its recorded choice when the tuning cache has one)."""
    b = S.Batch(pbs, prims, steps)
    if launch:
        b.set_launch(*launch)
    else:
        b.autotune(trial)
    return b


def run_batch(S, pbs, prims, steps, launch, K, W, progress=0, trial=2048, b=None):
    """Build a batch (unless given one), warm up W steps (then reset) and time exactly K
    steps.  Returns (batch, wall s, average launch ms, launches, event ms)."""
    if b is None:
        b = make_batch(S, pbs, prims, steps, launch, trial)
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


# ---------------------------------------------------------------------------
# accuracy of the tube batch against the bitwise FP64 kernel
# ---------------------------------------------------------------------------
class TubeReferee:
    """The FP64 kernel's run of every tube (bitwise the FP64 oracle), and a second FP64
    run from initial temperatures perturbed by PERTURB (relative): a tube whose two FP64
    runs differ by more than the bar cannot be pinned to 1e-6 by FP64 itself."""

    def __init__(self, S, pbs, prims, steps):
        import numpy as np
        pf = [S.problem(1, 2, p.nx, precision=S.F64) for p in pbs]
        self.ref = S.Batch(pf, prims, steps)
        self.ref.autotune(2048)  # the recorded choice when the tuning cache has one
        t0 = time.perf_counter()
        self.ref.run()
        self.seconds = time.perf_counter() - t0
        self.launch = list(self.ref.get_launch())
        self.bad = self.ref.check_state()
        pert = []
        for p in prims:
            q = p.copy()
            q[2] = q[2] * (1.0 + PERTURB)
            pert.append(q)
        self.alt = S.Batch(pf, pert, steps)
        self.alt.set_launch(*self.launch)
        self.alt.run()
        self.bad2 = self.alt.check_state()
        self.n = len(pbs)
        self.reproducible = None

    def finish(self):
        """The per-tube comparison of the two FP64 runs (thousands of small kernels, so the
        caller runs it after any timed window, not between the runs and the window)."""
        import numpy as np
        if self.reproducible is not None:
            return
        self.self_l1 = np.full(self.n, np.inf)
        for k in range(self.n):
            if self.bad[k] == 0 and self.bad2[k] == 0:
                self.self_l1[k] = self.alt.compare(k, self.ref)["l1_rho"]
        self.alt.close()
        # FP64 pins the tube's answer to a tenth of the bar or better
        self.reproducible = self.self_l1 <= 0.1 * L1_BAR

    def close(self):
        self.finish()
        self.ref.close()


def tube_accuracy(S, b, pbs, referee):
    """L1 / Linf of rho of every tube of batch b (run to its end) against the FP64 referee."""
    import numpy as np
    referee.finish()
    bad = b.check_state()
    n = len(pbs)
    l1 = np.full(n, np.nan)
    linf = np.full(n, np.nan)
    for k in range(n):
        if referee.bad[k] == 0 and bad[k] == 0:
            e = b.compare(k, referee.ref)
            l1[k], linf[k] = e["l1_rho"], e["linf_rho"]
    size = np.array([p.nx for p in pbs])
    ok = np.isfinite(l1)
    rep = ok & referee.reproducible
    bands = {}
    for lo, hi in ((256, 1024), (1024, 4096), (4096, 16384), (16384, 65537)):
        m = rep & (size >= lo) & (size < hi)
        if m.any():
            bands["N_%d_%d" % (lo, hi - 1 if hi == 65537 else hi)] = {
                "tubes": int(m.sum()), "l1_rho_median": float(np.median(l1[m])), "l1_rho_max": float(l1[m].max()),
                "frac_le_1e-6": float((l1[m] <= L1_BAR).mean())}
    fails = rep & (l1 > L1_BAR)
    irr = ok & ~referee.reproducible & np.isfinite(referee.self_l1)
    spread = (l1[irr] / np.maximum(referee.self_l1[irr], 1e-300)) if irr.any() else None
    return {
        "referee": "FP64 kernel (bitwise the FP64 oracle), every tube to t = 0.2",
        "tubes": n,
        "tubes_nonfinite_fp64": int((referee.bad > 0).sum()),
        "tubes_nonfinite_this_form_only": int(((bad > 0) & (referee.bad == 0)).sum()),
        "tubes_fp64_irreproducible": int((~referee.reproducible & (referee.bad == 0)).sum()),
        "irreproducible_rule": "FP64 from T*(1+1e-15) differs from FP64 by L1(rho) > 1e-7 (a tenth of the bar)",
        "reproducible_tubes": int(rep.sum()),
        "l1_rho_median": float(np.median(l1[rep])) if rep.any() else None,
        "l1_rho_max": float(l1[rep].max()) if rep.any() else None,
        "linf_rho_max": float(linf[rep].max()) if rep.any() else None,
        "frac_l1_rho_le_1e-6": float((l1[rep] <= L1_BAR).mean()) if rep.any() else None,
        "tubes_over_1e-6": int(fails.sum()),
        "all_finite_tubes": {"count": int(ok.sum()), "l1_rho_median": float(np.median(l1[ok])) if ok.any() else None,
                             "frac_l1_rho_le_1e-6": float((l1[ok] <= L1_BAR).mean()) if ok.any() else None},
        "irreproducible_tubes_l1_over_fp64_self_spread_median": float(np.median(spread)) if spread is not None else None,
        "by_size_reproducible": bands,
    }


# ---------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------
def restore_affinity():
    """The calling thread (or, as a preexec_fn, a child process) back on every CPU this
    process was given at start."""
    if hasattr(os, "sched_setaffinity"):
        os.sched_setaffinity(0, ALLOWED_CPUS)


def bind_threads():
    """OpenMP threads bound close together, one per CPU this process may use
    (OMP_PROC_BIND=close, OMP_PLACES = an explicit list of those CPUs), unless the caller
    set them: for this process (the oracle's OpenMP runs, set before the oracle library
    starts its runtime) and the reference runs.  The list is explicit because OMP_PLACES=cores
    inside a restricted CPU set put every thread of the LLVM OpenMP runtime on one place."""
    cpus = ALLOWED_CPUS
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", ",".join("{%d}" % c for c in cpus))
    return "threads bound: OMP_PROC_BIND=%s over %d CPUs of the affinity mask" % (os.environ["OMP_PROC_BIND"], len(cpus))


def cpu_baseline(S, pbs, prims, steps, nthreads=1):
    """Oracle REF32 (the reference's arithmetic and pass structure) on every 256th tube of
    the batch for up to 2000 steps each: serially (the base-c loop), or with one OpenMP
    worksharing loop per pass on nthreads threads (the base-omp pattern)."""
    bound = bind_threads()
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
    restore_affinity()  # the OpenMP runtime bound this thread to its first place
    how = "1 thread" if nthreads == 1 else "%d OpenMP threads (%s), one worksharing loop per pass" % (nthreads, bound)
    return {"value": total / secs, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "sample": "%d tubes (every 256th of the batch, N=%d..%d), min(steps,2000) steps each, "
                      "%.3g cell-updates in %.1f s; oracle REF32 = reference float arithmetic and "
                      "unfused pass structure, gcc -O3, %s, step loop only; host %s" % (
                          len(idx), min(pbs[k].nx for k in idx), max(pbs[k].nx for k in idx), total, secs,
                          how, cpu_model())}


def cpu_rate(pb, prim, steps):
    """cell-updates/s of the oracle's REF32 port (reference arithmetic and pass structure,
    gcc -O3, one core) on one problem for a given number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    op = oracle.problem(pb.dim, pb.order, pb.nx, pb.ny, bc=pb.bc, lx=pb.lx, ly=pb.ly)
    _, _, el = oracle.run(op, oracle.REF32, prim, steps, nthreads=1, timed=True)
    return {"value": pb.nx * pb.ny * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the same problem, %.1f s" % (steps, el)}


def reference_cpu(cores):
    """The reference solvers themselves, compiled from their own sources into oracle/_ref by
    oracle/Makefile (when that build exists), run by ref_harness.c's "time" mode: their
    initial data and stage sequence, wall time of the step loop only."""
    runs = [
        ("omp_on_2nd_order_base_shll", "base-omp/2nd_order_base_shll.c (2D second order, MC limiter, four "
         "contacts, 1024^2), OpenMP on, gcc-equivalent -O3 (clang)", [(1, 8), (cores, 30)]),
        ("base_shll_n1024", "base-c/base_shll.c at N = 1024 (configuration 1D Sod, first order), gcc -O3",
         [(1, 1639)]),
        ("base_shll_n65536", "base-c/base_shll.c at N = 65536 (first order; the reference has no 1D "
         "second-order solver), gcc -O3", [(1, 3000)]),
        ("base_shll_2d_n4096", "base-c/base_shll_2d.c at 4096^2 (the 2D configuration's size, the file's "
         "implosion data), gcc -O3", [(1, 3)]),
    ]
    out = {}
    bound = bind_threads()
    for exe, what, plan in runs:
        path = os.path.join(ROOT, "oracle", "_ref", exe + ".run")
        if not os.path.exists(path):
            continue
        for nth, steps in plan:
            # threaded runs share the host with other work: the median of three
            secs = []
            for _ in range(1 if nth == 1 else 3):
                with tempfile.TemporaryDirectory() as td:
                    r = subprocess.run([path, "time", str(steps), str(nth)], cwd=td, capture_output=True,
                                       text=True, timeout=600, preexec_fn=restore_affinity)
                m = re.search(r"cells (\d+) steps (\d+) threads (\d+) seconds ([0-9.eE+-]+)", r.stdout)
                if r.returncode != 0 or not m:
                    log("reference timing %s failed: %s" % (exe, r.stderr.strip()[-200:]))
                    break
                cells, st, th = int(m.group(1)), int(m.group(2)), int(m.group(3))
                secs.append(float(m.group(4)))
            if not secs:
                continue
            sec = sorted(secs)[len(secs) // 2]
            out["%s_%dt" % (exe, th)] = {
                "value": cells * st / sec, "unit": "cell-updates/s", "cores": th, "kind": "reference",
                "sample": "%s: %d cells x %d steps in %.2f s on %d thread%s (step loop only%s); host %s" % (
                    what, cells, st, sec, th, "" if th == 1 else "s (%s)" % bound,
                    "" if len(secs) == 1 else ", median of %d runs" % len(secs), cpu_model())}
    return out


# ---------------------------------------------------------------------------
# single-instance configurations
# ---------------------------------------------------------------------------
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
        shape = s.get_launch()
    with S.Solver(pb, ic) as s2:
        s2.set_launch(*shape)
        t0 = time.perf_counter()
        n = s2.run_to_time(0.2)
        dt = time.perf_counter() - t0
    out["c0_1d_o1_f64_N1024"] = {"steps": n, "seconds": dt, "cell_updates_per_s": 1024 * n / dt,
                                 "launch": list(shape)}
    if cpu:
        out["c0_1d_o1_f64_N1024"]["cpu_baseline"] = cpu_rate(pb, ic, n)
    # 1D second order, 65536 cells, t = 0.2: FP64, the fast FP64 form and both compensated forms
    n1 = 65536
    ic = S.ic_tube(n1)
    sol = {}
    for name, prec in (("f64", S.F64), ("f64f", S.F64F), ("c32", S.C32), ("c32k", S.C32K)):
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
    for name in ("f64f", "c32", "c32k"):
        e = sol[name].compare(sol["f64"])
        out["c1_1d_o2_%s_N65536" % name]["error_vs_f64"] = e
        out["c1_1d_o2_%s_N65536" % name]["meets_l1_1e-6"] = bool(e["l1_rho"] <= L1_BAR)
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
        s.autotune(256)  # the recorded choice when the tuning cache has one
        shape = s.get_launch()
        s.step(20)
        s.close()
        s = S.Solver(S.problem(2, 1, n2, n2, precision=prec), ic)
        s.set_launch(*shape)
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
        e = sol[name].compare(sol["f64"])
        out["c2_2d_o1_%s_4096sq" % name]["error_vs_f64"] = e
        out["c2_2d_o1_%s_4096sq" % name]["meets_l1_1e-6"] = bool(e["l1_rho"] <= L1_BAR)
    if cpu:
        out["c2_2d_o1_f64_4096sq"]["cpu_baseline"] = cpu_rate(S.problem(2, 1, n2, n2), ic, 2)
    for s in sol.values():
        s.close()
    # 64 uneven 2D domains, first order, 1000 steps, both FP32-compensated forms; the FP64
    # batch of the same domains is the accuracy referee
    pbs, prims = S.domain_batch(DOMAIN_SEED, NDOMAINS, precision=S.C32K)
    cells = sum(p.nx * p.ny for p in pbs)
    ref = S.Batch([S.problem(2, 1, p.nx, p.ny, precision=S.F64, lx=p.lx, ly=p.ly) for p in pbs], prims,
                  [DOMAIN_STEPS] * NDOMAINS)
    ref.autotune(256)
    t1 = time.perf_counter()
    ref.run()
    dt64 = time.perf_counter() - t1
    bad64 = int((ref.check_state() > 0).sum())
    for name, prec in (("c32k", S.C32K), ("c32", S.C32)):
        pbp = [S.problem(2, 1, p.nx, p.ny, precision=prec, lx=p.lx, ly=p.ly) for p in pbs]
        b = S.Batch(pbp, prims, [DOMAIN_STEPS] * NDOMAINS)
        b.autotune(256)  # the recorded choice when the tuning cache has one
        b.advance(10)
        b.reset()
        t0 = time.perf_counter()
        b.run()
        dt = time.perf_counter() - t0
        errs = [b.compare(k, ref)["l1_rho"] / (pbs[k].lx * pbs[k].ly) for k in range(NDOMAINS)]
        launch = list(b.get_launch())
        nonfinite = bad64 + int((b.check_state() > 0).sum())
        b.close()
        out["c4_2d_o1_%s_64domains" % name] = {"steps": DOMAIN_STEPS, "cells": cells, "seconds": dt,
                                               "launch": launch,
                                               "cell_updates_per_s": cells * DOMAIN_STEPS / dt,
                                               "f64_batch_cell_updates_per_s": cells * DOMAIN_STEPS / dt64,
                                               "l1_rho_vs_f64_per_unit_area_max": max(errs),
                                               "l1_rho_vs_f64_per_unit_area_median": sorted(errs)[NDOMAINS // 2],
                                               "meets_l1_1e-6": bool(max(errs) <= L1_BAR),
                                               "nonfinite_domains": nonfinite}
    ref.close()
    # the reference's OpenMP solver's own problem (base-omp/2nd_order_base_shll.c: 2D second
    # order, MC limiter theta 1.25, four contacts, 1024^2, outflow walls, t = 0.3): REF32 is its
    # arithmetic bit for bit, so its rate compares with the reference timed on the host
    n3 = 1024
    ic = S.ic_four_contacts(n3, n3)
    sol = {}
    for name, prec in (("ref32", S.REF32), ("f64", S.F64), ("c32k", S.C32K)):
        pb = S.problem(2, 2, n3, n3, precision=prec, limiter=S.LIM_MC, theta=1.25)
        s = S.Solver(pb, ic)
        s.autotune(256)
        shape = s.get_launch()
        s.close()
        s = S.Solver(pb, ic)
        s.set_launch(*shape)
        s.step(8)
        s.close()
        s = S.Solver(pb, ic)
        s.set_launch(*shape)
        t0 = time.perf_counter()
        n = s.run_to_time(0.3)
        dt = time.perf_counter() - t0
        sol[name] = s
        out["omp_problem_2d_o2_mc_%s_1024sq" % name] = {"steps": n, "seconds": dt,
                                                        "cell_updates_per_s": n3 * n3 * n / dt,
                                                        "launch": list(shape)}
    for name in ("ref32", "c32k"):
        e = sol[name].compare(sol["f64"])
        out["omp_problem_2d_o2_mc_%s_1024sq" % name]["error_vs_f64"] = e
        out["omp_problem_2d_o2_mc_%s_1024sq" % name]["meets_l1_1e-6"] = bool(e["l1_rho"] <= L1_BAR)
    for s in sol.values():
        s.close()
    return out


# ---------------------------------------------------------------------------
# 2D domains workload
# ---------------------------------------------------------------------------
DOMAIN_SEED = 4242
NDOMAINS = 64
DOMAIN_STEPS = 1000


def cpu_baseline_2d(pbs, prims, nthreads=1, steps=40):
    """Oracle REF32 first-order 2D (the reference's arithmetic and pass structure) on every
    8th domain of the batch for a few steps each."""
    bound = bind_threads()
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
    restore_affinity()
    how = "1 thread" if nthreads == 1 else "%d OpenMP threads (%s), one worksharing loop per pass" % (nthreads, bound)
    return {"value": total / secs, "unit": "cell-updates/s", "cores": nthreads, "kind": "port",
            "sample": "%d domains (every 8th of the batch), %d steps each, %.3g cell-updates in %.1f s; oracle "
                      "REF32 = reference float arithmetic and unfused pass structure, gcc -O3, %s; host %s" % (
                          len(idx), steps, total, secs, how, cpu_model())}


def roofline(upd, launches, ev_ms, avg_ms, B, shape, traffic_summary, valu_summary, kernel):
    """Roofline object of the timed step kernel: measured HBM traffic against the 8 TB/s
    peak (frac), the VALU issue fraction, and the algorithmic bytes rate."""
    t = ev_ms * 1e-3
    per_launch = upd / max(1, launches)
    tpu = profile_figure(traffic_summary, "hbm_bytes_per_cell_update", shape)
    vpu = profile_figure(valu_summary, "valu_insts_per_cell_update", shape)
    ceil = profile_figure(valu_summary, "valu_mix_ceiling_per_simd_per_ns", shape)
    simds = simd_count()
    traffic = tpu * per_launch if tpu is not None else None
    achieved = tpu * upd / t / 1e9 if (tpu is not None and t > 0) else None
    valu = None
    if vpu is not None and ceil is not None and simds and t > 0:
        rate = vpu * upd / t  # wavefront VALU instructions per second
        peak = ceil * 1e9 * simds
        valu = {"achieved": rate, "peak": peak, "unit": "wavefront VALU instructions/s", "frac": rate / peak,
                "insts_per_cell_update": vpu,
                "ceiling": "the kernel's interior-loop instruction mix weighted by the measured per-class issue "
                           "rates (%.4f per SIMD per ns, %d SIMDs; profiles/%s, profiles/valu_issue_ceiling.md)"
                           % (ceil, simds, valu_summary)}
    k = shape[1]
    return {
        "bound": "valu" if k > 1 else "hbm",
        "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": achieved / PEAK_HBM_GBS if achieved is not None else None,
        "traffic": traffic,
        "valu": valu,
        "algorithmic_GBps": B * upd / t / 1e9 if t > 0 else None,
        "algorithmic_bytes_per_cell_update": B,
        "kernel": kernel,
        "avg_launch_ms": avg_ms,
        "measured_copy_GBps": copy_ceiling(),
        "note": "k = %d steps per launch read and write HBM once: the kernel is bound by VALU issue. achieved = "
                "measured HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE per cell-update at this launch shape, "
                "profiles/%s) x cell-updates / HIP-event time; traffic = those bytes per launch; "
                "algorithmic_GBps = %d B per cell-update (state read and written once) / event time; valu = "
                "SQ_INSTS_VALU per cell-update (profiles/%s) x cell-updates / event time against the measured "
                "issue ceiling" % (k, traffic_summary, B, valu_summary),
    }


def main_domains(S, args):
    """64 uneven first-order 2D domains (side lengths log-uniform in [256, 2048], 1-4
    circles each), 1000 steps, FP32-compensated, one launch per k steps over the
    (domain, tile) work list; the FP64 batch of the same domains is the accuracy referee."""
    form = args.form if args.form in ("c32k", "c32", "f64") else "c32k"
    prec = S.PREC_NAMES[form]
    pbs, prims = S.domain_batch(DOMAIN_SEED, NDOMAINS, precision=prec)
    steps = [DOMAIN_STEPS] * NDOMAINS
    K = args.steps if args.steps > 0 else DOMAIN_STEPS
    launch = tuple(int(x) for x in args.launch.split(",")) if args.launch else None
    b, wall, avg_ms, launches, ev_ms = run_batch(S, pbs, prims, steps, launch, K, args.warmup, args.progress, 256)
    upd = updates_done(pbs, steps, K)
    lau = list(b.get_launch())
    B = bytes_per_update(2, prec)
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    acc = None
    if not args.no_accuracy:
        if K < DOMAIN_STEPS:
            b.run()
        ref = S.Batch([S.problem(2, 1, p.nx, p.ny, precision=S.F64, lx=p.lx, ly=p.ly) for p in pbs], prims, steps)
        ref.autotune(256)
        t0 = time.perf_counter()
        ref.run()
        dt64 = time.perf_counter() - t0
        errs = [b.compare(k, ref)["l1_rho"] / (pbs[k].lx * pbs[k].ly) for k in range(NDOMAINS)]
        acc = {"referee": "FP64 batch of the same domains, 1000 steps (its launch shape %s, %.3e cell-updates/s)" % (
                   list(ref.get_launch()), updates_done(pbs, steps, DOMAIN_STEPS) / dt64),
               "l1_rho_per_unit_area_max": max(errs), "l1_rho_per_unit_area_median": sorted(errs)[NDOMAINS // 2],
               "meets_l1_1e-6": bool(max(errs) <= L1_BAR),
               "nonfinite_domains": int((b.check_state() > 0).sum() + (ref.check_state() > 0).sum())}
        ref.close()
    b.close()
    kern = ("shll::k_step2d_tb<%d, %d, %d> (first order, %d steps per launch)" % (prec, lau[1], lau[2], lau[1])
            if lau[1] > 1 else "shll::k_step2d<%d, 1, %d, 0>" % (prec, lau[2]))
    result = {
        "metric": METRIC, "value": upd / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": FORMS[form][3], "data": "synthetic",
        "config": {"workload": "batched 2D first-order SHLL domains, " + FORMS[form][1],
                   "domains": NDOMAINS, "seed": DOMAIN_SEED, "side_range": [256, 2048],
                   "cells": int(sum(p.nx * p.ny for p in pbs)), "cell_updates_timed": int(upd),
                   "launches": launches, "parallelism": "1 GPU",
                   "launch_shape": {"rows_per_item": rows_per_item(lau), "steps_per_launch": lau[1],
                                    "waves_per_block": lau[2], "graph_steps": lau[3]}},
        "roofline": roofline(upd, launches, ev_ms, avg_ms, B, lau,
                             "k_step2d_tb_%s_domains64_hbm_traffic.md" % form,
                             "k_step2d_tb_%s_domains64_sq_counters.md" % form, kern),
        "accuracy": acc,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_2d(pbs, prims)
        result["cpu_baseline_omp"] = cpu_baseline_2d(pbs, prims, nthreads=host_cores())
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


def main_sod65536(S, args):
    """One 65536-cell Sod tube, second order, to t = 0.2 (104858 steps; the reference's 1D
    Sod data, base_shll.c:54-60, with the outflow walls of 2nd_order_base_shll.c): a single
    instance of 1e5 dependent steps, bound by the latency of one wavefront per SIMD."""
    form = args.form if args.form in ("f64f", "f64", "c32k", "c32", "ref32") else "f64f"
    prec = S.PREC_NAMES[form]
    n = 65536
    pb = S.problem(1, 2, n, precision=prec)
    ic = S.ic_tube(n)
    total = S.steps_to_time(pb, 0.2)
    K = min(args.steps, total) if args.steps > 0 else total
    s = S.Solver(pb, ic)
    if args.launch:
        s.set_launch(*(int(x) for x in args.launch.split(",")))
    else:
        s.autotune(8000)  # the recorded choice when the tuning cache has one
    lau = list(s.get_launch())
    if args.warmup > 0:
        raw0 = s.state_raw()
        s.step(args.warmup)
        s.set_state_raw(raw0)
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(K)
    wall = time.perf_counter() - t0
    avg_ms, launches, ev_ms = s.last_timing()
    upd = n * K
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    acc = None
    if not args.no_accuracy and form != "f64":
        with S.Solver(S.problem(1, 2, n, precision=S.F64), ic) as r:
            r.autotune(8000)
            r.step(K)
            e = s.compare(r)
        acc = {"referee": "FP64 kernel (bitwise the FP64 oracle), the same %d steps" % K, "l1_rho": e["l1_rho"],
               "linf_rho": e["linf_rho"], "meets_l1_1e-6": bool(e["l1_rho"] <= L1_BAR)}
    s.close()
    B = bytes_per_update(1, prec)
    kern = "shll::k_step1d<%d, 2, %d, %d, 0> (%s, %d cells/lane, %d waves/block)" % (
        prec, lau[0], lau[2], FORMS[form][1], lau[0], lau[2])
    tag = "k_step1d_%s_o2_sod65536" % form
    result = {
        "metric": METRIC, "value": upd / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": FORMS[form][3], "data": "synthetic",
        "config": {"workload": "one 65536-cell Sod tube, second order, outflow walls, t = 0.2, " + FORMS[form][1],
                   "nx": n, "steps_to_t_end": int(total), "cell_updates_timed": int(upd), "launches": launches,
                   "parallelism": "1 GPU",
                   "launch_shape": {"cells_per_lane": lau[0], "steps_per_launch": lau[1], "waves_per_block": lau[2],
                                    "graph_steps": lau[3]}},
        "roofline": roofline(upd, launches, ev_ms, avg_ms, B, lau, tag + "_hbm_traffic.md", tag + "_sq_counters.md",
                             kern),
        "accuracy": acc,
        "cpu_baseline": None if args.no_cpu else cpu_rate(S.problem(1, 2, n), ic, 400),
    }
    print(json.dumps(result))


def main_circle4096(S, args):
    """One 4096^2 first-order 2D domain (the circle of configuration c2: rho 10 inside r = 0.25,
    reflective walls, base_shll_2d.c's scheme), 2000 steps."""
    form = args.form if args.form in ("f64", "c32k", "c32", "ref32") else "c32k"
    prec = S.PREC_NAMES[form]
    n = 4096
    pb = S.problem(2, 1, n, n, precision=prec)
    ic = S.ic_circle(n, n)
    total = 2000
    K = min(args.steps, total) if args.steps > 0 else total
    s = S.Solver(pb, ic)
    if args.launch:
        s.set_launch(*(int(x) for x in args.launch.split(",")))
    else:
        s.autotune(256)  # the recorded choice when the tuning cache has one
    lau = list(s.get_launch())
    if args.warmup > 0:
        raw0 = s.state_raw()
        s.step(args.warmup)
        s.set_state_raw(raw0)
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(K)
    wall = time.perf_counter() - t0
    avg_ms, launches, ev_ms = s.last_timing()
    upd = n * n * K
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    acc = None
    if not args.no_accuracy and form != "f64":
        with S.Solver(S.problem(2, 1, n, n, precision=S.F64), ic) as r:
            r.autotune(256)
            r.step(K)
            e = s.compare(r)
        acc = {"referee": "FP64 kernel (bitwise the FP64 oracle), the same %d steps" % K, "l1_rho": e["l1_rho"],
               "linf_rho": e["linf_rho"], "meets_l1_1e-6": bool(e["l1_rho"] <= L1_BAR)}
    s.close()
    B = bytes_per_update(2, prec)
    k = lau[1]
    kern = ("shll::k_step2d_tb<%d, %d, %d> (first order, %d steps per launch)" % (prec, k, lau[2], k)
            if k > 1 else "shll::k_step2d<%d, 1, %d, 0>" % (prec, lau[2]))
    tag = ("k_step2d_tb_%s_circle4096" % form) if k > 1 else ("k_step2d_%s_circle4096" % form)
    result = {
        "metric": METRIC, "value": upd / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": FORMS[form][3], "data": "synthetic",
        "config": {"workload": "one 4096^2 first-order 2D domain (circle, reflective walls), 2000 steps, " +
                               FORMS[form][1], "nx": n, "ny": n, "cell_updates_timed": int(upd), "launches": launches,
                   "parallelism": "1 GPU",
                   "launch_shape": {"rows_per_item": rows_per_item(lau), "steps_per_launch": k,
                                    "waves_per_block": lau[2], "graph_steps": lau[3]}},
        "roofline": roofline(upd, launches, ev_ms, avg_ms, B, lau, tag + "_hbm_traffic.md", tag + "_sq_counters.md",
                             kern),
        "accuracy": acc,
        "cpu_baseline": None if args.no_cpu else cpu_rate(S.problem(2, 1, n, n), ic, 2),
    }
    print(json.dumps(result))


def main_omp2d(S, args, ref_cpu):
    """The reference OpenMP solver's own problem, one instance: base-omp/2nd_order_base_shll.c,
    2D second order, MC limiter theta = 1.25 (:44), its four contacts (:149-167) on 1024^2
    (:21-22), outflow walls, to t = 0.3 (:43).  REF32 by default: the file's arithmetic bit
    for bit, so the rate compares with that file timed on the host (cpu_reference)."""
    form = args.form if args.form in ("ref32", "f64", "c32k", "c32") else "ref32"
    prec = S.PREC_NAMES[form]
    n = 1024
    pb = S.problem(2, 2, n, n, precision=prec, limiter=S.LIM_MC, theta=1.25)
    ic = S.ic_four_contacts(n, n)
    total = S.steps_to_time(pb, 0.3)
    K = min(args.steps, total) if args.steps > 0 else total
    s = S.Solver(pb, ic)
    if args.launch:
        s.set_launch(*(int(x) for x in args.launch.split(",")))
    else:
        s.autotune(256)  # the recorded choice when the tuning cache has one
    lau = list(s.get_launch())
    if args.warmup > 0:
        # warm-up steps (graph capture included), then back to the initial state
        raw0 = s.state_raw()
        s.step(args.warmup)
        s.set_state_raw(raw0)
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(K)
    wall = time.perf_counter() - t0
    avg_ms, launches, ev_ms = s.last_timing()
    upd = n * n * K
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    acc = None
    if not args.no_accuracy and form != "f64":
        with S.Solver(S.problem(2, 2, n, n, precision=S.F64, limiter=S.LIM_MC, theta=1.25), ic) as r:
            r.autotune(256)
            r.step(K)
            e = s.compare(r)
        acc = {"referee": "FP64 kernel (bitwise the FP64 oracle), the same %d steps" % K,
               "l1_rho": e["l1_rho"], "linf_rho": e["linf_rho"], "max_rel_rho": e["max_rel_rho"],
               "meets_l1_1e-6": bool(e["l1_rho"] <= L1_BAR)}
    s.close()
    B = bytes_per_update(2, prec)
    k = lau[1]
    kern = ("shll::k_step2d<%d, 2, %d, 1> (second order, MC limiter, one step per launch)" % (prec, lau[2])
            if k == 1 else "shll::k_step2d_tb2<%d, 2, %d, 1> (second order, MC limiter, 2 steps per launch)"
            % (prec, lau[2]))
    tag = "k_step2d_%s_o2_mc1024" % form if k == 1 else "k_step2d_tb2_%s_o2_mc1024" % form
    result = {
        "metric": METRIC, "value": upd / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": FORMS[form][3], "data": "synthetic",
        "config": {"workload": "the reference OpenMP solver's problem (2D second order, MC limiter theta 1.25, "
                               "four contacts, outflow walls), " + FORMS[form][1],
                   "nx": n, "ny": n, "t_end": 0.3, "steps_to_t_end": int(total), "cell_updates_timed": int(upd),
                   "launches": launches, "parallelism": "1 GPU",
                   "launch_shape": {"rows_per_item": rows_per_item(lau), "steps_per_launch": k,
                                    "waves_per_block": lau[2], "graph_steps": lau[3]}},
        "roofline": roofline(upd, launches, ev_ms, avg_ms, B, lau, tag + "_hbm_traffic.md", tag + "_sq_counters.md",
                             kern),
        "accuracy": acc,
    }
    if k == 1:
        result["roofline"]["bound"] = "hbm"
        result["roofline"]["note"] = ("one step per launch: the state is read and written every step. " +
                                      result["roofline"]["note"].split(": the kernel is bound by VALU issue. ", 1)[-1])
    if ref_cpu:
        result["cpu_reference"] = ref_cpu
        omp = ref_cpu.get("omp_on_2nd_order_base_shll_%dt" % host_cores())
        if omp and form == "ref32":
            result["vs_reference_openmp_all_cores"] = result["value"] / omp["value"]
    result["cpu_baseline"] = (ref_cpu or {}).get("omp_on_2nd_order_base_shll_1t")
    print(json.dumps(result))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="time steps to time (0: the whole job)")
    ap.add_argument("--warmup", type=int, default=128,
                    help="untimed warm-up time steps (then reset); the default is one 128-step graph chunk")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 referee of the batch")
    ap.add_argument("--progress", type=int, default=0,
                    help="advance the timed run in pieces of this many steps and log each (profiling runs)")
    ap.add_argument("--launch", default="",
                    help="C,k,waves,graph: C cells per lane (1D) or rows per work item (2D), k steps per launch, "
                         "waves per block, steps per graph (0: direct launches); default: the recorded autotuner "
                         "choice")
    ap.add_argument("--calib-copy", type=int, default=0, help="bytes of a device copy to run first (profiling)")
    ap.add_argument("--autotune", action="store_true", help="re-run the autotuner instead of its recorded choices")
    ap.add_argument("--form", default="f64f", choices=sorted(FORMS),
                    help="arithmetic of the timed batch: f64f (FP64 face form, meets L1 1e-6; default), c32k / c32 "
                         "(FP32-compensated), f64 (bitwise FP64)")
    ap.add_argument("--workload", default="tubes", choices=["tubes", "domains", "omp2d", "sod65536", "circle4096"],
                    help="tubes: the 4096-tube second-order batch (headline); domains: 64 uneven 2D "
                         "first-order domains, 1000 steps; omp2d: the reference OpenMP solver's own "
                         "problem (2D second order, MC limiter, 1024^2, t = 0.3; --form ref32 by default); "
                         "sod65536: one 65536-cell second-order Sod tube to t = 0.2; circle4096: one 4096^2 "
                         "first-order 2D domain, 2000 steps")
    args = ap.parse_args()
    if not args.autotune:
        os.environ.setdefault("SHLL_TUNE_CACHE", TUNE_CACHE)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the reference's own solvers are timed first, on a quiet host: inside the full run their
    # 16-thread timing came out 1.7x slower than from a fresh process
    ref_cpu = None
    if not args.no_cpu and args.workload in ("tubes", "omp2d"):
        t0 = time.perf_counter()
        ref_cpu = reference_cpu(host_cores())
        log("reference solvers timed on the host in %.1f s" % (time.perf_counter() - t0))
    import shll_amd as S

    if args.calib_copy > 0:
        calibration_copy(args.calib_copy)
    if args.workload == "domains":
        main_domains(S, args)
        return
    if args.workload == "omp2d":
        main_omp2d(S, args, ref_cpu)
        return
    if args.workload == "sod65536":
        main_sod65536(S, args)
        return
    if args.workload == "circle4096":
        main_circle4096(S, args)
        return
    prec = S.PREC_NAMES[args.form]
    t0 = time.perf_counter()
    pbs, prims, steps = S.tube_batch(SEED, NTUBES, precision=prec)
    total_steps = max(steps)
    log("generated %d tubes in %.2fs, %d cells, max steps %d" % (len(pbs), time.perf_counter() - t0,
                                                                 sum(p.nx for p in pbs), total_steps))
    K = min(args.steps, total_steps) if args.steps > 0 else total_steps
    launch = tuple(int(x) for x in args.launch.split(",")) if args.launch else None
    referee = acc = None
    forms_out = {}
    # the timed batch is built and uploaded first, then the FP64 referee (every tube, twice)
    # runs, then the W warm-up steps and the timed window follow with no idle gap: the GPU
    # lowers its clocks within 50 ms of idling and takes tens of ms of load to raise them
    b = make_batch(S, pbs, prims, steps, launch)
    if not args.no_accuracy:
        t1 = time.perf_counter()
        referee = TubeReferee(S, pbs, prims, steps)
        log("FP64 referee and its perturbed twin in %.1f s" % (time.perf_counter() - t1))
    b, wall, avg_ms, launches, ev_ms = run_batch(S, pbs, prims, steps, launch, K, args.warmup, args.progress, b=b)
    upd = updates_done(pbs, steps, K)
    value = upd / wall
    B = bytes_per_update(1, prec)
    lau = list(b.get_launch())
    log("timed %d steps: %.3e cell-updates in %.3f s (events %.1f ms, %d launches)" % (K, upd, wall, ev_ms, launches))
    whole = {"steps": int(K), "seconds": wall, "cell_updates_per_s": value}
    if K < total_steps and referee is not None:
        # the rest of the job, timed apart, for the whole job's rate (window + rest)
        t1 = time.perf_counter()
        b.run()
        rest = time.perf_counter() - t1
        whole = {"steps": int(total_steps), "seconds": wall + rest,
                 "cell_updates_per_s": updates_done(pbs, steps, total_steps) / (wall + rest)}
    if referee is not None:
        acc = tube_accuracy(S, b, pbs, referee)
        acc["timed_steps_then_run_to_end"] = K < total_steps
    b.close()
    if not args.no_extra:
        # every other form of the same batch: full run at its recorded shape, rate and accuracy
        for name in ("f64f", "c32k", "c32", "f64"):
            if name == args.form:
                continue
            if name == "f64" and referee is not None:
                forms_out[name] = {"form": FORMS[name][1], "cell_updates_per_s": updates_done(pbs, steps, total_steps)
                                   / referee.seconds, "launch_shape": referee.launch,
                                   "accuracy": "bitwise the FP64 oracle (the referee)"}
                continue
            pp = [S.problem(1, 2, p.nx, precision=S.PREC_NAMES[name]) for p in pbs]
            bo, wo, ao, lo_, _ = run_batch(S, pp, prims, steps, None, total_steps, args.warmup)
            forms_out[name] = {"form": FORMS[name][1], "cell_updates_per_s": updates_done(pbs, steps, total_steps) / wo,
                               "avg_launch_ms": ao, "launches": lo_, "launch_shape": list(bo.get_launch())}
            if referee is not None and name != "f64":
                a = tube_accuracy(S, bo, pbs, referee)
                forms_out[name]["accuracy"] = a
                forms_out[name]["meets_l1_1e-6_on_reproducible_tubes"] = a["tubes_over_1e-6"] == 0
            log("%s: %.3e cell-updates/s" % (name, forms_out[name]["cell_updates_per_s"]))
            bo.close()
    if referee is not None:
        referee.close()
    kern = "shll::k_step1d<%d, 2, %d, %d, 0> (%s, second order minmod, %d cells/lane, %d waves/block)" % (
        prec, lau[0], lau[2], FORMS[args.form][1], lau[0], lau[2])
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
        "dtype": FORMS[args.form][3],
        "data": "synthetic",
        "config": {
            "workload": "batched 1D second-order SHLL shock tubes, " + FORMS[args.form][1],
            "tubes": NTUBES, "seed": SEED, "cells": int(sum(p.nx for p in pbs)),
            "N_range": [256, 65536], "t_end": 0.2, "max_steps": int(total_steps),
            "cell_updates_timed": int(upd), "launches": launches, "parallelism": "1 GPU",
            "launch_shape": {"cells_per_lane": lau[0], "steps_per_launch": lau[1], "waves_per_block": lau[2],
                             "graph_steps": lau[3]},
        },
        "whole_job": dict(whole, note=("the timed window of K steps, then the rest of the job to t = 0.2 (every "
                                       "tube to its own step count) timed apart; the FP64 accuracy referee runs "
                                       "before both") if whole["steps"] == total_steps and K < total_steps else
                          "the timed window (the rest of the job is not run without the accuracy referee)"
                          if K < total_steps else "the timed window is the whole job"),
        "roofline": roofline(upd, launches, ev_ms, avg_ms, B, lau,
                             "k_step1d_%s_o2_tubes4096_hbm_traffic.md" % args.form,
                             "k_step1d_%s_o2_tubes4096_sq_counters.md" % args.form, kern),
        "accuracy": acc,
    }
    # the same per-cell-update figures over the whole job's wall time (host gaps included)
    r = result["roofline"]
    if r["valu"] is not None:
        result["whole_job"]["valu_frac"] = r["valu"]["insts_per_cell_update"] * whole["cell_updates_per_s"] / \
            r["valu"]["peak"]
    if r["achieved"] is not None:
        result["whole_job"]["hbm_frac"] = r["achieved"] / (upd / (ev_ms * 1e-3)) * whole["cell_updates_per_s"] / \
            PEAK_HBM_GBS
    if not args.no_cpu:
        cores = host_cores()
        result["cpu_reference"] = ref_cpu
        result["cpu_baseline"] = cpu_baseline(S, pbs, prims, steps)
        result["cpu_baseline_omp"] = cpu_baseline(S, pbs, prims, steps, nthreads=cores)
    else:
        result["cpu_baseline"] = None
    if not args.no_extra:
        result["extra"] = {"c3_forms": forms_out}
        result["extra"].update(extra_single(S, cpu=not args.no_cpu))
        omp = (result.get("cpu_reference") or {}).get("omp_on_2nd_order_base_shll_%dt" % host_cores())
        gpu = result["extra"].get("omp_problem_2d_o2_mc_ref32_1024sq")
        if omp and gpu:
            # the same problem in the same arithmetic: the GPU against the reference's OpenMP
            gpu["vs_reference_openmp_all_cores"] = gpu["cell_updates_per_s"] / omp["value"]
    print(json.dumps(result))


if __name__ == "__main__":
    main()
