#!/usr/bin/env python3
"""Run the library's autotuner (shll_autotune / shll_batch_autotune) on the benchmark
problems and record its choices in a tuning cache (one line per device, kernel and
shape class: "<key>|C k waves graph"; in 2D C is the rows per work item, 1 the default).
The library reads the cache named by $SHLL_TUNE_CACHE; bench.py points it at
tune/gfx950.txt so timed runs use the recorded choice instead of re-tuning.

Usage (on the GPU):  python shll-sve-cfd_amd/tune/make_tune_cache.py OUT.txt [--only PREC ...]
  --only f64f c32k ...: tune only the problems of these precisions and append their lines to
  OUT.txt (an existing cache keeps its other lines)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    args = sys.argv[1:]
    only = None
    if "--only" in args:
        i = args.index("--only")
        only, args = args[i + 1:], args[:i]
    out = os.path.abspath(args[0] if args else os.path.join(HERE, "gfx950.txt"))
    if only is None and os.path.exists(out):
        os.remove(out)
    os.environ["SHLL_TUNE_CACHE"] = out
    import shll_amd as S

    def wanted(prec):
        return only is None or any(S.PREC_NAMES[o] == prec for o in only)

    # batched second-order tubes (the benchmark batch: both compensated forms and both FP64
    # forms); the trial window (2048 steps, every tube of >= 1280 cells still running) weights
    # the tube sizes roughly as the whole run does
    for prec in (S.C32, S.C32K, S.F64, S.F64F):
        if not wanted(prec):
            continue
        pbs, prims, steps = S.tube_batch(20261202, 4096, precision=prec)
        b = S.Batch(pbs, prims, steps)
        b.autotune(2048)
        print("batch prec%d o2:" % prec, b.get_launch(), flush=True)
        b.close()
    # single instances: 1D first order 1024 cells, 1D second order 65536 cells
    # (a 65536-cell trial of 2000 steps lasts ~3 ms per candidate, short enough for run-to-run
    # noise to pick among shapes within 5% of each other, and 8000-step trials ranked graph
    # lengths unlike the 104858-step run; 30000 steps)
    for pb, trial in ((S.problem(1, 1, 1024, precision=S.F64), 400),
                      (S.problem(1, 2, 65536, precision=S.F64), 30000),
                      (S.problem(1, 2, 65536, precision=S.F64F), 30000),
                      (S.problem(1, 2, 65536, precision=S.C32), 30000),
                      (S.problem(1, 2, 65536, precision=S.C32K), 30000)):
        if not wanted(pb.precision):
            continue
        with S.Solver(pb, S.ic_tube(pb.nx)) as s:
            s.autotune(trial)
            print("single dim1 order%d prec%d N=%d:" % (pb.order, pb.precision, pb.nx), s.get_launch(), flush=True)
    # 2D first order: the 64-domain batch (both compensated forms and FP64) and one 4096^2 domain
    # (every precision); a 256-step trial covers 64 launches of 4 steps
    for prec in (S.C32K, S.C32, S.F64):
        if not wanted(prec):
            continue
        pbs, prims = S.domain_batch(4242, 64, precision=prec)
        b = S.Batch(pbs, prims, [1000] * 64)
        b.autotune(256)
        print("batch dim2 order1 prec%d:" % prec, b.get_launch(), flush=True)
        b.close()
    for prec in (S.F64, S.C32, S.C32K):
        if not wanted(prec):
            continue
        with S.Solver(S.problem(2, 1, 4096, 4096, precision=prec), S.ic_circle(4096, 4096)) as s:
            s.autotune(256)
            print("single dim2 order1 prec%d 4096^2:" % prec, s.get_launch(), flush=True)
    # 2D second order: the reference OpenMP solver's problem (MC limiter, theta 1.25, four
    # contacts, 1024^2; base-omp/2nd_order_base_shll.c), every precision; its default items give
    # about one wavefront per SIMD, so the autotuner also tries shorter ones (1000-step trial)
    for prec in (S.REF32, S.F64, S.C32, S.C32K):
        if not wanted(prec):
            continue
        pb = S.problem(2, 2, 1024, 1024, precision=prec, limiter=S.LIM_MC, theta=1.25)
        with S.Solver(pb, S.ic_four_contacts(1024, 1024)) as s:
            s.autotune(1000)
            print("single dim2 order2 MC prec%d 1024^2:" % prec, s.get_launch(), flush=True)
    print(open(out).read())


if __name__ == "__main__":
    main()
