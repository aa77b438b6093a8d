#!/usr/bin/env python3
"""Write the golden vectors of the reference solvers (run once, in a container that
has the reference tree at /root/reference; the vectors are committed, the reference
is not).

Each case runs one reference solver, compiled from its own source by oracle/Makefile
(`make -C oracle ref` -> oracle/_ref/*.run; base-c files with gcc, the base-omp MC-limiter
file with clang, cross-checked on base-c/2nd_order_base_shll.c), either through the reference's own main()
or through its stage functions for a chosen number of steps, optionally after
replacing its initial primitive arrays with deterministic data from icgen.py.
The binaries dump the reference's raw float arrays; this script stores
  1D cases       : every array (p, u, a) at every requested step,
  2D cases       : a SHA-256 digest of every array plus values at a fixed cell sample,
  y-uniform cases: additionally one x line (all lines are checked equal here),
into tests/golden/<case>.npz, and a manifest.json with provenance.

Usage:  python tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import icgen  # noqa: E402

NX2 = 256  # the reference's compiled-in 2D size (base_shll_2d.c:21-22, 2nd_order_base_shll.c:29-30)
N1 = 256   # the reference's compiled-in 1D size (base_shll.c:16)
NMC = 1024  # base-omp/2nd_order_base_shll.c:30-31 (MC limiter file)
# compiled-in grid side of each scheme's reference file
SIDE = {"1d_o1": N1, "2d_o1": NX2, "2d_o2": NX2, "2d_o2mc": NMC}

# name, binary, scheme, ic spec (None = the reference's own), steps (None = run main())
CASES = [
    ("o1_1d_main", "base_shll", "1d_o1", None, None),
    ("o1_1d_sod", "base_shll", "1d_o1", None, [0, 1, 10, 100]),
    ("o1_1d_rand", "base_shll", "1d_o1", {"kind": "random_1d", "seed": 11, "block": 1}, [0, 1]),
    ("o1_1d_blocks", "base_shll", "1d_o1", {"kind": "random_1d", "seed": 12, "block": 16}, [0, 1, 10, 50]),
    ("o1_2d_main", "base_shll_2d", "2d_o1", None, None),
    ("o1_2d_implosion", "base_shll_2d", "2d_o1", None, [0, 1, 10]),
    ("o1_2d_circle", "base_shll_2d", "2d_o1", {"kind": "circle"}, [0, 1, 10, 50]),
    ("o1_2d_rand", "base_shll_2d", "2d_o1", {"kind": "random_2d", "seed": 21, "block": 1}, [0, 1]),
    ("o1_2d_blocks", "base_shll_2d", "2d_o1", {"kind": "random_2d", "seed": 22, "block": 16}, [0, 1, 5]),
    ("o2_2d_main", "2nd_order_base_shll", "2d_o2", None, None),
    ("o2_2d_fourshocks", "2nd_order_base_shll", "2d_o2", None, [0, 1, 10, 50]),
    ("o2_2d_rand", "2nd_order_base_shll", "2d_o2", {"kind": "random_2d", "seed": 31, "block": 1}, [0, 1]),
    ("o2_2d_blocks", "2nd_order_base_shll", "2d_o2", {"kind": "random_2d", "seed": 32, "block": 16}, [0, 1, 5]),
    ("o2_2d_sody", "2nd_order_base_shll", "2d_o2", {"kind": "sod_y_uniform"}, [0, 1, 10, 100, 410]),
    ("o2_2d_yuni", "2nd_order_base_shll", "2d_o2", {"kind": "random_2d_yuni", "seed": 33, "block": 16},
     [0, 1, 10, 40]),
    # base-omp/2nd_order_base_shll.c: MC limiter (alpha = 1.25), four contacts, 1024^2
    ("mc_2d_contacts", "omp_2nd_order_base_shll", "2d_o2mc", None, [0, 1, 10, 50]),
    ("mc_2d_rand", "omp_2nd_order_base_shll", "2d_o2mc", {"kind": "random_2d", "seed": 41, "block": 1}, [0, 1]),
    ("mc_2d_blocks", "omp_2nd_order_base_shll", "2d_o2mc", {"kind": "random_2d", "seed": 42, "block": 16},
     [0, 1, 5]),
    ("mc_2d_sody", "omp_2nd_order_base_shll", "2d_o2mc", {"kind": "sod_y_uniform"}, [0, 1, 10, 60]),
    ("mc_2d_yuni", "omp_2nd_order_base_shll", "2d_o2mc", {"kind": "random_2d_yuni", "seed": 43, "block": 16},
     [0, 1, 10, 40]),
]
# the clang recipe of the MC file applied to base-c/2nd_order_base_shll.c must give the gcc
# build's arrays (the MC file cannot be built with gcc, see oracle/Makefile)
CLANG_CHECK = ("clang_2nd_order_base_shll", "2nd_order_base_shll", [0, 1, 10, 50])


def make_ic(spec, n):
    k = spec["kind"]
    if k == "random_1d":
        return icgen.random_1d(n, spec["seed"], spec["block"])
    if k == "random_2d":
        return icgen.random_2d(n, n, spec["seed"], spec["block"])
    if k == "random_2d_yuni":
        return icgen.random_2d(n, n, spec["seed"], spec["block"], y_uniform=True)
    if k == "sod_y_uniform":
        return icgen.sod_2d_y_uniform(n, n)
    if k == "circle":
        return icgen.circle_2d(n, n)
    raise ValueError(k)


def load_dump(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    assert hdr[0] == 0x4C4C4853, path
    nv, n, st = int(hdr[1]), int(hdr[2]), int(hdr[3])
    arr = np.frombuffer(raw[16:], dtype=np.float32)
    p = arr[: nv * n].reshape(nv, n).copy()
    u = arr[nv * n: 2 * nv * n].reshape(nv, n).copy()
    a = arr[2 * nv * n:].copy()
    return p, u, a, st


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def store_snapshot(out, k, scheme, spec, p, u, a):
    if scheme == "1d_o1":
        out["p_%d" % k], out["u_%d" % k], out["a_%d" % k] = p, u, a
        return
    nv = p.shape[0]
    side = SIDE[scheme]
    digs = [digest(p[v]) for v in range(nv)] + [digest(u[v]) for v in range(nv)] + [digest(a)]
    out["dig_%d" % k] = np.array(digs)
    idx = out["idx"]
    out["ps_%d" % k], out["us_%d" % k], out["as_%d" % k] = p[:, idx], u[:, idx], a[idx]
    if spec is not None and spec["kind"] in ("sod_y_uniform", "random_2d_yuni"):
        P = p.reshape(nv, side, side)
        U = u.reshape(nv, side, side)
        A = a.reshape(side, side)
        same = all(np.array_equal(X, X[..., :1].repeat(side, axis=-1)) for X in (P, U, A))
        if not same:
            raise SystemExit("y-uniform case %s: lines differ at step %d" % (spec["kind"], k))
        out["prow_%d" % k], out["urow_%d" % k], out["arow_%d" % k] = P[:, :, 0], U[:, :, 0], A[:, 0]


def clang_crosscheck(refbin):
    """Run the clang build and the gcc build of base-c/2nd_order_base_shll.c on the same
    steps; every dumped array must be bitwise equal."""
    clang_bin, gcc_bin, steps = CLANG_CHECK
    arrays = {}
    for b in (clang_bin, gcc_bin):
        with tempfile.TemporaryDirectory() as td:
            subprocess.run([os.path.join(refbin, b + ".run"), "steps", os.path.join(td, "s"), "-",
                            ",".join(map(str, steps))], cwd=td, check=True, capture_output=True)
            arrays[b] = [load_dump(os.path.join(td, "s_%d.bin" % k)) for k in steps]
    for k, x, y in zip(steps, arrays[clang_bin], arrays[gcc_bin]):
        for xa, ya in zip(x[:3], y[:3]):
            if not np.array_equal(xa, ya):
                raise SystemExit("clang and gcc builds of 2nd_order_base_shll.c differ at step %d" % k)
    return {"binaries": [clang_bin, gcc_bin], "steps": steps, "bitwise_equal": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "REFDIR=" + args.ref])
    refbin = os.path.join(ROOT, "oracle", "_ref")
    manifest = {
        "generator": "tests/golden/make_golden.py",
        "reference_sources": {},
        "build": "oracle/Makefile target 'ref': gcc -O3 -Dmain=reference_main (flags of base-c/makefile:2-4)",
        "cases": {},
    }
    for f in ("base-c/base_shll.c", "base-c/base_shll_2d.c", "base-c/2nd_order_base_shll.c",
              "base-omp/2nd_order_base_shll.c"):
        manifest["reference_sources"][f] = hashlib.sha256(open(os.path.join(args.ref, f), "rb").read()).hexdigest()
    manifest["build_mc"] = ("oracle/Makefile target 'ref': base-omp/2nd_order_base_shll.c compiled with clang -O3 "
                            "-ffp-contract=off -ftrivial-auto-var-init=zero (its IC reads an uninitialised "
                            "local, :47/:136), OpenMP off (stages called outside a parallel region)")
    manifest["clang_crosscheck"] = clang_crosscheck(refbin)
    mpath = os.path.join(HERE, "manifest.json")
    if args.only and os.path.exists(mpath):
        manifest["cases"] = json.load(open(mpath))["cases"]
    for name, binary, scheme, spec, steps in CASES:
        if args.only and name != args.only:
            continue
        exe = os.path.join(refbin, binary + ".run")
        out = {}
        side = SIDE[scheme]
        if scheme != "1d_o1":
            out["idx"] = icgen.sample_indices(side, side)
        with tempfile.TemporaryDirectory() as td:
            if steps is None:
                res = subprocess.run([exe, "main", os.path.join(td, "m.bin")], cwd=td, check=True,
                                     capture_output=True, text=True)
                p, u, a, st = load_dump(os.path.join(td, "m.bin"))
                store_snapshot(out, st, scheme, spec, p, u, a)
                ks = [st]
                out["stdout"] = np.frombuffer(res.stdout.encode(), dtype=np.uint8)
                rd = os.path.join(td, "results.dat")
                if os.path.exists(rd):
                    out["results_dat"] = np.frombuffer(open(rd, "rb").read(), dtype=np.uint8)
            else:
                icpath = "-"
                if spec is not None:
                    ic = make_ic(spec, side).astype(np.float32)
                    icpath = os.path.join(td, "ic.bin")
                    ic.tofile(icpath)
                subprocess.run([exe, "steps", os.path.join(td, "s"), icpath, ",".join(map(str, steps))],
                               cwd=td, check=True, capture_output=True)
                ks = steps
                for k in steps:
                    p, u, a, st = load_dump(os.path.join(td, "s_%d.bin" % k))
                    assert st == k
                    store_snapshot(out, k, scheme, spec, p, u, a)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        manifest["cases"][name] = {"binary": binary, "scheme": scheme, "ic": spec, "steps": ks, "side": side,
                                   "file": name + ".npz", "bytes": os.path.getsize(path)}
        print("%-18s steps=%s  %d bytes" % (name, ks, os.path.getsize(path)), flush=True)
    json.dump(manifest, open(mpath, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
