"""GPU: the benchmark configurations at their full sizes, checked through properties that
hold at any size (the oracle would take hours on them).

c1  65536-cell Sod tube, second order, 104858 steps (t = 0.2): no wave reaches the walls, so
    the wall fluxes of mass and energy are exactly zero and those of momentum exactly the
    pressures 10 and 1 of the two end states -- the momentum sum grows by exactly 9 per unit
    time -- and the density converges to Toro's exact solution (L1 falls with N).
c2  4096^2 circle, first order, reflective walls, 2000 steps: mass and energy conserved, the
    symmetric data keep zero net momentum and their mirror symmetry, and the compensated form
    stays within L1(rho) 1e-6 of FP64.
c3  the 4096-tube second-order batch (seed 20261202), F64F, every tube to t = 0.2: tubes run
    in the batch equal the same tubes run alone, bitwise (the largest, the smallest and four
    more).
c4  the 64 uneven first-order 2D domains, 1000 steps: per-domain conservation, the compensated
    batch within L1(rho) 1e-6 of FP64 per unit area, batch runs equal single runs bitwise.

Conservation bounds are relative to the conserved total, about 10x above what was measured:
FP64 forms ~1e-13 (c1, 104858 steps), C32K ~1e-11 (c1) to ~2e-10 (a c4 domain, 1000 steps: the
Kahan update rounds each cell's net face term once per step); the figures are printed.
References: base_shll.c:92-127 and 2nd_order_base_shll.c:212-256,436-458 (the wall rules and
updates whose fluxes these ledgers sum), base_shll_2d.c:150-236."""
import os
import sys

import numpy as np
import pytest

import shll_amd as S

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_invariants import exact_riemann_rho  # noqa: E402

pytestmark = pytest.mark.gpu


def banner(*a):
    print("[gpu-case]", *a, flush=True)


def sod_run(prec, n, order=2):
    pb = S.problem(1, order, n, precision=prec)
    with S.Solver(pb, S.ic_tube(n)) as s:
        s.autotune(2000)  # the recorded choice when the tuning cache has one
        c0 = s.conserved_sums()
        steps = s.run_to_time(0.2)
        c1 = s.conserved_sums()
        rho = s.primitives()[0]
        bad = s.check_state()
    return pb, steps, c0, c1, rho, bad


@pytest.mark.parametrize("prec,tol,mtol", [(S.F64, 1e-12, 1e-14), (S.F64F, 1e-12, 1e-14), (S.C32K, 1e-9, 1e-11)])
def test_c1_full_size_ledgers_and_convergence(prec, tol, mtol):
    n = 65536
    banner("c1 full size", prec)
    pb, steps, c0, c1, rho, bad = sod_run(prec, n)
    assert steps == 104858 and bad == 0
    dt = 0.125 / n  # DT = 0.125*DX (base_shll.c:84)
    t = steps * dt
    dm = abs(c1[0] - c0[0]) / c0[0]
    de = abs(c1[2] - c0[2]) / c0[2]
    dp = abs(c1[1] - 9.0 * t) / (9.0 * t)
    print("c1 prec %d: mass %.2e, energy %.2e, momentum ledger %.2e (relative)" % (prec, dm, de, dp))
    assert c0[1] == 0.0
    assert dm <= tol and de <= tol
    assert dp <= mtol
    # convergence to the exact solution: at 4x the cells the L1 error at least halves
    x = (np.arange(n) + 0.5) / n
    ex, _, _ = exact_riemann_rho(x, t, 10.0, 10.0, 1.0, 1.0)
    e_full = np.abs(rho - ex).sum() / n
    m = n // 4
    _, steps_m, _, _, rho_m, _ = sod_run(prec, m)
    xm = (np.arange(m) + 0.5) / m
    exm, _, _ = exact_riemann_rho(xm, steps_m * 0.125 / m, 10.0, 10.0, 1.0, 1.0)
    e_quarter = np.abs(rho_m - exm).sum() / m
    print("c1 prec %d: L1(rho) vs exact %.3e at N=%d, %.3e at N=%d" % (prec, e_full, n, e_quarter, m))
    assert e_full < 0.5 * e_quarter and e_full < 1e-3


@pytest.mark.parametrize("prec,k,tol", [(S.F64, 2, 1e-14), (S.C32K, 4, 1e-10)])
def test_c2_full_size_conservation_and_symmetry(prec, k, tol):
    n = 4096
    banner("c2 full size", prec)
    pb = S.problem(2, 1, n, n, precision=prec)
    with S.Solver(pb, S.ic_circle(n, n)) as s:
        s.set_launch(1, k, 1, 64)
        c0 = s.conserved_sums()
        s.step(2000)
        c1 = s.conserved_sums()
        p = s.primitives()
        assert s.check_state() == 0
    dm = abs(c1[0] - c0[0]) / c0[0]
    de = abs(c1[3] - c0[3]) / c0[3]
    rho = p[0].reshape(n, n)
    mir_x = np.abs(rho - rho[::-1, :]).max() / rho.max()
    mir_y = np.abs(rho - rho[:, ::-1]).max() / rho.max()
    print("c2 prec %d: mass %.2e, energy %.2e, momentum (%.2e, %.2e), mirror x %.2e, y %.2e"
          % (prec, dm, de, c1[1], c1[2], mir_x, mir_y))
    assert dm <= tol and de <= tol
    # zero net momentum: the walls' pressure impulses cancel pairwise on symmetric data
    assert abs(c1[1]) <= 1e-15 and abs(c1[2]) <= 1e-15
    assert mir_x <= 1e-13 and mir_y <= 1e-13


def test_c2_full_size_compensated_accuracy():
    n = 4096
    banner("c2 full size accuracy")
    sol = {}
    for prec, k in ((S.F64, 2), (S.C32K, 4)):
        s = S.Solver(S.problem(2, 1, n, n, precision=prec), S.ic_circle(n, n))
        s.set_launch(1, k, 1, 64)
        s.step(2000)
        sol[prec] = s
    e = sol[S.C32K].compare(sol[S.F64])
    for s in sol.values():
        s.close()
    print("c2 C32K vs F64: L1(rho) %.3e, Linf(rho) %.3e" % (e["l1_rho"], e["linf_rho"]))
    assert e["l1_rho"] <= 1e-6


def test_c3_full_batch_equals_single_runs():
    banner("c3 full batch")
    pbs, prims, steps = S.tube_batch(20261202, 4096, precision=S.F64F)
    b = S.Batch(pbs, prims, steps)
    b.autotune(2048)
    b.run()
    sizes = np.array([p.nx for p in pbs])
    rng = np.random.default_rng(7)
    pick = [int(sizes.argmax()), int(sizes.argmin())] + [int(k) for k in rng.choice(len(pbs), 4, replace=False)]
    for k in pick:
        with S.Solver(pbs[k], prims[k]) as s:
            s.autotune(2000)
            s.step(steps[k])
            np.testing.assert_array_equal(b.conserved(k), s.conserved(), err_msg="tube %d (N=%d)" % (k, sizes[k]))
    print("c3: tubes %s (N=%s) equal their single runs" % (pick, [int(sizes[k]) for k in pick]))
    b.close()


@pytest.mark.parametrize("prec,k,tol", [(S.F64, 2, 1e-13), (S.C32K, 4, 2e-9)])
def test_c4_full_batch_conservation_and_single_runs(prec, k, tol):
    banner("c4 full batch", prec)
    pbs, prims = S.domain_batch(4242, 64, precision=prec)
    b = S.Batch(pbs, prims, [1000] * 64)
    b.set_launch(1, k, 1, 64)
    c0 = b.conserved_sums()
    b.run()
    c1 = b.conserved_sums()
    assert (b.check_state() == 0).all()
    dm = np.abs(c1[:, 0] - c0[:, 0]) / c0[:, 0]
    de = np.abs(c1[:, 3] - c0[:, 3]) / c0[:, 3]
    print("c4 prec %d: worst mass %.2e, worst energy %.2e" % (prec, dm.max(), de.max()))
    assert dm.max() <= tol and de.max() <= tol
    areas = np.array([p.nx * p.ny for p in pbs])
    for d in (int(areas.argmin()), int(areas.argmax())):
        with S.Solver(pbs[d], prims[d]) as s:
            s.set_launch(1, k, 4, 0)
            s.step(1000)
            np.testing.assert_array_equal(b.conserved(d), s.conserved(), err_msg="domain %d" % d)
    b.close()


def test_c4_full_batch_compensated_accuracy():
    banner("c4 full batch accuracy")
    out = {}
    for prec, k in ((S.F64, 2), (S.C32K, 4)):
        pbs, prims = S.domain_batch(4242, 64, precision=prec)
        b = S.Batch(pbs, prims, [1000] * 64)
        b.set_launch(1, k, 1, 64)
        b.run()
        out[prec] = b
    pbs = out[S.F64].pbs
    errs = []
    for d in range(64):
        e = out[S.C32K].compare(d, out[S.F64])
        errs.append(e["l1_rho"] / (pbs[d].lx * pbs[d].ly))  # L1 is summed over the domain's area
    for b in out.values():
        b.close()
    print("c4 C32K vs F64: L1(rho) per unit area, worst %.3e, median %.3e" % (max(errs), sorted(errs)[32]))
    assert max(errs) <= 1e-6
