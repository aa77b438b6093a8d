"""Pins of the oracle's F64F mode (the FP64 face-term form of the fast FP64 kernel) and the
round-off amplification of the second-order scheme that makes FP32 flux arithmetic miss
L1(rho) <= 1e-6 against FP64 on long second-order runs.

F64F has no reference counterpart; it is pinned by invariants written from the equations
and by its distance to the F64 mode, which is FP64 round-off only: a wrong sign, index or
dropped term in any of its formulas shows as a difference of 1e-4 or more."""
import numpy as np
import pytest

import icgen
from test_oracle_invariants import exact_riemann_rho, textbook_flux_1d


@pytest.mark.parametrize("order,bc", [(1, 0), (1, 1), (2, 1), (2, 0)])
def test_f64f_rest_state_is_a_fixed_point(orc, order, bc):
    pb = orc.problem(1, order, 64, bc=bc)
    prim = np.array([[3.25] * 64, [0.0] * 64, [1.5] * 64])
    c0 = orc.stage(pb, orc.F64F, prim)["cons"]
    _, c = orc.run(pb, orc.F64F, prim, 25)
    np.testing.assert_array_equal(c, c0)


def test_f64f_split_flux_against_the_textbook_flux(orc):
    """F+ + F- = F, the sound speed, and F+/- = F/2 +/- aU/2 at rest, from the equations."""
    n = 500
    prim = icgen.random_1d(n, 3, 1)
    pb = orc.problem(1, 1, n)
    st = orc.stage(pb, orc.F64F, prim)
    F, U, a = textbook_flux_1d(st["prim"])
    np.testing.assert_allclose(st["cons"], U, rtol=1e-15)
    np.testing.assert_allclose(st["fp"] + st["fm"], F, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(st["a"], a, rtol=1e-14)
    prim0 = prim.copy()
    prim0[1] = 0.0
    st0 = orc.stage(pb, orc.F64F, prim0)
    F0, U0, a0 = textbook_flux_1d(st0["prim"])
    np.testing.assert_allclose(st0["fp"], 0.5 * F0 + 0.5 * a0 * U0, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(st0["fm"], 0.5 * F0 - 0.5 * a0 * U0, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("order,bc", [(1, 0), (2, 1), (2, 0)])
def test_f64f_mirror_symmetry(orc, order, bc):
    n = 1024
    pb = orc.problem(1, order, n, bc=bc)
    ic = orc.ic_tube(n, 0.5, 3.0, 0.25, 1.25, 1.0, -0.5, 1.0)
    icm = ic[:, ::-1].copy()
    icm[1] *= -1
    p, _ = orc.run(pb, orc.F64F, ic, 800)
    pm, _ = orc.run(pb, orc.F64F, icm, 800)
    np.testing.assert_allclose(pm[0][::-1], p[0], rtol=1e-12)
    np.testing.assert_allclose(-pm[1][::-1], p[1], atol=1e-12 * np.abs(p[1]).max())


@pytest.mark.parametrize("order", [1, 2])
def test_f64f_conservation_and_wall_impulse(orc, order):
    """Reflective walls: mass and energy conserved to FP64 round-off over 3000 steps, and
    one step changes the momentum by the wall impulse -lam*(2 F+_1(last) - 2 F-_1(first))
    (base_shll.c:95-110)."""
    pb = orc.problem(1, order, 2048, bc=0)
    ic = orc.ic_tube(2048)
    c0 = orc.stage(pb, orc.F64F, ic)["cons"].sum(axis=1)
    _, c = orc.run(pb, orc.F64F, ic, 3000)
    c1 = c.sum(axis=1)
    assert abs(c1[0] - c0[0]) <= 1e-13 * c0[0]
    assert abs(c1[2] - c0[2]) <= 1e-13 * c0[2]
    n = 300
    pb = orc.problem(1, order, n, bc=0)
    prim = icgen.random_1d(n, 5, block=10)
    st = orc.stage(pb, orc.F64F, prim)
    lam = orc.constants(pb, orc.F64)["dtdx"]
    _, c = orc.run(pb, orc.F64F, prim, 1)
    d = c.sum(axis=1) - st["cons"].sum(axis=1)
    impulse = -lam * (2.0 * st["fp"][1, -1] - 2.0 * st["fm"][1, 0])
    assert abs(d[1] - impulse) <= 1e-13 * (np.abs(st["cons"][1]).sum() + abs(impulse))
    assert abs(impulse) > 1e-3


@pytest.mark.parametrize("order,lim", [(1, 0), (2, 0), (2, 1)])
def test_f64f_is_fp64_roundoff_away_from_f64(orc, order, lim):
    """Short runs on random data with both wall types: elementwise relative 1e-13 from F64."""
    n = 777
    for bc in (0, 1):
        pb = orc.problem(1, order, n, bc=bc, limiter=lim)
        prim = icgen.random_1d(n, 11 + bc, block=9)
        a, ca = orc.run(pb, orc.F64, prim, 60)
        b, cb = orc.run(pb, orc.F64F, prim, 60)
        np.testing.assert_allclose(cb, ca, rtol=1e-13, atol=1e-13 * np.abs(ca).max())
        assert not np.array_equal(cb, ca)  # a different association, not a copy of F64


def test_f64f_converges_to_the_exact_riemann_solution(orc):
    errs = []
    for n in (256, 512, 1024):
        pb = orc.problem(1, 2, n)
        steps = orc.steps_to_time(pb, 0.2)
        t = steps * orc.constants(pb, orc.F64)["dt"]
        p, _ = orc.run(pb, orc.F64F, orc.ic_tube(n), steps)
        ex, _, _ = exact_riemann_rho((np.arange(n) + 0.5) / n, t, 10.0, 10.0, 1.0, 1.0)
        errs.append(np.abs(p[0] - ex).sum() / n)
    assert all(b < 0.8 * a for a, b in zip(errs, errs[1:])), errs


def test_second_order_amplifies_roundoff_and_only_fp64_fluxes_meet_the_bar(orc):
    """Second-order Sod, 4096 cells, t = 0.2 (6554 steps), outflow walls:
    a 1e-15 relative perturbation of the initial temperature grows to L1(rho) above 1e-12 in
    FP64 (more than 1000 times), while first order damps it (below 10 times).  So flux
    arithmetic with FP32's 6e-8 rounding cannot stay within L1 1e-6 of FP64 on such runs
    (the compensated C32K form misses it), and the FP64 face form (F64F) meets it with
    orders of magnitude to spare."""
    n = 4096
    amp = {}
    for order in (1, 2):
        pb = orc.problem(1, order, n, bc=1)
        steps = orc.steps_to_time(pb, 0.2)
        ic = orc.ic_tube(n)
        icp = ic.copy()
        icp[2] *= 1.0 + 1e-15
        a, _ = orc.run(pb, orc.F64, ic, steps, nthreads=4)
        b, _ = orc.run(pb, orc.F64, icp, steps, nthreads=4)
        amp[order] = np.abs(a[0] - b[0]).sum() / n / 1e-15
        if order == 2:
            ref = a
            c, _ = orc.run(pb, orc.C32K, ic, steps)
            f, _ = orc.run(pb, orc.F64F, ic, steps)
            l1_c32k = np.abs(c[0] - ref[0]).sum() / n
            l1_f64f = np.abs(f[0] - ref[0]).sum() / n
    assert amp[2] > 1e3 and amp[1] < 10, amp
    assert l1_c32k > 1e-6, l1_c32k
    assert l1_f64f < 1e-10, l1_f64f


def test_f64f_tolerance_covers_a_faithful_reciprocal_square_root(orc):
    """The kernels take 1/sqrt(gamma R T) from a hardware seed and two Newton steps: within
    2 ulp of the correctly rounded value the oracle uses (tests/test_gpu_rounding.py), so the
    GPU parity tests hold F64F to a tolerance, not to bits.  Moving the oracle's value by a
    pseudo-random 0-2 ulp on the long cases of those tests moves the result by differences of
    the size the GPU shows (3.4e-14 at most, scaled by each variable's largest magnitude),
    inside their 3e-13 bound and far below the 1e-4 a wrong index, sign or term gives."""
    def scaled_err(a, b):
        scale = np.abs(b).max(axis=1)
        scale[scale == 0] = 1.0
        return float((np.abs(a - b).max(axis=1) / scale).max())

    cases = [(orc.problem(1, 2, 1024), orc.ic_tube(1024), 1639)]
    for bc in (0, 1):
        cases.append((orc.problem(1, 2, 500, bc=bc), icgen.random_1d(500, 82, block=20), 2500))
    errs = []
    for pb, ic, n in cases:
        _, exact = orc.run(pb, orc.F64F, ic, n)
        try:
            orc.f64f_rsq_nudge(2)
            _, nudged = orc.run(pb, orc.F64F, ic, n)
        finally:
            orc.f64f_rsq_nudge(0)
        _, again = orc.run(pb, orc.F64F, ic, n)
        np.testing.assert_array_equal(again, exact)  # the switch is off again
        errs.append(scaled_err(nudged, exact))
    print("F64F with 1/sqrt moved by 0-2 ulp: scaled differences", ["%.2e" % e for e in errs])
    assert all(1e-16 < e <= 3e-13 for e in errs), errs
