"""Invariants that pin the oracle's FP64 and FP32-compensated modes independently of the
golden vectors: each would fail under a plausible mistake (a dropped term, a wrong sign in
a wall table, a swapped neighbour, a transposed direction)."""
import math

import numpy as np
import pytest

import icgen

GAMMA, R = 1.4, 1.0
CV = R / (GAMMA - 1.0)


def textbook_flux_1d(prim):
    """Euler x flux from primitives, written from the equations (not the oracle's code)."""
    rho, u, T = prim
    p = rho * R * T
    E = rho * (CV * T + 0.5 * u * u)
    return np.stack([rho * u, rho * u * u + p, u * (E + p)]), np.stack([rho, rho * u, E]), np.sqrt(GAMMA * R * T)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("dim,order,bc", [(1, 1, 0), (1, 1, 1), (1, 2, 1), (1, 2, 0), (2, 1, 0), (2, 2, 1)])
def test_rest_state_is_a_fixed_point(orc, mode, dim, order, bc):
    """A uniform gas at rest must stay exactly at rest with any wall treatment."""
    n = 64 if dim == 1 else 24
    pb = orc.problem(dim, order, n, n if dim == 2 else 1, bc=bc)
    nc = orc.ncells(pb)
    prim = np.array([[3.25] * nc, [0.0] * nc] + ([[0.0] * nc] if dim == 2 else []) + [[1.5] * nc])
    st0 = orc.stage(pb, mode, prim)
    _, c = orc.run(pb, mode, prim, 25)
    np.testing.assert_array_equal(c, st0["cons"])


def test_split_flux_identity_and_rest_closed_form(orc):
    """F+ + F- = F (Z1 - Z3 = 1) against the textbook flux; at M = 0, F+/- = F/2 +/- a U/2."""
    n = 500
    prim = icgen.random_1d(n, 3, 1)
    pb = orc.problem(1, 1, n)
    st = orc.stage(pb, orc.F64, prim)
    F, U, a = textbook_flux_1d(st["prim"])
    np.testing.assert_allclose(st["fp"] + st["fm"], F, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(st["a"], a, rtol=1e-15)
    prim0 = prim.copy()
    prim0[1] = 0.0
    st0 = orc.stage(pb, orc.F64, prim0)
    F0, U0, a0 = textbook_flux_1d(st0["prim"])
    np.testing.assert_allclose(st0["fp"], 0.5 * F0 + 0.5 * a0 * U0, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(st0["fm"], 0.5 * F0 - 0.5 * a0 * U0, rtol=1e-15, atol=1e-15)
    # supersonic to the right (M >= 1): F- vanishes up to rounding (Z3 = (M-1)/2, Z2 = a(1-M^2)/2)
    prim1 = prim.copy()
    prim1[1] = np.sqrt(GAMMA * R * prim1[2])
    st1 = orc.stage(pb, orc.F64, prim1)
    F1, _, _ = textbook_flux_1d(st1["prim"])
    assert np.max(np.abs(st1["fm"]) / np.abs(F1).max(axis=0)) < 1e-13


@pytest.mark.parametrize("mode", [0, 2, 3])
@pytest.mark.parametrize("order,bc", [(1, 0), (2, 1)])
def test_mirror_symmetry(orc, mode, order, bc):
    """Reflecting the data about x = 1/2 reflects the solution (rho even, u odd)."""
    n = 1024
    pb = orc.problem(1, order, n, bc=bc)
    ic = orc.ic_tube(n, 0.5, 3.0, 0.25, 1.25, 1.0, -0.5, 1.0)
    icm = ic[:, ::-1].copy()
    icm[1] *= -1
    p, _ = orc.run(pb, mode, ic, 800)
    pm, _ = orc.run(pb, mode, icm, 800)
    tol = 1e-12 if mode == orc.F64 else 2e-5
    np.testing.assert_allclose(pm[0][::-1], p[0], rtol=tol)
    np.testing.assert_allclose(-pm[1][::-1], p[1], atol=tol * np.abs(p[1]).max())


@pytest.mark.parametrize("mode,tol", [(0, 1e-14), (2, 1e-13), (3, 1e-10)])
def test_conservation_with_reflective_walls(orc, mode, tol):
    """Mass and energy are conserved; F64 to round-off, C32 (telescoping face terms) too;
    C32K's Kahan update to ~1e-12 (the net face term of a cell is rounded once)."""
    pb = orc.problem(1, 1, 2048)
    ic = orc.ic_tube(2048)
    c0 = orc.stage(pb, mode, ic)["cons"].sum(axis=1)
    _, c = orc.run(pb, mode, ic, 3000)
    c1 = c.sum(axis=1)
    assert abs(c1[0] - c0[0]) <= tol * c0[0] * 10
    assert abs(c1[2] - c0[2]) <= tol * c0[2] * 10
    pb2 = orc.problem(2, 1, 64, 48)
    ic2 = icgen.random_2d(64, 48, 8, block=8)
    c0 = orc.stage(pb2, mode, ic2)["cons"].sum(axis=1)
    _, c = orc.run(pb2, mode, ic2, 60)
    c1 = c.sum(axis=1)
    assert abs(c1[0] - c0[0]) <= tol * c0[0] * 10
    assert abs(c1[3] - c0[3]) <= tol * c0[3] * 10


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_momentum_change_equals_wall_impulse(orc, seed):
    """Reflective walls exert pressure: one step changes the momentum sum by exactly the
    wall fluxes, -lambda*(2*F+_1(last cell) - 2*F-_1(first cell)) (base_shll.c:95-110),
    while mass and energy wall fluxes vanish."""
    n = 300
    pb = orc.problem(1, 1, n)
    prim = icgen.random_1d(n, seed, block=10)
    st = orc.stage(pb, orc.F64, prim)
    lam = orc.constants(pb, orc.F64)["dtdx"]
    _, c = orc.run(pb, orc.F64, prim, 1)
    d = c.sum(axis=1) - st["cons"].sum(axis=1)
    impulse = -lam * (2.0 * st["fp"][1, -1] - 2.0 * st["fm"][1, 0])
    scale = np.abs(st["cons"][1]).sum() + abs(impulse)
    assert abs(d[1] - impulse) <= 1e-13 * scale
    assert abs(d[0]) <= 1e-13 * st["cons"][0].sum()
    assert abs(d[2]) <= 1e-13 * st["cons"][2].sum()
    assert abs(impulse) > 1e-3  # the test would notice a zero (wrong-sign) wall flux


@pytest.mark.parametrize("order", [1, 2])
def test_2d_on_uniform_data_equals_1d(orc, order):
    """2D on y-uniform data with v = 0 equals the 1D run along x, and on x-uniform data
    with u = 0 equals the 1D run along y (F64, bitwise)."""
    nx, ny, steps = 96, 12, 40
    bc = 0 if order == 1 else 1
    line = icgen.random_1d(nx, 77, block=8)  # rho, u, T along one axis
    p1, c1 = orc.run(orc.problem(1, order, nx, bc=bc), orc.F64, line, steps)
    # y-uniform: fields depend on i
    prim = np.zeros((4, nx * ny))
    prim[0] = np.repeat(line[0], ny)
    prim[1] = np.repeat(line[1], ny)
    prim[3] = np.repeat(line[2], ny)
    p2, c2 = orc.run(orc.problem(2, order, nx, ny, bc=bc), orc.F64, prim, steps)
    P = p2.reshape(4, nx, ny)
    for j in range(ny):
        np.testing.assert_array_equal(P[0, :, j], p1[0])
        np.testing.assert_array_equal(P[1, :, j], p1[1])
        np.testing.assert_array_equal(P[3, :, j], p1[2])
    # x-uniform: fields depend on j, the velocity is v.  Square cells (lx = ny/nx): the
    # reference's DT follows DX only (base_shll_2d.c:134-136)
    prim = np.zeros((4, ny * nx))
    prim[0] = np.tile(line[0], ny)
    prim[2] = np.tile(line[1], ny)
    prim[3] = np.tile(line[2], ny)
    p3, _ = orc.run(orc.problem(2, order, ny, nx, bc=bc, lx=ny / nx), orc.F64, prim, steps)
    Q = p3.reshape(4, ny, nx)
    for i in range(ny):
        np.testing.assert_array_equal(Q[0, i], p1[0])
        np.testing.assert_array_equal(Q[2, i], p1[1])
        np.testing.assert_array_equal(Q[3, i], p1[2])


def exact_riemann_rho(x, t, rhoL, pL, rhoR, pR, x0=0.5, g=GAMMA):
    """Density of the exact solution of the Riemann problem with uL = uR = 0
    (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4)."""
    aL, aR = math.sqrt(g * pL / rhoL), math.sqrt(g * pR / rhoR)

    def f(p, rho, pk, a):
        if p > pk:
            A, B = 2 / ((g + 1) * rho), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * a / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    p = 0.5 * (pL + pR)
    for _ in range(100):
        h = 1e-7 * p
        fp = f(p, rhoL, pL, aL) + f(p, rhoR, pR, aR)
        d = (f(p + h, rhoL, pL, aL) + f(p + h, rhoR, pR, aR) - fp) / h
        p_new = max(1e-12, p - fp / d)
        if abs(p_new - p) < 1e-14 * p:
            break
        p = p_new
    u = 0.5 * (f(p, rhoR, pR, aR) - f(p, rhoL, pL, aL))
    rhoLs = rhoL * (p / pL) ** (1 / g)
    aLs = aL * (p / pL) ** ((g - 1) / (2 * g))
    rhoRs = rhoR * ((p / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pR + 1))
    S = aR * math.sqrt((g + 1) / (2 * g) * p / pR + (g - 1) / (2 * g))
    out = np.empty_like(x)
    for k, xx in enumerate(x):
        s = (xx - x0) / t
        if s < -aL:
            out[k] = rhoL
        elif s < u - aLs:
            out[k] = rhoL * (2 / (g + 1) + (g - 1) / ((g + 1) * aL) * (-s)) ** (2 / (g - 1))
        elif s < u:
            out[k] = rhoLs
        elif s < S:
            out[k] = rhoRs
        else:
            out[k] = rhoR
    return out, p, u


def test_sod_converges_to_the_exact_riemann_solution(orc):
    """First-order Sod (rho 10:1, T 1:1) converges to Toro's exact solution in L1."""
    rho, pstar, ustar = exact_riemann_rho(np.array([0.5]), 0.2, 10.0, 10.0, 1.0, 1.0)
    assert abs(pstar - 2.84816) < 2e-4 and abs(ustar - 0.97167) < 2e-4
    errs = []
    for n in (256, 512, 1024, 2048):
        pb = orc.problem(1, 1, n)
        steps = orc.steps_to_time(pb, 0.2)
        t = steps * orc.constants(pb, orc.F64)["dt"]
        p, _ = orc.run(pb, orc.F64, orc.ic_tube(n), steps, nthreads=4)
        x = (np.arange(n) + 0.5) / n
        ex, _, _ = exact_riemann_rho(x, t, 10.0, 10.0, 1.0, 1.0)
        errs.append(np.abs(p[0] - ex).sum() / n)
    assert all(b < 0.8 * a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 0.05, errs


def test_minmod_brute_force(orc):
    """minmod(l, r) = 0 if l*r < 0, else the smaller magnitude, ties to r (reference modes);
    the C32 limiter is the median of (l, r, 0)."""
    vals = [-3.0, -1.5, -1.0, -0.0, 0.0, 0.5, 1.0, 1.5, 3.0, 1e-30, -1e-30, 2e-25, -3e-25]
    for l in vals:
        for r in vals:
            lf, rf = float(np.float32(l)), float(np.float32(r))
            want = sorted([lf, rf, 0.0])[1]
            got = orc.minmod(orc.C32, l, r)
            assert got == want, (l, r, got, want)
            if lf * rf > 0:
                assert got == (lf if abs(lf) < abs(rf) else rf)
    # opposite signs whose float product underflows are clipped too
    assert orc.minmod(orc.C32, 1e-30, -1e-30) == 0.0
    for mode in (orc.F64, orc.REF32):
        f = float if mode == orc.F64 else (lambda x: float(np.float32(x)))
        for l in vals:
            for r in vals:
                lf, rf = f(l), f(r)
                prod = lf * rf if mode == orc.F64 else float(np.float32(lf) * np.float32(rf))
                want = 0.0 if prod < 0 else (lf if abs(lf) < abs(rf) else rf)
                got = orc.minmod(mode, l, r)
                assert got == want and (want == 0 or math.copysign(1, got) == math.copysign(1, want)), (mode, l, r)
    # the float product of two tiny slopes of opposite sign underflows to -0: no clipping
    assert orc.minmod(orc.REF32, 1e-30, -1e-30) == float(np.float32(-1e-30))
    assert orc.minmod(orc.F64, 1e-30, -1e-30) == 0.0


@pytest.mark.parametrize("mode", [2, 3])
def test_c32_first_order_accuracy(orc, mode):
    """The compensated schemes stay within L1(rho) <= 1e-6 of FP64 on first-order Sod."""
    pb = orc.problem(1, 1, 1024)
    ic = orc.ic_tube(1024)
    pf, _ = orc.run(pb, orc.F64, ic, 1639)
    pc, _ = orc.run(pb, mode, ic, 1639)
    assert np.abs(pc[0] - pf[0]).sum() / 1024 <= 1e-6


def test_kahan_and_twosum_updates_agree(orc):
    """The two compensated updates differ only in how the same face terms are added:
    their states stay within a few FP32 ulps of each other (first order, 400 steps)."""
    pb = orc.problem(1, 1, 512)
    ic = icgen.random_1d(512, 21, block=16)
    _, c2 = orc.run(pb, orc.C32, ic, 400)
    _, c3 = orc.run(pb, orc.C32K, ic, 400)
    assert np.abs(c3 - c2).max() <= 1e-6 * np.abs(c2).max()
    assert not np.array_equal(c2, c3)  # they are different schemes


def textbook_mc(dl, dr, theta):
    """Generalised MC limiter as a three-way minmod (van Leer 1977; LeVeque, Finite Volume
    Methods for Hyperbolic Problems, 6.12): minmod(theta*dl, (dl+dr)/2, theta*dr)."""
    c = [theta * dl, 0.5 * (dl + dr), theta * dr]
    if all(x > 0 for x in c):
        return min(c)
    if all(x < 0 for x in c):
        return max(c)
    return 0.0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mc_slope_is_the_textbook_limiter(orc, mode):
    """On dyadic data (every operation exact) the MC slope of each mode equals the
    three-way minmod form for theta in {1.25, 1.5, 2}, and reduces to minmod at theta = 1."""
    vals = [-3.0, -1.5, -1.0, -0.25, 0.0, 0.5, 1.0, 1.75, 4.0]
    for theta in (1.0, 1.25, 1.5, 2.0):
        for ql in vals:
            for qc in vals:
                for qr in vals:
                    dl, dr = qc - ql, qr - qc
                    got = orc.slope(mode, orc.LIM_MC, theta, ql, qc, qr)
                    assert got == textbook_mc(dl, dr, theta), (mode, theta, ql, qc, qr, got)
                    if theta == 1.0:
                        assert got == orc.slope(mode, orc.LIM_MINMOD, theta, ql, qc, qr)
    # an MC slope exceeds the minmod slope where the data are smooth but uneven
    assert orc.slope(mode, orc.LIM_MC, 1.25, 0.0, 1.0, 3.0) == 1.25
    assert orc.slope(mode, orc.LIM_MINMOD, 1.25, 0.0, 1.0, 3.0) == 1.0


@pytest.mark.parametrize("mode", [2, 3])
def test_c32_mc_accuracy(orc, mode):
    """With the MC limiter the compensated schemes stay within L1(rho) <= 1e-6 of FP64."""
    pb = orc.problem(1, 2, 1024, limiter=orc.LIM_MC)
    ic = orc.ic_tube(1024)
    pf, _ = orc.run(pb, orc.F64, ic, 1639)
    pc, _ = orc.run(pb, mode, ic, 1639)
    assert np.abs(pc[0] - pf[0]).sum() / 1024 <= 1e-6


def test_mc_is_sharper_than_minmod(orc):
    """Second-order Sod at t = 0.2: the MC limiter's L1 error against the exact solution is
    below minmod's (it clips less at smooth extrema), and both are below first order's."""
    n = 1024
    errs = {}
    for name, order, lim in (("o1", 1, 0), ("minmod", 2, 0), ("mc", 2, 1)):
        pb = orc.problem(1, order, n, bc=1, limiter=lim)
        steps = orc.steps_to_time(pb, 0.2)
        t = steps * orc.constants(pb, orc.F64)["dt"]
        p, _ = orc.run(pb, orc.F64, orc.ic_tube(n), steps)
        ex, _, _ = exact_riemann_rho((np.arange(n) + 0.5) / n, t, 10.0, 10.0, 1.0, 1.0)
        errs[name] = np.abs(p[0] - ex).sum() / n
    assert errs["mc"] < errs["minmod"] < errs["o1"], errs


@pytest.mark.parametrize("mode", [0, 2])
def test_mc_mirror_symmetry_and_conservation(orc, mode):
    """MC limiter, reflective walls: the mirrored problem gives the mirrored solution and
    mass and energy are conserved."""
    n = 1024
    pb = orc.problem(1, 2, n, bc=0, limiter=orc.LIM_MC)
    ic = orc.ic_tube(n, 0.5, 3.0, 0.25, 1.25, 1.0, -0.5, 1.0)
    icm = ic[:, ::-1].copy()
    icm[1] *= -1
    c0 = orc.stage(pb, mode, ic)["cons"].sum(axis=1)
    p, c = orc.run(pb, mode, ic, 800)
    pm, _ = orc.run(pb, mode, icm, 800)
    tol = 1e-10 if mode == orc.F64 else 2e-5
    np.testing.assert_allclose(pm[0][::-1], p[0], rtol=tol)
    np.testing.assert_allclose(-pm[1][::-1], p[1], atol=tol * np.abs(p[1]).max())
    c1 = c.sum(axis=1)
    assert abs(c1[0] - c0[0]) <= 1e-12 * c0[0] and abs(c1[2] - c0[2]) <= 1e-12 * c0[2]
