"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same
inputs.  F64 and REF32 are compared bitwise (the oracle's reference association and
IEEE rounding are reproduced exactly); C32 is compared bitwise against the oracle's
restatement of the compensated scheme and by its distance to F64; F64F, whose kernels
take 1/sqrt(gamma R T) from a hardware seed and two Newton steps (within 2 ulp) where
the oracle rounds correctly, within a stated round-off tolerance (assert_oracle_match)."""
import os

import numpy as np
import pytest

import icgen
import shll_amd as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = {S.F64: 0, S.REF32: 1, S.C32: 2, S.C32K: 3, S.F64F: 4}  # library precision -> oracle mode
COMP = (S.C32, S.C32K)  # the (hi, lo) compensated forms


def banner(*a):
    print("[gpu-case]", *a, flush=True)


def opb(orc, pb):
    return orc.problem(pb.dim, pb.order, pb.nx, pb.ny, bc=pb.bc, lx=pb.lx, ly=pb.ly, limiter=pb.limiter,
                       theta=pb.theta)


def gpu_run(pb, prim, nsteps, launch=None):
    with S.Solver(pb, prim) as s:
        if launch:
            s.set_launch(*launch)
        s.step(nsteps)
        raw = s.state_raw()
        return s.primitives(), s.conserved(), raw


def oracle_run(orc, pb, prim, nsteps):
    o = opb(orc, pb)
    mode = MODES[pb.precision]
    p, c = orc.run(o, mode, prim, nsteps, nthreads=4)
    raw = orc.run_c32_raw(o, prim, nsteps, mode) if pb.precision in COMP else None
    return p, c, raw


# F64F against the oracle: the largest difference of any value, relative to the largest
# magnitude of its variable in the oracle's array.  Tolerances are about 8-10x above the
# largest difference the MI355X measured (every comparison prints its own); a wrong index,
# sign or dropped term in the kernels shows as 1e-4 or more.
F64F_TOL_SHORT = 5e-14  # up to 310 steps (measured 6.0e-15)
F64F_TOL_LONG = 3e-13  # 548 to 2500 steps (measured 3.4e-14)
F64F_TOL_STAGE = 2e-15  # split fluxes and sound speed of one state (measured 2.7e-16)


def scaled_err(got, want):
    g = np.asarray(got, dtype=np.float64)
    w = np.asarray(want, dtype=np.float64).reshape(g.shape)
    if g.ndim == 1:
        g, w = g[None, :], w[None, :]
    g, w = g.reshape(g.shape[0], -1), w.reshape(w.shape[0], -1)
    scale = np.abs(w).max(axis=1)
    scale[scale == 0] = 1.0
    return float((np.abs(g - w).max(axis=1) / scale).max())


def assert_oracle_match(prec, got, want, what, tol=F64F_TOL_SHORT):
    """Bitwise for every form but F64F; F64F within tol of the oracle (scaled_err)."""
    if prec != S.F64F:
        np.testing.assert_array_equal(got, want, err_msg=what)
        return
    e = scaled_err(got, want)
    print("[f64f-vs-oracle] %s: %.2e" % (what, e), flush=True)
    assert e <= tol, "%s: F64F is %.3e from the oracle (scaled), above %.0e" % (what, e, tol)


def assert_same(pb, gpu, ora, what, tol=F64F_TOL_SHORT):
    gp, gc, graw = gpu
    op, oc, oraw = ora
    # compare physical states only: unstable data (CFL > 1) ends in NaN, which says nothing
    assert np.isfinite(oc).all(), what + ": the oracle run left the stable regime"
    if pb.precision in COMP:
        np.testing.assert_array_equal(graw[0], oraw[0], err_msg=what + " hi")
        np.testing.assert_array_equal(graw[1], oraw[1], err_msg=what + " lo")
    assert_oracle_match(pb.precision, gc, oc, what + " conserved", tol)
    assert_oracle_match(pb.precision, gp, op, what + " primitives", tol)


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
@pytest.mark.parametrize("order", [1, 2])
def test_1d_sod_full_run(orc, prec, order):
    """Sod, 1024 cells to t = 0.2 (1639 steps)."""
    pb = S.problem(1, order, 1024, precision=prec)
    n = S.steps_to_time(pb, 0.2)
    assert n == 1639
    ic = S.ic_tube(1024)
    banner("1d sod", prec, order)
    assert_same(pb, gpu_run(pb, ic, n), oracle_run(orc, pb, ic, n), "sod p%d o%d" % (prec, order))


@pytest.mark.parametrize("launch", [(1, 1, 1, 0), (1, 8, 4, 64), (2, 3, 1, 0), (4, 16, 4, 128), (8, 5, 4, 20),
                                    (2, 24, 1, 0), (8, 1, 1, 8), (8, 60, 1, 0), (2, 2, 4, 0)])
@pytest.mark.parametrize("prec", [S.F64, S.C32, S.C32K])
def test_1d_launch_shapes_are_bitwise_identical(orc, launch, prec):
    """Every (cells/lane, steps/launch, waves/block, graph) gives the oracle's bits."""
    n = 3001
    pb = S.problem(1, 2, n, precision=prec)
    ic = icgen.random_1d(n, 5, block=37)
    banner("1d shapes", launch, prec)
    assert_same(pb, gpu_run(pb, ic, 150, launch), oracle_run(orc, pb, ic, 150), "shape %s" % (launch,))


@pytest.mark.parametrize("launch,sizes", [((2, 3, 4, 0), (127, 255, 602, 858, 1082)),
                                          ((4, 7, 4, 0), (255, 767, 1234, 1746)),
                                          ((8, 16, 4, 0), (511, 1535, 2464, 3488))])
@pytest.mark.parametrize("bc", [0, 1])
def test_1d_cooperative_blocks_wall_at_a_wavefront_edge(orc, launch, sizes, bc):
    """Blocks of 4 wavefronts with 2+ cells per lane share edge fluxes through LDS; the
    tiles of these sizes would put ghost cell n on a wavefront's first cell, which the
    tiler moves one cell on.  Both walls, second order, bitwise."""
    for n in sizes:
        pb = S.problem(1, 2, n, precision=S.C32K, bc=bc)
        ic = icgen.random_1d(n, 60 + bc, block=7)
        banner("1d coop", launch, n, bc)
        assert_same(pb, gpu_run(pb, ic, 40, launch), oracle_run(orc, pb, ic, 40), "coop %s n=%d" % (launch, n))


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32K, S.F64F])
def test_1d_long_launches_without_halo(orc, prec):
    """A tube that fits one segment (or one cooperative block) has no halo, so any number of
    steps fits one launch: Sod at 1024 cells to t = 0.2 in one 1639-step launch (8 cells per
    lane, a block of 4 wavefronts), and a 500-cell tube in launches of 1024 steps (one
    wavefront), both walls; bitwise (F64F: bitwise its own run at one step per launch, and
    within round-off of the oracle)."""
    def check(pb, ic, n, launch, what):
        got = gpu_run(pb, ic, n, launch)
        assert_same(pb, got, oracle_run(orc, pb, ic, n), what, F64F_TOL_LONG)
        if prec == S.F64F:
            np.testing.assert_array_equal(got[1], gpu_run(pb, ic, n, (1, 1, 1, 0))[1], err_msg=what)

    pb = S.problem(1, 1, 1024, precision=prec)
    banner("long launch 1024", prec)
    check(pb, S.ic_tube(1024), 1639, (8, 4096, 4, 8192), "1639-step launch")
    for bc in (0, 1):
        pb = S.problem(1, 2, 500, precision=prec, bc=bc)
        banner("long launch 500", prec, bc)
        check(pb, icgen.random_1d(500, 82, block=20), 2500, (8, 1024, 1, 0), "1024-step launches")


def test_1d_impossible_launch_shape_is_rejected():
    pb = S.problem(1, 2, 3001)
    with S.Solver(pb, S.ic_tube(3001)) as s:
        before = s.get_launch()
        with pytest.raises(S.ShllError) as e:
            s.set_launch(2, 64, 1, 0)  # 128-cell segments, 128-cell halo each side
        assert e.value.code == S.EINVAL
        with pytest.raises(S.ShllError) as e:
            s.set_launch(2, 1, 1, 4097)  # a graph of more than 4096 launches
        assert e.value.code == S.EINVAL
        assert s.get_launch() == before
        s.step(3)


@pytest.mark.parametrize("order,bc", [(1, 0), (1, 1), (2, 0), (2, 1)])
@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_1d_walls_random_states(orc, order, bc, prec):
    n = 700
    pb = S.problem(1, order, n, precision=prec, bc=bc)
    ic = icgen.random_1d(n, 40 + order, block=1)
    banner("1d walls", order, bc, prec)
    assert_same(pb, gpu_run(pb, ic, 3, (2, 2, 1, 0)), oracle_run(orc, pb, ic, 3), "walls")
    ic = icgen.random_1d(n, 50 + order, block=25)
    assert_same(pb, gpu_run(pb, ic, 60, (4, 7, 4, 0)), oracle_run(orc, pb, ic, 60), "walls blocky")


def test_1d_reference_golden_and_results_file(tmp_path):
    """REF32 on the GPU reproduces the reference's own run (base_shll.c, N=256, 410 steps)
    bitwise, and its results.dat byte for byte."""
    g = np.load(os.path.join(GOLD, "o1_1d_main.npz"), allow_pickle=False)
    pb = S.problem(1, 1, 256, precision=S.REF32)
    banner("1d golden")
    with S.Solver(pb, S.ic_tube(256)) as s:
        assert s.run_to_time(0.2) == 410
        p = s.primitives().astype(np.float32)
        u = s.state_raw()
        np.testing.assert_array_equal(p, g["p_410"])
        np.testing.assert_array_equal(u, g["u_410"])
        s.write_results(tmp_path / "results.dat")
        assert (tmp_path / "results.dat").read_bytes() == bytes(g["results_dat"])


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
@pytest.mark.parametrize("order", [1, 2])
def test_2d_random_and_circle(orc, prec, order):
    banner("2d", prec, order)
    pb = S.problem(2, order, 70, 150, precision=prec)
    ic = icgen.random_2d(70, 150, 9, block=1)
    assert_same(pb, gpu_run(pb, ic, 2), oracle_run(orc, pb, ic, 2), "2d rough")
    ic = icgen.random_2d(70, 150, 13, block=13)
    assert_same(pb, gpu_run(pb, ic, 25), oracle_run(orc, pb, ic, 25), "2d blocky")
    pb = S.problem(2, order, 96, 96, precision=prec)
    ic = S.ic_circle(96, 96)
    assert_same(pb, gpu_run(pb, ic, 60), oracle_run(orc, pb, ic, 60), "2d circle")


@pytest.mark.parametrize("name,order,steps", [("o1_2d_main", 1, 205), ("o2_2d_fourshocks", 2, 50)])
def test_2d_reference_golden(name, order, steps):
    """REF32 on the GPU reproduces the reference's 2D solvers at 256^2 bitwise."""
    import hashlib
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    pb = S.problem(2, order, 256, 256, precision=S.REF32)
    ic = S.ic_implosion(256, 256) if order == 1 else S.ic_four_shocks(256, 256)
    banner("2d golden", name)
    with S.Solver(pb, ic) as s:
        s.step(steps)
        p = s.primitives().astype(np.float32)
        u = s.state_raw()
    dig = [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in list(p) + list(u)]
    assert dig == list(g["dig_%d" % steps][:8])


def test_2d_launch_shapes(orc):
    pb = S.problem(2, 2, 200, 130, precision=S.F64)
    ic = icgen.random_2d(200, 130, 3, block=9)
    ora = oracle_run(orc, pb, ic, 12)
    for launch in [(1, 1, 1, 0), (1, 1, 4, 4), (1, 1, 4, 0)]:
        banner("2d shapes", launch)
        assert_same(pb, gpu_run(pb, ic, 12, launch), ora, "2d shape %s" % (launch,))


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32K])
def test_2d_rows_per_item(orc, prec):
    """2D work items of any height (cells_per_lane in 2D: the owned rows of an item, 2 and
    more; 0 and 1 the default) give the oracle's bits: second order (MC, both walls) at one
    step per launch, the compensated form's two-step pipeline, first order at 4 steps, and a
    batch; heights above the grid, and bad heights rejected."""
    for order, lim, k, rows in ((2, S.LIM_MC, 1, (2, 3, 5, 7, 300)), (1, S.LIM_MINMOD, 4, (2, 3, 17, 64))):
        for bc in (0, 1):
            pb = S.problem(2, order, 150, 130, precision=prec, bc=bc, limiter=lim)
            ic = icgen.random_2d(150, 130, 71 + bc, block=9)
            ora = oracle_run(orc, pb, ic, 9)
            for r in rows:
                banner("2d rows", prec, order, bc, r)
                assert_same(pb, gpu_run(pb, ic, 9, (r, k, 4 if r % 2 else 1, 0)), ora, "2d rows %d o%d" % (r, order))
    if prec == S.C32K:
        pb = S.problem(2, 2, 203, 70, precision=prec, bc=1)
        ic = icgen.random_2d(203, 70, 77, block=9)
        ora = oracle_run(orc, pb, ic, 7)
        for r in (2, 5, 12):
            assert_same(pb, gpu_run(pb, ic, 7, (r, 2, 1, 4)), ora, "2d o2 k2 rows %d" % r)
    pbs, prims = S.domain_batch(59, 4, smin=20, smax=120, precision=prec, order=2)
    b = S.Batch(pbs, prims, [11, 5, 9, 2])
    b.set_launch(6, 1, 4, 4)
    assert b.get_launch() == (6, 1, 4, 4)
    b.run()
    for k in range(4):
        ref = oracle_run(orc, pbs[k], prims[k], [11, 5, 9, 2][k])
        assert_oracle_match(prec, b.conserved(k), ref[1], "rows batch instance %d" % k)
    b.close()
    with S.Solver(S.problem(2, 2, 40, 30, precision=prec), S.ic_circle(40, 30)) as s:
        for bad in (-1, 4097):
            with pytest.raises(S.ShllError) as e:
                s.set_launch(bad, 1, 1, 0)
            assert e.value.code == S.EINVAL


@pytest.mark.parametrize("prec", [S.F64, S.C32, S.C32K])
@pytest.mark.parametrize("order", [1, 2])
def test_1d_batch_equals_single_runs(orc, prec, order):
    """A batch of uneven tubes: every instance equals its own oracle run, and finished
    instances stay frozen while longer ones continue."""
    pbs, prims, steps = S.tube_batch(777, 12, nmin=64, nmax=3000, precision=prec, order=order)
    steps = [max(1, s // 8) for s in steps]
    b = S.Batch(pbs, prims, steps)
    # cooperative blocks: 2 cells per lane, or the FP64 batch's tuned shape (4 cells, k = 8)
    b.set_launch(*((4, 8, 4, 32) if prec == S.F64 else (2, 8, 4, 32)))
    banner("1d batch", prec, order)
    b.run()
    for k in range(len(pbs)):
        ref = oracle_run(orc, pbs[k], prims[k], steps[k])
        np.testing.assert_array_equal(b.conserved(k), ref[1], err_msg="instance %d" % k)
        np.testing.assert_array_equal(b.primitives(k), ref[0], err_msg="instance %d" % k)
    b.close()


@pytest.mark.parametrize("dim", [1, 2])
def test_batch_piecewise_advance_with_graph_chunks(orc, dim):
    """Advancing in uneven pieces (graph chunks starting on odd and even launches, direct
    launches between them) gives every instance its oracle result."""
    if dim == 1:
        pbs, prims, steps = S.tube_batch(4321, 9, nmin=100, nmax=2500, precision=S.F64)
        steps = [max(1, s // 10) for s in steps]
        launch = (2, 8, 4, 32)   # 4 launches per graph chunk
    else:
        pbs, prims = S.domain_batch(77, 3, smin=30, smax=90, precision=S.F64)
        steps = [41, 17, 29]
        launch = (1, 1, 4, 4)    # 4 launches per graph chunk
    b = S.Batch(pbs, prims, steps)
    b.set_launch(*launch)
    banner("batch pieces", dim)
    for n in (3, 50, 9, 1):
        b.advance(n)
    b.run()
    for k in range(len(pbs)):
        ref = oracle_run(orc, pbs[k], prims[k], steps[k])
        np.testing.assert_array_equal(b.conserved(k), ref[1], err_msg="instance %d" % k)
    b.reset()
    b.run()
    for k in range(len(pbs)):
        ref = oracle_run(orc, pbs[k], prims[k], steps[k])
        np.testing.assert_array_equal(b.conserved(k), ref[1], err_msg="instance %d after reset" % k)
    b.close()


@pytest.mark.parametrize("prec", [S.C32, S.C32K])
def test_2d_batch_equals_single_runs(orc, prec):
    pbs, prims = S.domain_batch(31, 4, smin=40, smax=120, precision=prec)
    steps = [30, 11, 20, 5]
    b = S.Batch(pbs, prims, steps)
    banner("2d batch")
    b.run()
    for k in range(len(pbs)):
        ref = oracle_run(orc, pbs[k], prims[k], steps[k])
        np.testing.assert_array_equal(b.conserved(k), ref[1], err_msg="domain %d" % k)
    b.close()


@pytest.mark.parametrize("prec,tol", [(S.F64, 1e-13), (S.C32, 1e-13), (S.C32K, 1e-9)])
def test_conservation_reflective(prec, tol):
    """Reflective walls: mass and energy are conserved (F64 to round-off; C32's face
    update telescopes, so to ~1e-14; C32K's Kahan update to ~1e-11)."""
    pb = S.problem(1, 1, 4096, precision=prec)
    banner("conservation", prec)
    with S.Solver(pb, S.ic_tube(4096)) as s:
        s0 = s.conserved_sums()
        s.step(3000)
        s1 = s.conserved_sums()
        assert s.check_state() == 0
    assert abs(s1[0] - s0[0]) <= tol * abs(s0[0])
    assert abs(s1[2] - s0[2]) <= tol * abs(s0[2])
    pb = S.problem(2, 1, 128, 128, precision=prec)
    with S.Solver(pb, S.ic_circle(128, 128)) as s:
        s0 = s.conserved_sums()
        s.step(300)
        s1 = s.conserved_sums()
    assert abs(s1[0] - s0[0]) <= tol * abs(s0[0])
    assert abs(s1[3] - s0[3]) <= tol * abs(s0[3])


@pytest.mark.parametrize("prec", [S.C32, S.C32K])
def test_c32_accuracy_against_f64_on_gpu(prec):
    """FP32-compensated first-order Sod (1024 cells) vs the FP64 kernel: L1(rho) <= 1e-6."""
    pb64 = S.problem(1, 1, 1024, precision=S.F64)
    pb32 = S.problem(1, 1, 1024, precision=prec)
    ic = S.ic_tube(1024)
    banner("c32 accuracy")
    with S.Solver(pb64, ic) as a, S.Solver(pb32, ic) as b:
        a.run_to_time(0.2)
        b.run_to_time(0.2)
        e = b.compare(a)
    assert e["l1_rho"] <= 1e-6, e
    assert e["nonfinite"] == 0


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
@pytest.mark.parametrize("dim", [1, 2])
def test_checkpoint_restart_is_bitwise(prec, dim):
    """A run stopped after n steps, saved with state_raw and restarted in a new context with
    set_state_raw, ends bit for bit where the uninterrupted run ends."""
    if dim == 1:
        pb, ic = S.problem(1, 2, 2000, precision=prec), icgen.random_1d(2000, 12, block=40)
    else:
        pb, ic = S.problem(2, 2, 96, 80, precision=prec), S.ic_four_shocks(96, 80)
    banner("restart", prec, dim)
    with S.Solver(pb, ic) as s:
        s.step(37)
        raw = s.state_raw()
        s.step(45)
        want = s.state_raw()
    with S.Solver(pb, S.ic_tube(2000) if dim == 1 else S.ic_implosion(96, 80)) as s:
        s.set_state_raw(raw)
        s.step(45)
        got = s.state_raw()
    for g, w in zip(got if isinstance(got, tuple) else (got,), want if isinstance(want, tuple) else (want,)):
        assert g.tobytes() == w.tobytes()


@pytest.mark.parametrize("dim", [1, 2])
def test_full_results_file_holds_primitives_and_conserved(tmp_path, dim):
    pb = S.problem(dim, 1, 64, 48 if dim == 2 else 1, precision=S.F64)
    ic = S.ic_tube(64) if dim == 1 else S.ic_circle(64, 48)
    with S.Solver(pb, ic) as s:
        s.step(30)
        p, c = s.primitives(), s.conserved()
        s.write_results_full(tmp_path / "full.dat")
        s.write_results(tmp_path / "short.dat")
    r = S.read_results(tmp_path / "full.dat", 64, 48 if dim == 2 else 1)
    short = S.read_results(tmp_path / "short.dat", 64, 48 if dim == 2 else 1)
    names = ["rho", "u", "T"] if dim == 1 else ["rho", "u", "v", "T"]
    for k, n in enumerate(names):
        np.testing.assert_array_equal(r[n].ravel(), short[n].ravel())
        np.testing.assert_allclose(r[n].ravel(), p[k], rtol=1e-6)
        np.testing.assert_allclose(r["u%d" % k].ravel(), c[k], rtol=1e-6, atol=1e-30)


# ---------------------------------------------------------------------------
# MC limiter (base-omp/2nd_order_base_shll.c:44,317-325,385-392)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_1d_mc_limiter(orc, prec, bc):
    n = 1500
    banner("1d mc", prec, bc)
    for theta, launch in ((1.25, (2, 2, 1, 0)), (1.25, (8, 5, 4, 20)), (2.0, (4, 7, 4, 0))):
        pb = S.problem(1, 2, n, precision=prec, bc=bc, limiter=S.LIM_MC, theta=theta)
        ic = icgen.random_1d(n, 60 + bc, block=1)
        assert_same(pb, gpu_run(pb, ic, 3, launch), oracle_run(orc, pb, ic, 3), "mc rough %s" % (launch,))
        ic = icgen.random_1d(n, 70 + bc, block=29)
        assert_same(pb, gpu_run(pb, ic, 80, launch), oracle_run(orc, pb, ic, 80), "mc blocky %s" % (launch,))
    pb = S.problem(1, 2, 1024, precision=prec, limiter=S.LIM_MC)
    ic = S.ic_tube(1024)
    assert_same(pb, gpu_run(pb, ic, 1639), oracle_run(orc, pb, ic, 1639), "mc sod")


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_2d_mc_limiter(orc, prec):
    banner("2d mc", prec)
    pb = S.problem(2, 2, 70, 150, precision=prec, limiter=S.LIM_MC)
    ic = icgen.random_2d(70, 150, 19, block=1)
    assert_same(pb, gpu_run(pb, ic, 2), oracle_run(orc, pb, ic, 2), "2d mc rough")
    ic = icgen.random_2d(70, 150, 23, block=11)
    assert_same(pb, gpu_run(pb, ic, 25), oracle_run(orc, pb, ic, 25), "2d mc blocky")
    pb = S.problem(2, 2, 128, 96, precision=prec, limiter=S.LIM_MC)
    ic = S.ic_four_contacts(128, 96)
    assert_same(pb, gpu_run(pb, ic, 60), oracle_run(orc, pb, ic, 60), "2d mc four contacts")


@pytest.mark.parametrize("name", ["mc_2d_contacts", "mc_2d_rand", "mc_2d_blocks"])
def test_2d_mc_reference_golden(name):
    """REF32 with the MC limiter reproduces base-omp/2nd_order_base_shll.c at its own
    1024^2 size bitwise (digests of every array after each recorded step)."""
    import hashlib
    import json
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    case = json.load(open(os.path.join(GOLD, "manifest.json")))["cases"][name]
    n = case["side"]
    pb = S.problem(2, 2, n, n, precision=S.REF32, limiter=S.LIM_MC, theta=1.25)
    spec = case["ic"]
    ic = S.ic_four_contacts(n, n) if spec is None else icgen.random_2d(n, n, spec["seed"], spec["block"])
    banner("2d mc golden", name)
    with S.Solver(pb, ic) as s:
        done = 0
        for k in case["steps"]:
            s.step(k - done)
            done = k
            p = s.primitives().astype(np.float32)
            u = s.state_raw()
            dig = [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in list(p) + list(u)]
            assert dig == list(g["dig_%d" % k][:8]), (name, k)


@pytest.mark.parametrize("prec", [S.F64, S.C32K])
def test_1d_mc_batch_equals_single_runs(orc, prec):
    """A batch of MC tubes with different theta gives each tube's single-run bits."""
    banner("1d mc batch", prec)
    sizes = [300, 1111, 4000]
    pbs = [S.problem(1, 2, n, precision=prec, limiter=S.LIM_MC, theta=th) for n, th in zip(sizes, (1.0, 1.25, 2.0))]
    ics = [icgen.random_1d(n, 90 + i, block=17) for i, n in enumerate(sizes)]
    steps = [40, 75, 60]
    b = S.Batch(pbs, ics, steps)
    b.set_launch(4, 8, 1, 16)
    b.run()
    for k, (pb, ic, ns) in enumerate(zip(pbs, ics, steps)):
        p, c, _ = oracle_run(orc, pb, ic, ns)
        np.testing.assert_array_equal(b.conserved(k), c, err_msg="tube %d" % k)
        np.testing.assert_array_equal(b.primitives(k), p, err_msg="tube %d" % k)
    b.close()


# ---------------------------------------------------------------------------
# 2D first order, several steps per launch (row-pipelined stages)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k", [2, 4, 8])
@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_2d_steps_per_launch(orc, prec, k):
    """k steps per launch give the oracle's bits: both walls, several row items and column
    strips, a last launch with fewer than k steps, graphs and direct launches."""
    for bc, (nx, ny), wpb, graph in ((0, (150, 130), 1, 0), (1, (203, 70), 4, 4 * k), (0, (9, 7), 4, 0)):
        pb = S.problem(2, 1, nx, ny, precision=prec, bc=bc)
        ic = icgen.random_2d(nx, ny, 31 + bc, block=9)
        steps = 3 * k + 1
        banner("2d k", prec, k, bc, nx, ny)
        assert_same(pb, gpu_run(pb, ic, steps, (1, k, wpb, graph)), oracle_run(orc, pb, ic, steps),
                    "2d k=%d %dx%d" % (k, nx, ny))
    pb = S.problem(2, 1, 96, 96, precision=prec)
    ic = S.ic_circle(96, 96)
    assert_same(pb, gpu_run(pb, ic, 60, (1, k, 4, 0)), oracle_run(orc, pb, ic, 60), "2d k circle")


def test_2d_steps_per_launch_rejected_in_second_order():
    """FP64 and REF32 second order run one step per launch (their stages need twice the
    registers); the compensated forms take 2."""
    pb = S.problem(2, 2, 64, 64)
    with S.Solver(pb, S.ic_four_shocks(64, 64)) as s:
        with pytest.raises(S.ShllError) as e:
            s.set_launch(1, 4, 4, 0)
        assert e.value.code == S.EINVAL
        with pytest.raises(S.ShllError):
            s.set_launch(1, 2, 4, 0)
        s.step(2)
    pb = S.problem(2, 2, 64, 64, precision=S.C32K)
    with S.Solver(pb, S.ic_four_shocks(64, 64)) as s:
        with pytest.raises(S.ShllError):
            s.set_launch(1, 4, 4, 0)
        s.set_launch(1, 2, 4, 0)
        s.step(3)


@pytest.mark.parametrize("k", [2, 4, 8])
@pytest.mark.parametrize("prec", [S.F64, S.C32K])
def test_2d_batch_steps_per_launch(orc, prec, k):
    """A batch of uneven domains at k steps per launch: each domain equals its own run,
    including domains that stop in the middle of a launch."""
    pbs, prims = S.domain_batch(57, 5, smin=20, smax=140, precision=prec)
    steps = [4 * k + 3, 2 * k, 11, 1, 5 * k]
    b = S.Batch(pbs, prims, steps)
    b.set_launch(1, k, 4, 2 * k)
    banner("2d batch k", prec, k)
    b.run()
    for i in range(len(pbs)):
        ref = oracle_run(orc, pbs[i], prims[i], steps[i])
        np.testing.assert_array_equal(b.conserved(i), ref[1], err_msg="domain %d" % i)
    b.close()


# ---------------------------------------------------------------------------
# edge cases: smallest grids, sizes around segment and strip edges, empty and
# one-instance batches, large grids
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_1d_smallest_and_segment_edge_sizes(orc, prec):
    """2..5 cells (both walls in one lane) and sizes around 64*C-cell segment edges, with
    launch shapes of 1, 2 and 8 cells per lane."""
    banner("1d sizes", prec)
    for order, bc in ((1, 0), (2, 1), (2, 0)):
        for n in (2, 3, 4, 5, 63, 64, 65, 66, 127, 130, 447, 449, 513):
            pb = S.problem(1, order, n, precision=prec, bc=bc)
            ic = icgen.random_1d(n, 100 + n, block=3)
            ora = oracle_run(orc, pb, ic, 21)
            for launch in ((1, 1, 1, 0), (2, 8, 4, 0), (8, 16, 1, 32)):
                assert_same(pb, gpu_run(pb, ic, 21, launch), ora, "n=%d o%d bc%d %s" % (n, order, bc, launch))


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_2d_smallest_grids(orc, prec):
    """Grids of 2 rows or columns and strips narrower than a wavefront, square cells (the
    reference's DT follows DX only), first order at 1 and 4 steps per launch, second order."""
    banner("2d sizes", prec)
    for nx, ny in ((2, 2), (2, 64), (64, 2), (3, 70), (65, 3), (5, 130), (130, 5)):
        h = 1.0 / max(nx, ny)
        ic = icgen.random_2d(nx, ny, 7, block=9)
        for order, ks in ((1, (1, 4)), (2, (1,))):
            for bc in (0, 1):
                pb = S.problem(2, order, nx, ny, precision=prec, bc=bc, lx=nx * h, ly=ny * h)
                ora = oracle_run(orc, pb, ic, 9)
                for k in ks:
                    assert_same(pb, gpu_run(pb, ic, 9, (1, k, 4, 0)), ora,
                                "%dx%d o%d bc%d k%d" % (nx, ny, order, bc, k))


@pytest.mark.parametrize("dim", [1, 2])
def test_batch_with_empty_and_tiny_instances(orc, dim):
    """A batch holding an instance with no steps (left at its initial state), the smallest
    instances and a single-launch one; and a batch of one instance."""
    banner("batch edges", dim)
    if dim == 1:
        sizes = [2, 3, 700, 64, 5000]
        pbs = [S.problem(1, 2, n, precision=S.C32K) for n in sizes]
        ics = [icgen.random_1d(n, 200 + i, block=2) for i, n in enumerate(sizes)]
        launch = (2, 8, 4, 16)
    else:
        sizes = [(2, 2), (3, 70), (90, 40), (64, 2), (150, 150)]
        pbs = []
        for nx, ny in sizes:
            h = 1.0 / max(nx, ny)
            pbs.append(S.problem(2, 1, nx, ny, precision=S.C32K, lx=nx * h, ly=ny * h))
        ics = [icgen.random_2d(nx, ny, 210 + i, block=5) for i, (nx, ny) in enumerate(sizes)]
        launch = (1, 4, 4, 8)
    steps = [0, 17, 40, 1, 33]
    b = S.Batch(pbs, ics, steps)
    b.set_launch(*launch)
    b.run()
    for i in range(len(pbs)):
        ref = oracle_run(orc, pbs[i], ics[i], steps[i])
        np.testing.assert_array_equal(b.conserved(i), ref[1], err_msg="instance %d" % i)
        np.testing.assert_array_equal(b.primitives(i), ref[0], err_msg="instance %d" % i)
    b.close()
    b = S.Batch(pbs[2:3], ics[2:3], [23])
    b.run()
    ref = oracle_run(orc, pbs[2], ics[2], 23)
    np.testing.assert_array_equal(b.conserved(0), ref[1])
    b.close()


def test_large_grids(orc):
    """A 4M-cell tube and an 8M-cell 2D grid, a few steps, against the oracle (bitwise)."""
    n = 1 << 22
    banner("large 1d", n)
    pb = S.problem(1, 2, n, precision=S.C32K)
    ic = icgen.random_1d(n, 301, block=4096)
    assert_same(pb, gpu_run(pb, ic, 5, (8, 16, 1, 0)), oracle_run(orc, pb, ic, 5), "1d 4M")
    nx, ny = 4096, 2048
    ic = icgen.random_2d(nx, ny, 302, block=256)
    for prec, k in ((S.F64, 2), (S.C32K, 4)):
        banner("large 2d", prec, k)
        pb = S.problem(2, 1, nx, ny, precision=prec, lx=1.0, ly=0.5)
        assert_same(pb, gpu_run(pb, ic, 6, (1, k, 4, 0)), oracle_run(orc, pb, ic, 6), "2d 8M k%d" % k)


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.C32, S.C32K])
def test_stage_fluxes_match_the_reference_stages(orc, prec):
    """Stage level: the sound speed and split fluxes the step kernels compute from a state
    equal the oracle's Compute_P_from_U + Compute_F_from_P (bitwise), 1D in both
    kinetic-energy associations and 2D, on rough random states."""
    banner("stage", prec)
    for pb, ic in ((S.problem(1, 1, 777, precision=prec), icgen.random_1d(777, 401, block=1)),
                   (S.problem(1, 2, 777, precision=prec), icgen.random_1d(777, 402, block=1)),
                   (S.problem(2, 1, 90, 70, precision=prec), icgen.random_2d(90, 70, 403, block=1))):
        with S.Solver(pb, ic) as s:
            got = s.stage_fluxes()
        want = orc.stage(opb(orc, pb), MODES[prec], ic)
        for key in got:
            np.testing.assert_array_equal(got[key], want[key].reshape(got[key].shape),
                                          err_msg="%s dim%d order%d" % (key, pb.dim, pb.order))


@pytest.mark.parametrize("lim", [S.LIM_MINMOD, S.LIM_MC])
@pytest.mark.parametrize("prec", [S.C32, S.C32K])
def test_2d_second_order_two_steps_per_launch(orc, prec, lim):
    """Second order, compensated forms, 2 steps per launch (two-row pipeline stages): the
    oracle's bits with both walls and limiters, several items and strips, small grids and
    a last launch of one step."""
    banner("2d o2 k2", prec, lim)
    for bc, (nx, ny), wpb, graph in ((1, (150, 130), 1, 0), (0, (203, 70), 4, 8), (1, (9, 7), 4, 0),
                                     (0, (2, 2), 1, 0), (1, (3, 70), 4, 0), (0, (70, 3), 1, 0)):
        h = 1.0 / max(nx, ny)
        pb = S.problem(2, 2, nx, ny, precision=prec, bc=bc, lx=nx * h, ly=ny * h, limiter=lim)
        ic = icgen.random_2d(nx, ny, 41 + bc, block=9)
        banner("  ", bc, nx, ny)
        assert_same(pb, gpu_run(pb, ic, 7, (1, 2, wpb, graph)), oracle_run(orc, pb, ic, 7),
                    "2d o2 k2 %dx%d bc%d" % (nx, ny, bc))
    pb = S.problem(2, 2, 128, 96, precision=prec, limiter=lim)
    ic = S.ic_four_contacts(128, 96)
    assert_same(pb, gpu_run(pb, ic, 40, (1, 2, 4, 0)), oracle_run(orc, pb, ic, 40), "2d o2 k2 contacts")
    pbs, prims = S.domain_batch(58, 4, smin=20, smax=120, precision=prec, order=2)
    for p in pbs:
        p.limiter = lim
    steps = [9, 4, 1, 6]
    b = S.Batch(pbs, prims, steps)
    b.set_launch(1, 2, 4, 4)
    b.run()
    for i in range(len(pbs)):
        np.testing.assert_array_equal(b.conserved(i), oracle_run(orc, pbs[i], prims[i], steps[i])[1],
                                      err_msg="batch domain %d" % i)
    b.close()
