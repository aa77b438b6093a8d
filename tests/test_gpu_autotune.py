"""GPU: the autotuner (shll_autotune / shll_batch_autotune).  It times candidate launch
shapes from a saved state, keeps the fastest, puts the state back, and records the choice
in $SHLL_TUNE_CACHE (one line per device, kernel and shape class); a second call with the
same key takes the recorded shape without timing.  The choice never changes results
(bitwise against the oracle).  In 2D, when the default work items give fewer than 4
wavefronts per SIMD, shorter items are among the candidates."""
import numpy as np
import pytest

import icgen
import shll_amd as S
from test_gpu_parity import banner, opb

pytestmark = pytest.mark.gpu


def read_cache(path):
    return [line.strip().split("|") for line in open(path) if line.strip()]


def test_single_1d_autotune_restores_the_state_and_records_its_choice(orc, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SHLL_TUNE_CACHE", str(cache))
    pb = S.problem(1, 2, 3000, precision=S.F64)
    ic = icgen.random_1d(3000, 21, block=30)
    banner("autotune 1d")
    with S.Solver(pb, ic) as s:
        s.step(7)
        before = s.state_raw()
        s.autotune(64)
        shape = s.get_launch()
        assert s.state_raw().tobytes() == before.tobytes()  # the trial runs are undone
        s.step(93)
        got = s.conserved()
    lines = read_cache(cache)
    assert len(lines) == 1 and " dim1 order2 prec0 batch0 " in lines[0][0] + " ", lines
    assert tuple(int(x) for x in lines[0][1].split()) == shape
    _, ref = orc.run(opb(orc, pb), orc.F64, ic, 100, nthreads=4)
    np.testing.assert_array_equal(got, ref)
    # the same key again: the recorded shape, and no second line
    with S.Solver(pb, ic) as s:
        s.autotune(64)
        assert s.get_launch() == shape
    assert len(read_cache(cache)) == 1


def test_2d_autotune_on_a_small_grid(orc, tmp_path, monkeypatch):
    """256^2, second order, MC: the default 16-row items are 80 work items (5 strips of 60
    columns, 16 bands), far fewer than 4 per SIMD, so items of 12, 8, 6, 4, 3 and 2 rows
    are candidates too; whichever wins, the run keeps the oracle's bits."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SHLL_TUNE_CACHE", str(cache))
    pb = S.problem(2, 2, 256, 256, precision=S.REF32, limiter=S.LIM_MC, theta=1.25)
    ic = S.ic_four_contacts(256, 256)
    banner("autotune 2d")
    with S.Solver(pb, ic) as s:
        s.autotune(40)
        rows, k, w, g = s.get_launch()
        s.step(30)
        got = s.conserved()
    assert k == 1 and w in (1, 4) and rows in (1, 12, 8, 6, 4, 3, 2), (rows, k, w, g)
    (line,) = read_cache(cache)
    assert " dim2 order2mc prec1 batch0 " in line[0] + " ", line
    _, ref = orc.run(opb(orc, pb), orc.REF32, ic, 30, nthreads=4)
    np.testing.assert_array_equal(got, ref)


def test_batch_autotune(orc, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SHLL_TUNE_CACHE", str(cache))
    pbs, prims, steps = S.tube_batch(4321, 6, nmin=300, nmax=5000, precision=S.C32K)
    steps = [max(1, st // 20) for st in steps]
    banner("autotune batch", [p.nx for p in pbs], steps)
    b = S.Batch(pbs, prims, steps)
    b.autotune(64)
    shape = b.get_launch()
    assert shape[1] >= 4, shape  # batches try 4 or more steps per launch
    b.run()
    for k in range(len(pbs)):
        _, ref = orc.run(opb(orc, pbs[k]), orc.C32K, prims[k], steps[k], nthreads=4)
        np.testing.assert_array_equal(b.conserved(k), ref, err_msg="instance %d" % k)
    b.close()
    (line,) = read_cache(cache)
    assert " batch1 " in line[0] + " ", line
