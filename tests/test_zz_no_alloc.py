"""GPU: the allocation-free conserved sums (shll_conserved_sums_noalloc) through the C ABI
give the bits of shll_conserved_sums, on two calls in a row -- the second on a smaller
problem, so it reuses the scratch the first call left behind -- and after the state
advances."""
import ctypes

import numpy as np
import pytest

import icgen
import shll_amd as S

pytestmark = pytest.mark.gpu


def sums_both(s):
    a = np.zeros(4)
    b = np.zeros(4)
    assert S.lib().shll_conserved_sums_noalloc(s.h, S._dp(b)) == S.OK, S.lib().shll_last_error()
    assert S.lib().shll_conserved_sums(s.h, S._dp(a)) == S.OK
    return a[: s.nv], b[: s.nv]


@pytest.mark.parametrize("prec", [S.F64, S.C32K, S.F64F])
def test_noalloc_sums_equal_the_allocating_entry_point(prec):
    print("[gpu-case] noalloc sums", prec, flush=True)
    big = S.problem(1, 2, 300000, precision=prec)   # the sums kernel's full 256-block grid
    small = S.problem(1, 2, 1000, precision=prec)   # 4 blocks
    with S.Solver(big, icgen.random_1d(300000, 71, block=50)) as s:
        want, got = sums_both(s)
        assert got.tobytes() == want.tobytes()
        s.step(20)
        want, got = sums_both(s)
        assert got.tobytes() == want.tobytes()
    with S.Solver(small, icgen.random_1d(1000, 72, block=5)) as s:
        # first call on this context: the allocating entry point has not run yet
        got = np.zeros(4)
        assert S.lib().shll_conserved_sums_noalloc(s.h, S._dp(got)) == S.OK
        want = s.conserved_sums()
        assert got[: s.nv].tobytes() == want.tobytes()
    if prec != S.F64F:
        pb2 = S.problem(2, 1, 300, 200, precision=prec, lx=1.0, ly=200 / 300)
        with S.Solver(pb2, S.ic_circle(300, 200)) as s:
            s.step(7)
            want, got = sums_both(s)
            assert got.tobytes() == want.tobytes()


def test_noalloc_sums_reject_a_null_context():
    out = np.zeros(4)
    assert S.lib().shll_conserved_sums_noalloc(None, S._dp(out)) == S.EINVAL
    assert S.lib().shll_last_error()
