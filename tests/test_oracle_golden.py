"""The CPU oracle's reference-precision mode against golden vectors written by the
compiled reference solvers (tests/golden/make_golden.py).  Bitwise."""
import hashlib
import json
import os

import numpy as np
import pytest

import icgen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
N2 = 256


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def side_of(case):
    return case.get("side", N2)


def ic_for(orc, case):
    spec = case["ic"]
    n = side_of(case)
    if spec is None:
        if case["scheme"] == "1d_o1":
            return orc.ic_tube(n)
        if case["scheme"] == "2d_o1":
            return orc.ic_implosion(n, n)
        if case["scheme"] == "2d_o2mc":
            return orc.ic_four_contacts(n, n)
        return orc.ic_four_shocks(n, n)
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


def problem_for(orc, scheme, n=N2):
    """The reference file of each scheme: base_shll.c, base_shll_2d.c, 2nd_order_base_shll.c,
    base-omp/2nd_order_base_shll.c (MC limiter, alpha = 1.25)."""
    if scheme == "1d_o1":
        return orc.problem(1, 1, n)
    if scheme == "2d_o1":
        return orc.problem(2, 1, n, n)
    if scheme == "2d_o2mc":
        return orc.problem(2, 2, n, n, limiter=orc.LIM_MC, theta=1.25)
    return orc.problem(2, 2, n, n)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


CASES = sorted(MANIFEST["cases"].items())


@pytest.mark.parametrize("name,case", CASES, ids=[c[0] for c in CASES])
def test_ref32_matches_reference(orc, name, case):
    g = load(name)
    pb = problem_for(orc, case["scheme"], side_of(case))
    ic = ic_for(orc, case)
    for k in case["steps"]:
        if k == 0:
            st = orc.stage(pb, orc.REF32, ic)
            p, u, a = (st["prim"].astype(np.float32), st["cons"].astype(np.float32),
                       st["a"].astype(np.float32))
        else:
            po, co = orc.run(pb, orc.REF32, ic, k, nthreads=4)
            p, u, a = po.astype(np.float32), co.astype(np.float32), None
        if case["scheme"] == "1d_o1":
            np.testing.assert_array_equal(p, g["p_%d" % k], err_msg="%s p step %d" % (name, k))
            np.testing.assert_array_equal(u, g["u_%d" % k], err_msg="%s u step %d" % (name, k))
            if a is not None:
                np.testing.assert_array_equal(a, g["a_%d" % k])
        else:
            nv = p.shape[0]
            dig = list(g["dig_%d" % k])
            idx = g["idx"]
            np.testing.assert_array_equal(p[:, idx], g["ps_%d" % k], err_msg="%s p sample step %d" % (name, k))
            np.testing.assert_array_equal(u[:, idx], g["us_%d" % k], err_msg="%s u sample step %d" % (name, k))
            assert [digest(p[v]) for v in range(nv)] == dig[:nv], (name, k, "prim digest")
            assert [digest(u[v]) for v in range(nv)] == dig[nv:2 * nv], (name, k, "cons digest")
            if a is not None:
                assert digest(a) == dig[2 * nv]


def test_reference_step_counts(orc):
    """README.md:89-99 (1D, t=0.2), :201-204 (2D, t=0.1), :247-248 (2D, t=0.8); the golden
    main() runs (410, 205 and 1639 steps)."""
    for n, steps in [(256, 410), (512, 820), (1024, 1639), (2048, 3277), (4096, 6554), (8192, 13108),
                     (16384, 26215), (32768, 52429), (65536, 104858)]:
        assert orc.steps_to_time(orc.problem(1, 1, n), 0.2) == steps
    for n, steps in [(256, 205), (512, 410), (1024, 820), (2048, 1639)]:
        assert orc.steps_to_time(orc.problem(2, 1, n, n), 0.1) == steps
    for n, steps in [(256, 1639), (512, 3277), (1024, 6554)]:
        assert orc.steps_to_time(orc.problem(2, 2, n, n), 0.8) == steps
    for name in ("o1_1d_main", "o1_2d_main", "o2_2d_main"):
        assert str(MANIFEST["cases"][name]["steps"][0]) in bytes(load(name)["stdout"]).decode()


@pytest.mark.parametrize("name", ["o2_2d_sody", "o2_2d_yuni", "mc_2d_sody", "mc_2d_yuni"])
def test_1d_second_order_is_the_2d_file_on_y_uniform_data(orc, name):
    """The derived 1D second-order scheme equals every x line of the reference's 2D
    second-order solver run on y-uniform data with v = 0 (bitwise); with either limiter."""
    case = MANIFEST["cases"][name]
    g = load(name)
    n = side_of(case)
    ic2 = ic_for(orc, case)
    ic1 = ic2.reshape(4, n, n)[[0, 1, 3], :, 0]
    lim = orc.LIM_MC if case["scheme"] == "2d_o2mc" else orc.LIM_MINMOD
    pb = orc.problem(1, 2, n, limiter=lim)
    for k in case["steps"]:
        if k == 0:
            st = orc.stage(pb, orc.REF32, ic1)
            p, u = st["prim"], st["cons"]
        else:
            p, u = orc.run(pb, orc.REF32, ic1, k)
        np.testing.assert_array_equal(p.astype(np.float32), g["prow_%d" % k][[0, 1, 3]])
        np.testing.assert_array_equal(u.astype(np.float32), g["urow_%d" % k][[0, 1, 3]])
        assert np.all(g["prow_%d" % k][2] == 0) and np.all(g["urow_%d" % k][2] == 0)


def test_initial_conditions_match_reference(orc):
    """The oracle's IC generators reproduce the reference's own step-0 arrays."""
    g = load("o1_1d_sod")
    np.testing.assert_array_equal(orc.ic_tube(256).astype(np.float32), g["p_0"])
    for name, ic in (("o1_2d_implosion", orc.ic_implosion(N2, N2)), ("o2_2d_fourshocks", orc.ic_four_shocks(N2, N2)),
                     ("mc_2d_contacts", orc.ic_four_contacts(1024, 1024))):
        gg = load(name)
        case = MANIFEST["cases"][name]
        st = orc.stage(problem_for(orc, case["scheme"], side_of(case)), orc.REF32, ic)
        assert [digest(st["prim"][v]) for v in range(4)] == list(gg["dig_0"][:4])


def test_circle_ic(orc):
    """Circle IC of the oracle equals icgen's, and the cell count matches a brute-force count."""
    a = orc.ic_circle(N2, N2)
    b = icgen.circle_2d(N2, N2)
    np.testing.assert_array_equal(a, b)
    cnt = 0
    for i in range(N2):
        for j in range(N2):
            x, y = (i + 0.5) / N2 - 0.5, (j + 0.5) / N2 - 0.5
            cnt += x * x + y * y < 0.0625
    assert int((a[0] == 10.0).sum()) == cnt
    assert abs(cnt / N2 ** 2 - np.pi * 0.0625) < 2e-3
