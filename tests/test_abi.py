"""C-ABI library checks that need no GPU: it loads, exports every entry point the
header declares, and its host-side logic (time-loop step counts, initial data,
argument validation) matches the oracle.  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

import shll_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shll.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shll_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = S.lib()
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(S.EXPORTS)
    assert b"gfx950" in L.shll_version()


def test_steps_to_time_matches_oracle(orc):
    for n in (256, 1000, 1024, 4097, 65536):
        for order in (1, 2):
            assert S.steps_to_time(S.problem(1, order, n), 0.2) == orc.steps_to_time(orc.problem(1, order, n), 0.2)
    assert S.steps_to_time(S.problem(2, 1, 256, 256), 0.1) == 205
    assert S.steps_to_time(S.problem(2, 2, 256, 256), 0.8) == 1639
    assert S.steps_to_time(S.problem(1, 1, 65536), 0.2) == 104858


def test_initial_data_matches_oracle(orc):
    np.testing.assert_array_equal(S.ic_tube(1024), orc.ic_tube(1024))
    np.testing.assert_array_equal(S.ic_tube(999, 0.37, 7.25, 0.5, 1.5, 1.0, -0.25, 1.0),
                                  orc.ic_tube(999, 0.37, 7.25, 0.5, 1.5, 1.0, -0.25, 1.0))
    np.testing.assert_array_equal(S.ic_implosion(100, 60), orc.ic_implosion(100, 60))
    np.testing.assert_array_equal(S.ic_four_shocks(64, 96), orc.ic_four_shocks(64, 96))
    np.testing.assert_array_equal(S.ic_circle(128, 96), orc.ic_circle(128, 96))
    for nx, ny in ((64, 96), (1024, 1024), (33, 17)):
        np.testing.assert_array_equal(S.ic_four_contacts(nx, ny), orc.ic_four_contacts(nx, ny))


def test_batch_generators_are_counter_based_and_in_range():
    a = [S.gen_tube(1234, k) for k in range(512)]
    b = S.gen_tube(1234, 77)
    assert (a[77].n, a[77].rhoL, a[77].TL, a[77].x0) == (b.n, b.rhoL, b.TL, b.x0)
    ns = np.array([t.n for t in a])
    assert ns.min() >= 256 and ns.max() <= 65536
    # log-uniform: median near sqrt(256*65536) = 4096
    assert 2500 < np.median(ns) < 6500
    assert all(2 <= t.rhoL <= 20 and 0.5 <= t.TL <= 2 and 0.3 <= t.x0 <= 0.7 for t in a)
    assert all(float(np.float32(t.rhoL)) == t.rhoL for t in a)
    d = [S.gen_domain(99, k) for k in range(64)]
    assert all(256 <= x.nx <= 2048 and 256 <= x.ny <= 2048 and 1 <= x.ncircles <= 4 for x in d)
    for x in d[:4]:
        p = S.ic_domain(x)
        assert p.shape == (4, x.nx * x.ny) and p[0].min() >= 1.0 and p[0].max() <= 20.0


def test_invalid_problems_are_rejected_before_any_device_call():
    prim = np.zeros((3, 8))
    h = ctypes.c_void_p()
    for pb in (S.problem(3, 1, 8), S.problem(1, 3, 8), S.problem(1, 1, 1), S.problem(1, 1, 8, precision=7),
               S.Problem(1, 1, 0, 0, 8, 2, 1.0, 1.0), S.problem(2, 1, 8, 1), S.problem(1, 1, 8, lx=0.0),
               S.problem(1, 2, 8, limiter=2), S.problem(1, 1, 8, limiter=S.LIM_MC),
               S.problem(1, 2, 8, limiter=S.LIM_MC, theta=0.5), S.problem(2, 2, 8, 8, limiter=S.LIM_MC, theta=2.5)):
        rc = S.lib().shll_init(ctypes.byref(pb), S._dp(prim), ctypes.byref(h))
        assert rc == S.EINVAL, (pb.dim, pb.order, pb.nx, pb.ny, rc)
        assert S.lib().shll_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_cpu_fallback_without_a_gpu():
    """Without a device the library refuses to run (there is no CPU path)."""
    pb = S.problem(1, 1, 64)
    with pytest.raises(S.ShllError) as e:
        S.Solver(pb, S.ic_tube(64))
    assert e.value.code == S.EDEVICE


def test_results_reader_reads_the_reference_results_file(tmp_path):
    """read_results parses the reference's own results.dat (base_shll.c, N = 256, 410 steps)
    into the arrays the golden run holds."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "o1_1d_main.npz"), allow_pickle=False)
    path = tmp_path / "results.dat"
    path.write_bytes(bytes(g["results_dat"]))
    r = S.read_results(path, 256)
    assert sorted(r) == ["T", "rho", "u", "x"]
    # "%e" keeps 7 significant digits
    np.testing.assert_allclose(r["rho"], g["p_410"][0], rtol=6e-7)
    np.testing.assert_allclose(r["u"], g["p_410"][1], rtol=6e-7, atol=1e-12)
    np.testing.assert_allclose(r["T"], g["p_410"][2], rtol=6e-7)
    np.testing.assert_allclose(r["x"], (np.arange(256) + 0.5) / 256, rtol=1e-6)
    with pytest.raises(ValueError):
        S.read_results(path, 255)


def test_steps_to_time_reports_a_float_time_loop_that_never_ends(orc):
    """At 2^24 cells DT = 2^-27 is half an ulp of the float time once time reaches 1/8, so
    the reference's loop `time += DT` (base_shll.c:200-218) stops growing before t = 0.2
    and would never end: the library says so (-1) at once instead of spinning."""
    import time
    t0 = time.perf_counter()
    assert S.steps_to_time(S.problem(1, 2, 1 << 24), 0.2) == -1
    assert time.perf_counter() - t0 < 5.0
    # the float sum stalls at 1/8: an end time below that still counts normally
    n = S.steps_to_time(S.problem(1, 2, 1 << 24), 0.1)
    assert n > 0 and n == orc.steps_to_time(orc.problem(1, 2, 1 << 24), 0.1)
    # and the largest sizes of the configurations are far from it
    assert S.steps_to_time(S.problem(1, 2, 65536), 0.2) == 104858


def test_integration_example_compiles_against_the_header(tmp_path):
    """The integration guide's gpu_main.c (its first C block) is valid C99 against include/shll.h
    (the GPU test links and runs it with the reference's own base_shll.c)."""
    import subprocess
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"^```c\n(.*?)^```$", text, flags=re.S | re.M)
    assert m and "int main(void)" in m.group(1)
    src = tmp_path / "gpu_main.c"
    src.write_text(m.group(1))
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
