"""GPU: the drop-in path exercised as C, and the momentum ledger on the GPU.

bin/shll_run is the reference's main() over the C ABI (host/shll_run.c).  Run with its
defaults it is base_shll.c (1D Sod, N = 256, t = 0.2, reflective walls, the reference's
float arithmetic): it must print the reference's own "Completed in 410 steps" and write
the reference's results.dat byte for byte.  With --dim 2 it is base_shll_2d.c (205 steps)
and with --dim 2 --order 2 it is 2nd_order_base_shll.c (1639 steps): the end states must
have the digests of the reference's own arrays.

Momentum ledger: reflective walls exert pressure, so one step changes the momentum sum by
exactly the wall fluxes (base_shll.c:95-110, base_shll_2d.c:150-216) while mass and
energy fluxes through the walls vanish."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import icgen
import shll_amd as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "shll-sve-cfd_amd", "bin", "shll_run")
GOLD = os.path.join(ROOT, "tests", "golden")


def run_driver(tmp_path, *args):
    print("[gpu-case] shll_run", *args, flush=True)
    r = subprocess.run([DRIVER, *args], cwd=tmp_path, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    return r.stdout


def test_driver_defaults_are_base_shll_c(tmp_path):
    g = np.load(os.path.join(GOLD, "o1_1d_main.npz"), allow_pickle=False)
    out = run_driver(tmp_path)
    assert out == bytes(g["stdout"])
    assert (tmp_path / "results.dat").read_bytes() == bytes(g["results_dat"])


@pytest.mark.parametrize("name,args,steps", [("o1_2d_main", ["--dim", "2"], 205),
                                             ("o2_2d_main", ["--dim", "2", "--order", "2"], 1639)])
def test_driver_2d_reference_runs(tmp_path, name, args, steps):
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    out = run_driver(tmp_path, *args, "--dump", "end.bin")
    assert out == bytes(g["stdout"])
    n = 256 * 256
    raw = (tmp_path / "end.bin").read_bytes()
    p = np.frombuffer(raw[: 4 * n * 8], dtype=np.float64).reshape(4, n).astype(np.float32)
    u = np.frombuffer(raw[4 * n * 8:], dtype=np.float32).reshape(4, n)
    dig = [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in list(p) + list(u)]
    assert dig == list(g["dig_%d" % steps][:8])
    # the six-column results.dat of the same run parses back to the same values (7 digits)
    rows = np.loadtxt(tmp_path / "results.dat")
    assert rows.shape == (n, 6)
    np.testing.assert_allclose(rows[:, 2], p[0], rtol=6e-7)
    np.testing.assert_allclose(rows[:, 5], p[3], rtol=6e-7)


def test_driver_json_problem_file(tmp_path):
    """The same 2nd_order_base_shll.c run from a JSON problem file: same stdout, same digests."""
    import json
    g = np.load(os.path.join(GOLD, "o2_2d_main.npz"), allow_pickle=False)
    (tmp_path / "run.json").write_text(json.dumps({"dim": 2, "order": 2, "nx": 256, "dump": "end.bin",
                                                   "out": "-"}))
    out = run_driver(tmp_path, "--config", "run.json")
    assert out == bytes(g["stdout"])
    n = 256 * 256
    raw = (tmp_path / "end.bin").read_bytes()
    u = np.frombuffer(raw[4 * n * 8:], dtype=np.float32).reshape(4, n)
    assert [hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest() for x in u] == list(g["dig_1639"][4:8])
    assert not (tmp_path / "results.dat").exists()


@pytest.mark.parametrize("prec", [S.F64, S.REF32, S.F64F, S.C32K])
@pytest.mark.parametrize("order", [1, 2])
def test_momentum_change_equals_wall_impulse_1d(prec, order):
    n = 300
    pb = S.problem(1, order, n, precision=prec, bc=S.BC_REFLECT)
    prim = icgen.random_1d(n, 5 + order, block=10)
    dx = 1.0 / n
    tol = 1e-13 if prec in (S.F64, S.F64F) else 2e-6
    with S.Solver(pb, prim) as s:
        st = s.stage_fluxes()
        c0 = s.conserved_sums()
        lam = 0.125
        s.step(1)
        c1 = s.conserved_sums()
    d = c1 - c0
    impulse = -lam * (2.0 * st["fp"][1, -1] - 2.0 * st["fm"][1, 0]) * dx
    scale = abs(c0[1]) + abs(impulse) + 1e-300
    assert abs(d[1] - impulse) <= tol * max(scale, abs(impulse) * 1e3), (d, impulse)
    assert abs(d[0]) <= tol * abs(c0[0]) and abs(d[2]) <= tol * abs(c0[2]), d
    assert abs(impulse) > 1e-6


def test_momentum_change_equals_wall_impulse_2d():
    nx, ny = 48, 40
    pb = S.problem(2, 1, nx, ny, precision=S.F64, lx=1.0, ly=ny / nx)
    prim = icgen.random_2d(nx, ny, 17, block=6)
    vol = (1.0 / nx) * ((ny / nx) / ny)
    with S.Solver(pb, prim) as s:
        st = s.stage_fluxes()
        c0 = s.conserved_sums()
        s.step(1)
        c1 = s.conserved_sums()
    fp = st["fp"].reshape(4, nx, ny)
    fm = st["fm"].reshape(4, nx, ny)
    hp = st["hp"].reshape(4, nx, ny)
    hm = st["hm"].reshape(4, nx, ny)
    lamx = 0.125
    lamy = 0.125 * (1.0 / nx) / ((ny / nx) / ny)  # DT / DY with square cells
    ix = -lamx * (2.0 * fp[1, -1, :].sum() - 2.0 * fm[1, 0, :].sum()) * vol
    iy = -lamy * (2.0 * hp[2, :, -1].sum() - 2.0 * hm[2, :, 0].sum()) * vol
    d = c1 - c0
    assert abs(d[1] - ix) <= 1e-13 * (np.abs(c0[1]) + abs(ix) + 1.0), (d[1], ix)
    assert abs(d[2] - iy) <= 1e-13 * (np.abs(c0[2]) + abs(iy) + 1.0), (d[2], iy)
    assert abs(d[0]) <= 1e-13 * c0[0] and abs(d[3]) <= 1e-13 * c0[3]
    assert abs(ix) > 1e-6 and abs(iy) > 1e-6


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "integration_base_shll.run")),
                    reason="built only where the reference tree was present at build time")
def test_integration_example_is_the_reference_file_on_the_gpu(tmp_path):
    """The integration guide's gpu_main.c linked with the reference's own base_shll.c (main renamed):
    the reference's stdout line and results.dat, byte for byte."""
    g = np.load(os.path.join(GOLD, "o1_1d_main.npz"), allow_pickle=False)
    exe = os.path.join(ROOT, "oracle", "_ref", "integration_base_shll.run")
    print("[gpu-case] integration example", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    assert r.stdout == bytes(g["stdout"])
    assert (tmp_path / "results.dat").read_bytes() == bytes(g["results_dat"])
