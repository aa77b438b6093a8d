"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
imports or calls it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F64, REF32, C32, C32K, F64F = 0, 1, 2, 3, 4
BC_REFLECT, BC_OUTFLOW = 0, 1
LIM_MINMOD, LIM_MC = 0, 1


class Problem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("bc", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("limiter", ctypes.c_int32),
        ("lx", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("theta", ctypes.c_double),
    ]


def problem(dim, order, nx, ny=1, bc=None, lx=1.0, ly=1.0, limiter=LIM_MINMOD, theta=1.25):
    """The reference pairs first order with reflective walls and second order with outflow.
    limiter LIM_MC with theta = 1.25 is base-omp/2nd_order_base_shll.c:44,317-325."""
    if bc is None:
        bc = BC_REFLECT if order == 1 else BC_OUTFLOW
    return Problem(dim, order, bc, nx, ny if dim == 2 else 1, limiter, lx, ly, theta)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Problem)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        L.orc_nvar.argtypes = [P]
        L.orc_steps_to_time.argtypes = [P, ctypes.c_float]
        L.orc_steps_to_time.restype = ctypes.c_int64
        L.orc_constants.argtypes = [P, ctypes.c_int, dp]
        L.orc_run.argtypes = [P, ctypes.c_int, dp, ctypes.c_int64, dp, dp, ctypes.c_int, dp]
        L.orc_run_c32_raw.argtypes = [P, ctypes.c_int, dp, ctypes.c_int64, fp, fp]
        L.orc_stage.argtypes = [P, ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_f64f_rsq_nudge.argtypes = [ctypes.c_int]
        L.orc_minmod.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_minmod.restype = ctypes.c_double
        L.orc_ic_tube.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7 + [dp]
        L.orc_ic_implosion.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_ic_four_shocks.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_ic_four_contacts.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_slope.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4
        L.orc_slope.restype = ctypes.c_double
        L.orc_ic_circle.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 9 + [dp]
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def nvar(pb):
    return 3 if pb.dim == 1 else 4


def ncells(pb):
    return pb.nx * pb.ny


def steps_to_time(pb, t_end):
    return int(lib().orc_steps_to_time(ctypes.byref(pb), ctypes.c_float(t_end)))


def constants(pb, mode):
    out = np.zeros(6)
    lib().orc_constants(ctypes.byref(pb), mode, _dp(out))
    return dict(zip(["gamma", "R", "cv", "dt", "dtdx", "dtdy"], out))


def run(pb, mode, prim, nsteps, nthreads=1, timed=False):
    """Returns (prim_out, cons_out) as (nvar, ncells) float64 arrays [, elapsed seconds]."""
    prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(nvar(pb), ncells(pb))
    po = np.empty_like(prim)
    co = np.empty_like(prim)
    el = ctypes.c_double(0.0)
    rc = lib().orc_run(ctypes.byref(pb), mode, _dp(prim), int(nsteps), _dp(po), _dp(co),
                       int(nthreads), ctypes.byref(el))
    if rc != 0:
        raise ValueError("orc_run failed (%d)" % rc)
    if timed:
        return po, co, el.value
    return po, co


def run_c32_raw(pb, prim, nsteps, mode=C32):
    """Raw (hi, lo) float planes of a C32 or C32K run."""
    prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(nvar(pb), ncells(pb))
    hi = np.empty(prim.shape, dtype=np.float32)
    lo = np.empty(prim.shape, dtype=np.float32)
    rc = lib().orc_run_c32_raw(ctypes.byref(pb), int(mode), _dp(prim), int(nsteps),
                               hi.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               lo.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc != 0:
        raise ValueError("orc_run_c32_raw failed (%d)" % rc)
    return hi, lo


def stage(pb, mode, prim):
    """One evaluation of U<-P, P<-U and the split fluxes.  Returns a dict of arrays."""
    prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(nvar(pb), ncells(pb))
    out = {k: np.zeros_like(prim) for k in ("cons", "prim", "fp", "fm", "hp", "hm")}
    out["a"] = np.zeros(ncells(pb))
    rc = lib().orc_stage(ctypes.byref(pb), mode, _dp(prim), _dp(out["cons"]), _dp(out["prim"]),
                         _dp(out["a"]), _dp(out["fp"]), _dp(out["fm"]), _dp(out["hp"]),
                         _dp(out["hm"]))
    if rc != 0:
        raise ValueError("orc_stage failed (%d)" % rc)
    if pb.dim == 1:
        del out["hp"], out["hm"]
    return out


def f64f_rsq_nudge(ulps):
    """Test switch: the F64F mode's 1/sqrt moved by a pseudo-random 0..ulps ulp (0: off)."""
    lib().orc_f64f_rsq_nudge(int(ulps))


def minmod(mode, l, r):
    return lib().orc_minmod(mode, float(l), float(r))


def slope(mode, limiter, theta, ql, qc, qr):
    return lib().orc_slope(mode, limiter, float(theta), float(ql), float(qc), float(qr))


def ic_four_contacts(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().orc_ic_four_contacts(nx, ny, _dp(out))
    return out


def ic_tube(n, x0=0.5, rhoL=10.0, uL=0.0, TL=1.0, rhoR=1.0, uR=0.0, TR=1.0):
    out = np.zeros((3, n))
    lib().orc_ic_tube(n, x0, rhoL, uL, TL, rhoR, uR, TR, _dp(out))
    return out


def ic_implosion(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().orc_ic_implosion(nx, ny, _dp(out))
    return out


def ic_four_shocks(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().orc_ic_four_shocks(nx, ny, _dp(out))
    return out


def ic_circle(nx, ny, lx=1.0, ly=1.0, cx=0.5, cy=0.5, r=0.25, rho_in=10.0, T_in=1.0,
              rho_out=1.0, T_out=1.0):
    out = np.zeros((4, nx * ny))
    lib().orc_ic_circle(nx, ny, lx, ly, cx, cy, r, rho_in, T_in, rho_out, T_out, _dp(out))
    return out
