"""Python binding of the MI355X SHLL library (lib/libshll_amd.so) through its C ABI.

A thin ctypes layer: every call goes to the HIP library; nothing here computes.
If the library is missing or no GPU is usable, calls raise (there is no CPU path).

Reference routine -> call (base-c/base_shll.c):
    Allocate_and_Init_Memory + Compute_U_from_P + Compute_P_from_U  ->  Solver(...)
    the step loop body {Compute_F_from_P; Update_U_from_F; Compute_P_from_U}
                                                               ->  Solver.step(n)
    while (time < TOTAL_TIME) {...}                              ->  Solver.run_to_time(t)
    Save_Results                                                 ->  Solver.primitives(),
                                                                     Solver.write_results(path)
    Free_Memory                                                  ->  Solver.close()
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.environ.get("SHLL_AMD_LIB", os.path.join(PKG, "lib", "libshll_amd.so"))

F64, REF32, C32, C32K, F64F = 0, 1, 2, 3, 4
BC_REFLECT, BC_OUTFLOW = 0, 1
LIM_MINMOD, LIM_MC = 0, 1
LIMITER_NAMES = {"minmod": LIM_MINMOD, "mc": LIM_MC}
PREC_NAMES = {"f64": F64, "ref32": REF32, "c32": C32, "c32k": C32K, "f64f": F64F}

OK, EINVAL, ENOMEM, EDEVICE, ENONFINITE, ESTATE = 0, 1, 2, 3, 4, 5


class ShllError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("shll error %d: %s" % (code, msg))
        self.code = code


class Problem(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("order", ctypes.c_int32), ("bc", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("limiter", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("theta", ctypes.c_double)]


class Errors(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "l1_rho", "l1_u", "l1_p", "linf_rho", "linf_p", "max_rel_rho", "max_rel_p", "max_du",
        "max_abs_u_ref", "nonfinite")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TubeSpec(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x0", ctypes.c_double), ("rhoL", ctypes.c_double),
                ("TL", ctypes.c_double), ("rhoR", ctypes.c_double), ("TR", ctypes.c_double)]


class DomainSpec(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("lx", ctypes.c_double),
                ("ly", ctypes.c_double), ("ncircles", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("cx", ctypes.c_double * 4), ("cy", ctypes.c_double * 4), ("r", ctypes.c_double * 4),
                ("rho", ctypes.c_double * 4), ("T", ctypes.c_double * 4)]


EXPORTS = [
    "shll_init", "shll_step", "shll_run_to_time", "shll_steps_to_time", "shll_get_primitives",
    "shll_get_conserved", "shll_state_bytes", "shll_get_state_raw", "shll_set_state_raw", "shll_write_results",
    "shll_write_results_full", "shll_stage_fluxes",
    "shll_conserved_sums", "shll_conserved_sums_noalloc", "shll_check_state", "shll_compare", "shll_compare_host", "shll_free",
    "shll_set_launch", "shll_get_launch", "shll_autotune", "shll_set_timing", "shll_last_timing",
    "shll_batch_init", "shll_batch_advance", "shll_batch_run", "shll_batch_get_primitives",
    "shll_batch_get_conserved", "shll_batch_conserved_sums", "shll_batch_set_launch",
    "shll_batch_autotune", "shll_batch_get_launch", "shll_batch_check_state", "shll_batch_compare",
    "shll_batch_set_timing", "shll_batch_last_timing",
    "shll_batch_total_cell_updates", "shll_batch_reset", "shll_batch_free", "shll_ic_tube",
    "shll_ic_implosion", "shll_ic_four_shocks", "shll_ic_four_contacts", "shll_ic_circle", "shll_gen_tube", "shll_gen_domain",
    "shll_ic_domain", "shll_last_error", "shll_version",
]

_lib = None


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ShllError(EDEVICE, "HIP library not built: %s (run __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i64, i32, c = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_int32, ctypes.c_int
    P = ctypes.POINTER(Problem)
    pp = ctypes.POINTER(vp)
    sig = {
        "shll_init": (c, [P, dp, pp]),
        "shll_step": (c, [vp, i64]),
        "shll_run_to_time": (c, [vp, ctypes.c_float, ctypes.POINTER(i64)]),
        "shll_steps_to_time": (i64, [P, ctypes.c_float]),
        "shll_get_primitives": (c, [vp, dp]),
        "shll_get_conserved": (c, [vp, dp]),
        "shll_state_bytes": (ctypes.c_size_t, [vp]),
        "shll_get_state_raw": (c, [vp, vp, ctypes.c_size_t]),
        "shll_set_state_raw": (c, [vp, vp, ctypes.c_size_t]),
        "shll_write_results": (c, [vp, ctypes.c_char_p]),
        "shll_write_results_full": (c, [vp, ctypes.c_char_p]),
        "shll_conserved_sums": (c, [vp, dp]),
        "shll_conserved_sums_noalloc": (c, [vp, dp]),
        "shll_stage_fluxes": (c, [vp, dp, dp, dp, dp, dp]),
        "shll_check_state": (c, [vp, ctypes.POINTER(i64)]),
        "shll_compare": (c, [vp, vp, ctypes.POINTER(Errors)]),
        "shll_compare_host": (c, [vp, dp, ctypes.POINTER(Errors)]),
        "shll_free": (None, [vp]),
        "shll_set_launch": (c, [vp, i32, i32, i32, i32]),
        "shll_get_launch": (c, [vp, ctypes.POINTER(i32)]),
        "shll_autotune": (c, [vp, i32]),
        "shll_set_timing": (c, [vp, i32]),
        "shll_last_timing": (c, [vp, dp, ctypes.POINTER(i64), dp]),
        "shll_batch_init": (c, [P, ctypes.POINTER(dp), ctypes.POINTER(i64), i32, pp]),
        "shll_batch_advance": (c, [vp, i64, ctypes.POINTER(i64)]),
        "shll_batch_run": (c, [vp]),
        "shll_batch_get_primitives": (c, [vp, i32, dp]),
        "shll_batch_get_conserved": (c, [vp, i32, dp]),
        "shll_batch_conserved_sums": (c, [vp, dp]),
        "shll_batch_set_launch": (c, [vp, i32, i32, i32, i32]),
        "shll_batch_autotune": (c, [vp, i32]),
        "shll_batch_get_launch": (c, [vp, ctypes.POINTER(i32)]),
        "shll_batch_check_state": (c, [vp, ctypes.POINTER(i64), ctypes.POINTER(i32)]),
        "shll_batch_compare": (c, [vp, i32, vp, i32, ctypes.POINTER(Errors)]),
        "shll_batch_set_timing": (c, [vp, i32]),
        "shll_batch_last_timing": (c, [vp, dp, ctypes.POINTER(i64), dp]),
        "shll_batch_total_cell_updates": (i64, [vp]),
        "shll_batch_reset": (c, [vp]),
        "shll_batch_free": (None, [vp]),
        "shll_ic_tube": (None, [i64] + [ctypes.c_double] * 7 + [dp]),
        "shll_ic_implosion": (None, [i64, i64, dp]),
        "shll_ic_four_shocks": (None, [i64, i64, dp]),
        "shll_ic_four_contacts": (None, [i64, i64, dp]),
        "shll_ic_circle": (None, [i64, i64] + [ctypes.c_double] * 9 + [dp]),
        "shll_gen_tube": (None, [ctypes.c_uint64, i64, i64, i64, ctypes.POINTER(TubeSpec)]),
        "shll_gen_domain": (None, [ctypes.c_uint64, i64, i64, i64, ctypes.POINTER(DomainSpec)]),
        "shll_ic_domain": (None, [ctypes.POINTER(DomainSpec), dp]),
        "shll_last_error": (ctypes.c_char_p, []),
        "shll_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise ShllError(rc, lib().shll_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def problem(dim, order, nx, ny=1, precision=F64, bc=None, lx=1.0, ly=1.0, limiter=LIM_MINMOD, theta=1.25):
    """The reference pairs first order with reflective walls and second order with outflow;
    limiter="mc" with theta = 1.25 is base-omp/2nd_order_base_shll.c:44,317-325."""
    if isinstance(precision, str):
        precision = PREC_NAMES[precision]
    if isinstance(limiter, str):
        limiter = LIMITER_NAMES[limiter]
    if bc is None:
        bc = BC_REFLECT if order == 1 else BC_OUTFLOW
    return Problem(dim, order, bc, precision, nx, ny if dim == 2 else 1, lx, ly, limiter, 0, theta)


def nvar(pb):
    return 3 if pb.dim == 1 else 4


def steps_to_time(pb, t_end):
    return int(lib().shll_steps_to_time(ctypes.byref(pb), ctypes.c_float(t_end)))


class Solver:
    """One instance on the GPU (shll_init ... shll_free)."""

    def __init__(self, pb, prim):
        self.pb = pb
        self.nv = nvar(pb)
        self.ncell = pb.nx * pb.ny
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(self.nv, self.ncell)
        h = ctypes.c_void_p()
        _check(lib().shll_init(ctypes.byref(pb), _dp(prim), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().shll_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n):
        _check(lib().shll_step(self.h, int(n)))

    def run_to_time(self, t_end):
        n = ctypes.c_int64(0)
        _check(lib().shll_run_to_time(self.h, ctypes.c_float(t_end), ctypes.byref(n)))
        return n.value

    def primitives(self):
        out = np.empty((self.nv, self.ncell))
        _check(lib().shll_get_primitives(self.h, _dp(out)))
        return out

    def conserved(self):
        out = np.empty((self.nv, self.ncell))
        _check(lib().shll_get_conserved(self.h, _dp(out)))
        return out

    def state_raw(self):
        """F64 / F64F: (nv, N) float64; REF32: (nv, N) float32; C32 / C32K: (hi, lo) float32 pair."""
        nb = lib().shll_state_bytes(self.h)
        buf = np.empty(nb, dtype=np.uint8)
        _check(lib().shll_get_state_raw(self.h, buf.ctypes.data_as(ctypes.c_void_p), nb))
        if self.pb.precision in (F64, F64F):
            return buf.view(np.float64).reshape(self.nv, self.ncell)
        f = buf.view(np.float32)
        if self.pb.precision == REF32:
            return f.reshape(self.nv, self.ncell)
        f = f.reshape(2, self.nv, self.ncell)
        return f[0], f[1]

    def set_state_raw(self, raw):
        """Restart from a checkpoint: raw as returned by state_raw() (F64 / REF32 array, or
        the (hi, lo) pair of the compensated forms)."""
        if isinstance(raw, tuple):
            raw = np.concatenate([np.ascontiguousarray(r).ravel() for r in raw])
        buf = np.ascontiguousarray(raw).view(np.uint8).ravel()
        _check(lib().shll_set_state_raw(self.h, buf.ctypes.data_as(ctypes.c_void_p), buf.size))

    def write_results(self, path):
        _check(lib().shll_write_results(self.h, str(path).encode()))

    def write_results_full(self, path):
        """Primitives and conserved variables (the ten-column 2D format of omp-vectors)."""
        _check(lib().shll_write_results_full(self.h, str(path).encode()))

    def conserved_sums(self):
        out = np.zeros(4)
        _check(lib().shll_conserved_sums(self.h, _dp(out)))
        return out[: self.nv]

    def stage_fluxes(self):
        """Sound speed and split fluxes of the current state, as the step kernels compute
        them: {"a": (n,), "fp", "fm" (and "hp", "hm" in 2D): (nvar, n)}."""
        n = self.pb.nx * self.pb.ny
        out = {"a": np.zeros(n), "fp": np.zeros((self.nv, n)), "fm": np.zeros((self.nv, n))}
        if self.pb.dim == 2:
            out["hp"], out["hm"] = np.zeros((self.nv, n)), np.zeros((self.nv, n))
        _check(lib().shll_stage_fluxes(self.h, _dp(out["a"]), _dp(out["fp"]), _dp(out["fm"]),
                                       _dp(out["hp"]) if "hp" in out else None,
                                       _dp(out["hm"]) if "hm" in out else None))
        return out

    def check_state(self):
        bad = ctypes.c_int64(0)
        rc = lib().shll_check_state(self.h, ctypes.byref(bad))
        if rc not in (OK, ENONFINITE):
            _check(rc)
        return int(bad.value)

    def compare(self, ref):
        e = Errors()
        if isinstance(ref, Solver):
            _check(lib().shll_compare(self.h, ref.h, ctypes.byref(e)))
        else:
            ref = np.ascontiguousarray(ref, dtype=np.float64).reshape(self.nv, self.ncell)
            _check(lib().shll_compare_host(self.h, _dp(ref), ctypes.byref(e)))
        return e.as_dict()

    def set_launch(self, cells_per_lane=None, steps_per_launch=8, waves_per_block=4, graph_steps=128):
        """cells_per_lane: 1D cells per lane (default 2); 2D owned rows per work item (default 1:
        16, 32 or 64 rows for 1, 2 or more steps per launch)."""
        if cells_per_lane is None:
            cells_per_lane = 2 if self.pb.dim == 1 else 1
        _check(lib().shll_set_launch(self.h, cells_per_lane, steps_per_launch, waves_per_block, graph_steps))

    def get_launch(self):
        out = (ctypes.c_int32 * 4)()
        _check(lib().shll_get_launch(self.h, out))
        return tuple(out)

    def autotune(self, trial_steps):
        _check(lib().shll_autotune(self.h, int(trial_steps)))

    def set_timing(self, on=True):
        _check(lib().shll_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        avg, tot, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _check(lib().shll_last_timing(self.h, ctypes.byref(avg), ctypes.byref(n), ctypes.byref(tot)))
        return avg.value, n.value, tot.value


class Batch:
    """Many independent instances advanced together (shll_batch_*)."""

    def __init__(self, pbs, prims, nsteps):
        n = len(pbs)
        self.pbs = list(pbs)
        arr = (Problem * n)(*pbs)
        self._prims = [np.ascontiguousarray(p, dtype=np.float64) for p in prims]
        pp = (ctypes.POINTER(ctypes.c_double) * n)(*[_dp(p) for p in self._prims])
        ns = np.ascontiguousarray(nsteps, dtype=np.int64)
        h = ctypes.c_void_p()
        _check(lib().shll_batch_init(arr, pp, ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n,
                                     ctypes.byref(h)))
        self.h = h
        self._prims = None
        self.nsteps = ns

    def close(self):
        if getattr(self, "h", None):
            lib().shll_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, n):
        rem = ctypes.c_int64(0)
        _check(lib().shll_batch_advance(self.h, int(n), ctypes.byref(rem)))
        return rem.value

    def run(self):
        _check(lib().shll_batch_run(self.h))

    def reset(self):
        _check(lib().shll_batch_reset(self.h))

    def primitives(self, k):
        pb = self.pbs[k]
        out = np.empty((nvar(pb), pb.nx * pb.ny))
        _check(lib().shll_batch_get_primitives(self.h, k, _dp(out)))
        return out

    def conserved(self, k):
        pb = self.pbs[k]
        out = np.empty((nvar(pb), pb.nx * pb.ny))
        _check(lib().shll_batch_get_conserved(self.h, k, _dp(out)))
        return out

    def conserved_sums(self):
        out = np.zeros(len(self.pbs) * 4)
        _check(lib().shll_batch_conserved_sums(self.h, _dp(out)))
        return out.reshape(-1, 4)

    def set_launch(self, cells_per_lane=None, steps_per_launch=16, waves_per_block=4, graph_steps=128):
        """As Solver.set_launch (1D default 2 cells per lane, 2D the default rows per item)."""
        if cells_per_lane is None:
            cells_per_lane = 2 if self.pbs[0].dim == 1 else 1
        _check(lib().shll_batch_set_launch(self.h, cells_per_lane, steps_per_launch, waves_per_block, graph_steps))

    def autotune(self, trial_steps):
        _check(lib().shll_batch_autotune(self.h, int(trial_steps)))

    def get_launch(self):
        out = (ctypes.c_int32 * 4)()
        _check(lib().shll_batch_get_launch(self.h, out))
        return tuple(out)

    def check_state(self):
        """Per-instance count of non-finite / non-positive cells."""
        bad = np.zeros(len(self.pbs), dtype=np.int64)
        nb = ctypes.c_int32(0)
        _check(lib().shll_batch_check_state(self.h, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                            ctypes.byref(nb)))
        return bad

    def compare(self, k, ref, kref=None):
        e = Errors()
        _check(lib().shll_batch_compare(self.h, k, ref.h, k if kref is None else kref, ctypes.byref(e)))
        return e.as_dict()

    def set_timing(self, on=True):
        _check(lib().shll_batch_set_timing(self.h, 1 if on else 0))

    def last_timing(self):
        avg, tot, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _check(lib().shll_batch_last_timing(self.h, ctypes.byref(avg), ctypes.byref(n), ctypes.byref(tot)))
        return avg.value, n.value, tot.value

    def total_cell_updates(self):
        return int(lib().shll_batch_total_cell_updates(self.h))


# ---- initial data (host generators of the library) ----
def ic_tube(n, x0=0.5, rhoL=10.0, uL=0.0, TL=1.0, rhoR=1.0, uR=0.0, TR=1.0):
    out = np.zeros((3, n))
    lib().shll_ic_tube(n, x0, rhoL, uL, TL, rhoR, uR, TR, _dp(out))
    return out


def ic_implosion(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().shll_ic_implosion(nx, ny, _dp(out))
    return out


def ic_four_shocks(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().shll_ic_four_shocks(nx, ny, _dp(out))
    return out


def ic_four_contacts(nx, ny):
    out = np.zeros((4, nx * ny))
    lib().shll_ic_four_contacts(nx, ny, _dp(out))
    return out


def ic_circle(nx, ny, lx=1.0, ly=1.0, cx=0.5, cy=0.5, r=0.25, rho_in=10.0, T_in=1.0, rho_out=1.0, T_out=1.0):
    out = np.zeros((4, nx * ny))
    lib().shll_ic_circle(nx, ny, lx, ly, cx, cy, r, rho_in, T_in, rho_out, T_out, _dp(out))
    return out


def gen_tube(seed, k, nmin=256, nmax=65536):
    s = TubeSpec()
    lib().shll_gen_tube(seed, k, nmin, nmax, ctypes.byref(s))
    return s


def gen_domain(seed, k, smin=256, smax=2048):
    s = DomainSpec()
    lib().shll_gen_domain(seed, k, smin, smax, ctypes.byref(s))
    return s


def ic_domain(spec):
    out = np.zeros((4, spec.nx * spec.ny))
    lib().shll_ic_domain(ctypes.byref(spec), _dp(out))
    return out


def tube_batch(seed, count, nmin=256, nmax=65536, precision=C32, t_end=0.2, order=2):
    """Problems, primitives and step counts of a seeded batch of shock tubes."""
    pbs, prims, steps = [], [], []
    for k in range(count):
        s = gen_tube(seed, k, nmin, nmax)
        pb = problem(1, order, s.n, precision=precision)
        pbs.append(pb)
        prims.append(ic_tube(s.n, s.x0, s.rhoL, 0.0, s.TL, s.rhoR, 0.0, s.TR))
        steps.append(steps_to_time(pb, t_end))
    return pbs, prims, steps


def domain_batch(seed, count, smin=256, smax=2048, precision=C32, order=1):
    pbs, prims = [], []
    for k in range(count):
        d = gen_domain(seed, k, smin, smax)
        pbs.append(problem(2, order, d.nx, d.ny, precision=precision, lx=d.lx, ly=d.ly))
        prims.append(ic_domain(d))
    return pbs, prims


def read_results(path, nx, ny=1):
    """Read a results.dat written by write_results / write_results_full (or by the
    reference's Save_Results) into named (nx, ny) arrays -- the reading half of the
    reference's plotting script (base-omp/fetch_local.py:13-27), without S3 or a plot."""
    data = np.loadtxt(path, ndmin=2)
    dim = 1 if ny == 1 else 2
    names = {(1, 4): ["x", "rho", "u", "T"],
             (1, 7): ["x", "rho", "u", "T", "u0", "u1", "u2"],
             (2, 6): ["x", "y", "rho", "u", "v", "T"],
             (2, 10): ["x", "y", "rho", "u", "v", "T", "u0", "u1", "u2", "u3"]}.get((dim, data.shape[1]))
    if names is None or data.shape[0] != nx * ny:
        raise ValueError("%s: %d rows x %d columns does not match a %dD results file of %d cells"
                         % (path, data.shape[0], data.shape[1], dim, nx * ny))
    shape = (nx,) if dim == 1 else (nx, ny)
    return {n: data[:, k].reshape(shape) for k, n in enumerate(names)}
