// shll_types.h -- tables and launch arguments shared by the host API and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shll {

enum Prec : int { PREC_F64 = 0, PREC_REF32 = 1, PREC_C32 = 2, PREC_C32K = 3, PREC_F64F = 4 };
enum Bc : int { BC_REFLECT = 0, BC_OUTFLOW = 1 };
// second-order slope limiters: minmod of the one-sided differences
// (2nd_order_base_shll.c:268-276) or the MC form minmod(half central difference,
// theta * minmod(one-sided differences)) (base-omp/2nd_order_base_shll.c:317-325)
enum Lim : int { LIM_MINMOD = 0, LIM_MC = 1 };

// Per-instance scheme constants, computed on the host (never folded on the device):
//   F64   : d_*  (gamma = 1.4, CV = R/(gamma-1), lam = DT/DX)
//   REF32 : f_*  (the reference's float constants, base_shll.c:17-22,83-84)
//   C32   : c_*  (the F64 constants rounded once to float; rcv = 1/CV)
//   F64F  : d_*, and d_rcv = 1/CV in double
// theta: the MC limiter's weight (the reference's float alpha = 1.25 in REF32)
struct Consts {
    double d_cv, d_gR, d_R, d_lamx, d_lamy, d_theta, d_rcv;
    float f_cv, f_gR, f_R, f_lamx, f_lamy, f_theta;
    float c_rcv, c_gR, c_R, c_lamx, c_lamy, c_theta;
};

// One instance (tube or 2D domain) of a launch.
struct Inst {
    int64_t off;     // first cell of the instance in every variable plane
    int64_t nsteps;  // steps this instance takes in total (batch); huge for single runs
    int32_t nx;      // cells in x
    int32_t ny;      // cells in y (1 in 1D)
    int32_t bc;      // Bc
    int32_t ke;      // kinetic-energy association of the reference file (0: base_shll.c, 1: 2D files)
    Consts k;
};

// One wavefront's share of a 1D launch: it loads cells [g0, g0 + 64*C) of its
// instance (cells outside [0, nx) get a benign state), advances them, and writes
// back the owned cells [x0, x0 + w), all of which are still exact after the
// launch's steps (g0 <= x0 - order*steps or g0 <= 0, and the same on the right).
struct Item1 {
    int32_t inst;
    int32_t g0;
    int32_t x0;
    int32_t w;
};

// One wavefront's share of a 2D launch: a strip of columns j in [j0, j0 + 64)
// (owned [jx0, jx0 + jw)) streamed along rows i in [i0, i0 + ilen) (owned
// [ix0, ix0 + iw)).
struct Item2 {
    int32_t inst;
    int32_t j0, jx0, jw;
    int32_t i0, ix0, iw, ilen;
};

struct StepArgs {
    const Inst* inst;
    const void* items;
    int32_t n_items;
    int32_t ksteps;      // steps of this launch (an instance does min(ksteps, nsteps - step_base))
    int64_t step_base;   // steps already taken before this launch
    int64_t plane;       // cells per variable plane
    const void* in0;     // state read: F64 double*, REF32 float*, C32 hi float*
    const void* in1;     //             C32 lo float*
    void* out0;          // state written
    void* out1;
    // Device-driven launches (constant graph nodes): when jctl is set, it points at the
    // device counter {j0, anchor, jmax}; this launch's index is j = j0 + jofs, it covers
    // n_items = active[j] items (0 if j >= jmax) and step_base = anchor + j * ksteps,
    // overriding the two host-side fields above.  active[] lists the item counts of the
    // launches of a regular grid of ksteps-step launches that starts at step `anchor`.
    const int64_t* jctl;
    const int32_t* active;
    int32_t jofs;
};

// launch index, item count and step base of a step launch
struct LaunchView {
    int32_t n_items;
    int64_t step_base;
};

__host__ __device__ inline LaunchView launch_view(const StepArgs& a) {
    LaunchView v{a.n_items, a.step_base};
    if (a.jctl) {
        const int64_t j = a.jctl[0] + a.jofs;
        v.n_items = j < a.jctl[2] ? a.active[j] : 0;
        v.step_base = a.jctl[1] + j * (int64_t)a.ksteps;
    }
    return v;
}

// XCD-aware work order.  The GPU hands workgroup b of a launch to XCD b % kXcds, and each
// XCD has its own L2.  With a grid of a multiple of kXcds blocks (the launch dispatchers
// round it up) and q = ceil(nb / kXcds), block b takes work w = (b % kXcds) * q + b / kXcds,
// so XCD x works on the contiguous run [x*q, (x+1)*q): neighbouring tiles, whose halos
// overlap, share an L2.  Every w < nb is taken by exactly one block b < kXcds * q.
// nb: blocks with work; -1: this block has none.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_work(int b, int nb) {
    const int q = (nb + kXcds - 1) / kXcds;
    const int r = b / kXcds;
    const int w = (b % kXcds) * q + r;
    return (r < q && w < nb) ? w : -1;
}
__host__ __device__ inline int xcd_grid(int nb) { return (nb + kXcds - 1) / kXcds * kXcds; }

}  // namespace shll
