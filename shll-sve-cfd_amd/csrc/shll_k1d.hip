// shll_k1d.hip -- fused 1D SHLL step kernels (first and second order) for gfx950.
//
// One wavefront owns a contiguous segment of 64*C cells, C consecutive cells per
// lane, held in registers for the whole launch.  A step is
//   primitive recovery + SHLL split fluxes of the lane's C cells   (Compute_P_from_U,
//                                                                   Compute_F_from_P)
//   neighbour fluxes of the two adjacent lanes by DPP wave shifts   (the gather of
//                                                                   Update_U_from_F)
//   boundary cells and the conservative update in registers         (Update_U_from_F)
// so a launch of k steps reads the state from HBM once and writes it once; the
// segment's outer order*k cells on each side are recomputed halo (temporal
// blocking without LDS and without barriers: wavefronts never wait on each other).
//
// References: base_shll.c:87-177 (first order, reflective), 2nd_order_base_shll.c
// :203-459 (second order, outflow; its x sweep on the 3-variable system is the 1D
// second-order scheme), base-omp/2nd_order_base_shll.c:317-325 (the MC limiter).
#include "shll_flux.h"

namespace shll {

// Wall modes of a wavefront's segment (a wave-uniform choice, see k_step1d):
//   WM_NONE   the segment is at least 3 cells from both ends: no boundary rule applies
//   WM_GHOST  ghost cells -1 and n (ghost_cells) carry the walls
//   WM_RULES  ghost cells plus the explicit zero-slope rule (reflective second order)
enum { WM_NONE = 0, WM_GHOST = 1, WM_RULES = 2 };

// slope of a split-flux window at position p (window index p+2) of global cell gg;
// with WM_RULES zero on the two boundary cells and beyond (2nd_order_base_shll.c:227-256)
template <class A, int WM, int LIM, class RT, int NW>
__device__ __forceinline__ RT slope_at(const RT (&W)[NW], int p, int gg, int n, RT theta) {
    RT s = limited_slope<A, LIM>(W[p + 1], W[p + 2], W[p + 3], theta);
    if constexpr (WM != WM_RULES) return s;
    return (gg <= 0 || gg >= n - 1) ? (RT)0 : s;
}

// window of a per-lane array q[C]: W[p+2] = value at lane position p, p in [-2, C+1]
template <class RT, int C>
__device__ __forceinline__ void make_window(const RT (&q)[C], RT (&W)[C + 4]) {
#pragma unroll
    for (int m = 0; m < C; ++m) W[m + 2] = q[m];
    if constexpr (C >= 2) {
        W[0] = from_left(q[C - 2]);
        W[1] = from_left(q[C - 1]);
        W[C + 2] = from_right(q[0]);
        W[C + 3] = from_right(q[1]);
    } else {
        W[1] = from_left(q[0]);
        W[0] = from_left(W[1]);
        W[3] = from_right(q[0]);
        W[4] = from_right(W[3]);
    }
}

// Walls as ghost cells: before each step the state of cell -1 becomes that of cell 0
// and the state of cell n that of cell n-1 -- copied for outflow walls, mirrored
// (momentum negated) for reflective ones.  Every operation of the split flux is odd
// or even in the velocity, so the ghosts' fluxes are bitwise the values the
// reference's boundary rules use for the missing neighbour: F+(-1) = F+(0) and
// F-(n) = F-(n-1) for outflow (2nd_order_base_shll.c:212-256), F+(-1) = s*F-(0) and
// F-(n) = s*F+(n-1) with s = (-1, +1, -1) for reflection (base_shll.c:93-110).  With
// copied ghosts every boundary slope also comes out zero through the limiter (one of
// its differences is exactly 0, and so is then the MC form's second argument), as the
// reference sets them; reflective second order
// needs the explicit rule (WM_RULES).  Cells beyond the ghosts only feed the ghosts'
// own updates, which are replaced again before the next step.
template <int C, class RT>
__device__ __forceinline__ void ghost_cells(RT (&q)[3][C], int gl, int n, bool refl) {
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const RT next = from_right(q[v][0]);     // cell gl+C
        const RT prev = from_left(q[v][C - 1]);  // cell gl-1
        RT out[C];
#pragma unroll
        for (int m = 0; m < C; ++m) {
            RT right = m + 1 < C ? q[v][m + 1] : next;  // cell g+1
            RT left = m > 0 ? q[v][m - 1] : prev;       // cell g-1
            if (v == 1) {
                right = refl ? -right : right;
                left = refl ? -left : left;
            }
            const int g = gl + m;
            out[m] = g == -1 ? right : (g == n ? left : q[v][m]);
        }
#pragma unroll
        for (int m = 0; m < C; ++m) q[v][m] = out[m];
    }
}

// ---------------------------------------------------------------------------
// F64 / REF32: the reference association, cell by cell
// ---------------------------------------------------------------------------
template <int PREC, int ORDER, int C, int WM, int LIM>
__device__ __forceinline__ void step_ref(typename Real<PREC>::T (&u)[3][C], const Inst& in, int gl) {
    using RT = typename Real<PREC>::T;
    using A = RefArith<RT>;
    const RT lam = PREC == PREC_F64 ? (RT)in.k.d_lamx : (RT)in.k.f_lamx;
    const RT th = PREC == PREC_F64 ? (RT)in.k.d_theta : (RT)in.k.f_theta;
    const int n = in.nx;
    const bool refl = in.bc == BC_REFLECT;
    if constexpr (WM != WM_NONE) ghost_cells<C>(u, gl, n, refl);
    RT fp[3][C], fm[3][C];
#pragma unroll
    for (int m = 0; m < C; ++m) {
        RT p[3], q[3], a;
        flux1<PREC, RT>(u[0][m], u[1][m], u[2][m], in, p, q, a);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            fp[v][m] = p[v];
            fm[v][m] = q[v];
        }
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        RT FP[C + 4], FM[C + 4];
        make_window<RT, C>(fp[v], FP);
        make_window<RT, C>(fm[v], FM);
#pragma unroll
        for (int m = 0; m < C; ++m) {
            const int g = gl + m;
            // FL = F+ of the left neighbour, FR = F- of the right one (ghosts at the walls)
            RT uu = u[v][m] - lam * (fp[v][m] - fm[v][m] + FM[m + 3] - FP[m + 1]);
            if constexpr (ORDER == 2) {
                // own slopes, the left neighbour's F+ slope and the right neighbour's F-
                // slope; every slope of a boundary or ghost cell is zero
                RT dp = slope_at<A, WM, LIM>(FP, m, g, n, th);
                RT dm = slope_at<A, WM, LIM>(FM, m, g, n, th);
                RT Ldf = slope_at<A, WM, LIM>(FP, m - 1, g - 1, n, th);
                RT Rdf = slope_at<A, WM, LIM>(FM, m + 1, g + 1, n, th);
                uu = (RT)((double)uu - 0.5 * (double)lam * (double)(dp + dm - Rdf - Ldf));
            }
            u[v][m] = uu;
        }
    }
}

// ---------------------------------------------------------------------------
// C32: float fluxes of the hi word, face terms t = lam*G added with TwoSum
// ---------------------------------------------------------------------------
// KAHAN: the face terms t[m] - t[m+1] enter each cell by one Kahan step (lo holds
// the part of the value not yet in hi) instead of the exact TwoSum chain
template <int ORDER, int C, int WM, bool KAHAN, int LIM>
__device__ __forceinline__ void step_c32(float (&hi)[3][C], float (&lo)[3][C], const Inst& in, int gl) {
    using A = C32Arith;
    const float lam = in.k.c_lamx, hlam = 0.5f * lam, th = in.k.c_theta;
    const int n = in.nx;
    if constexpr (WM != WM_NONE) ghost_cells<C>(hi, gl, n, in.bc == BC_REFLECT);
    float fp[3][C], fm[3][C];
#pragma unroll
    for (int m = 0; m < C; ++m) {
        float p[3], q[3], a;
        flux1<PREC_C32, float>(hi[0][m], hi[1][m], hi[2][m], in, p, q, a);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            fp[v][m] = p[v];
            fm[v][m] = q[v];
        }
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float FP[C + 4], FM[C + 4], t[C + 1];
        make_window<float, C>(fp[v], FP);
        make_window<float, C>(fm[v], FM);
#pragma unroll
        for (int f = 0; f <= C; ++f) {
            const int gf = gl + f;  // face gf lies between cells gf-1 and gf
            const float G = FP[f + 1] + FM[f + 2];
            if constexpr (ORDER == 2) {
                const float d = slope_at<A, WM, LIM>(FP, f - 1, gf - 1, n, th) - slope_at<A, WM, LIM>(FM, f, gf, n, th);
                t[f] = __builtin_fmaf(hlam, d, lam * G);  // lam*(G + d/2)
            } else {
                t[f] = lam * G;
            }
        }
#pragma unroll
        for (int m = 0; m < C; ++m) {
            if constexpr (KAHAN) {
                const float y = (t[m] - t[m + 1]) + lo[v][m];
                const float s = hi[v][m] + y;
                lo[v][m] = y - (s - hi[v][m]);
                hi[v][m] = s;
            } else {
                float a = hi[v][m], b = lo[v][m], e;
                two_sum(a, t[m], a, e);
                b = b + e;
                two_sum(a, -t[m + 1], a, e);
                b = b + e;
                fast_two_sum(a, b, hi[v][m], lo[v][m]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// ns steps of one segment from registers, then the owned cells back to HBM.  Each
// wall mode gets its own copy of loop and store, so the loop-carried state never
// has to be moved between the copies' registers.
// ---------------------------------------------------------------------------
template <int ORDER, int C, int WM, bool KAHAN, int LIM>
__device__ __forceinline__ void advance_c32(float (&hi)[3][C], float (&lo)[3][C], const Inst& in, const Item1& it,
                                            int gl, int ns, const StepArgs& a) {
    for (int s = 0; s < ns; ++s) step_c32<ORDER, C, WM, KAHAN, LIM>(hi, lo, in, gl);
    float* __restrict__ oh = (float*)a.out0 + in.off;
    float* __restrict__ ol = (float*)a.out1 + in.off;
    const int64_t P = a.plane;
#pragma unroll
    for (int m = 0; m < C; ++m) {
        const int g = gl + m;
        if (g >= it.x0 && g < it.x0 + it.w) {
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                oh[v * P + g] = hi[v][m];
                ol[v * P + g] = lo[v][m];
            }
        }
    }
}

template <int PREC, int ORDER, int C, int WM, int LIM>
__device__ __forceinline__ void advance_ref(typename Real<PREC>::T (&u)[3][C], const Inst& in, const Item1& it,
                                            int gl, int ns, const StepArgs& a) {
    using RT = typename Real<PREC>::T;
    for (int s = 0; s < ns; ++s) step_ref<PREC, ORDER, C, WM, LIM>(u, in, gl);
    RT* __restrict__ ou = (RT*)a.out0 + in.off;
    const int64_t P = a.plane;
#pragma unroll
    for (int m = 0; m < C; ++m) {
        const int g = gl + m;
        if (g >= it.x0 && g < it.x0 + it.w) {
#pragma unroll
            for (int v = 0; v < 3; ++v) ou[v * P + g] = u[v][m];
        }
    }
}

// ---------------------------------------------------------------------------
// the launch: one wavefront per Item1 (LIM: the second-order limiter)
// ---------------------------------------------------------------------------
template <int PREC, int ORDER, int C, int WPB, int LIM>
__global__ __launch_bounds__(64 * WPB) void k_step1d(StepArgs a) {
    using RT = typename Real<PREC>::T;
    const int wave = (int)blockIdx.x * WPB + uniform((int)threadIdx.x >> 6);
    const LaunchView lv = launch_view(a);
    if (wave >= lv.n_items) return;
    const int lane = (int)threadIdx.x & 63;
    const Item1 it = ((const Item1*)a.items)[wave];
    const Inst in = a.inst[it.inst];
    const int64_t left = in.nsteps - lv.step_base;
    const int ns = left < (int64_t)a.ksteps ? (int)left : a.ksteps;
    if (ns <= 0) return;
    const int n = in.nx;
    const int gl = it.g0 + lane * C;
    const int64_t P = a.plane;
    // wall rules can only apply if the segment reaches within 3 cells of an end
    const bool walls = it.g0 <= 2 || it.g0 + 64 * C >= n - 3;
    const int wm = !walls ? WM_NONE : (ORDER == 2 && in.bc == BC_REFLECT ? WM_RULES : WM_GHOST);
    constexpr int WR = ORDER == 2 ? WM_RULES : WM_GHOST;

    if constexpr (PREC == PREC_C32 || PREC == PREC_C32K) {
        constexpr bool K = PREC == PREC_C32K;
        const float* __restrict__ ih = (const float*)a.in0 + in.off;
        const float* __restrict__ il = (const float*)a.in1 + in.off;
        float hi[3][C], lo[3][C];
#pragma unroll
        for (int m = 0; m < C; ++m) {
            const int g = gl + m;
            const bool ok = g >= 0 && g < n;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                hi[v][m] = ok ? ih[v * P + g] : (v == 0 ? 1.0f : (v == 1 ? 0.0f : 2.5f));
                lo[v][m] = ok ? il[v * P + g] : 0.0f;
            }
        }
        if (wm == WM_NONE) advance_c32<ORDER, C, WM_NONE, K, LIM>(hi, lo, in, it, gl, ns, a);
        else if (ORDER == 1 || wm == WM_GHOST) advance_c32<ORDER, C, WM_GHOST, K, LIM>(hi, lo, in, it, gl, ns, a);
        else advance_c32<ORDER, C, WR, K, LIM>(hi, lo, in, it, gl, ns, a);
    } else {
        const RT* __restrict__ iu = (const RT*)a.in0 + in.off;
        RT u[3][C];
#pragma unroll
        for (int m = 0; m < C; ++m) {
            const int g = gl + m;
            const bool ok = g >= 0 && g < n;
#pragma unroll
            for (int v = 0; v < 3; ++v) u[v][m] = ok ? iu[v * P + g] : (RT)(v == 0 ? 1.0 : (v == 1 ? 0.0 : 2.5));
        }
        if (wm == WM_NONE) advance_ref<PREC, ORDER, C, WM_NONE, LIM>(u, in, it, gl, ns, a);
        else if (ORDER == 1 || wm == WM_GHOST) advance_ref<PREC, ORDER, C, WM_GHOST, LIM>(u, in, it, gl, ns, a);
        else advance_ref<PREC, ORDER, C, WR, LIM>(u, in, it, gl, ns, a);
    }
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <int PREC, int ORDER, int C, int LIM>
static hipError_t launch_c(int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    if (wpb == 1)
        hipLaunchKernelGGL((k_step1d<PREC, ORDER, C, 1, LIM>), dim3(nblocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_step1d<PREC, ORDER, C, 4, LIM>), dim3(nblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int PREC, int ORDER, int LIM>
static hipError_t launch_o(int C, int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    switch (C) {
    case 1: return launch_c<PREC, ORDER, 1, LIM>(wpb, nblocks, s, a);
    case 2: return launch_c<PREC, ORDER, 2, LIM>(wpb, nblocks, s, a);
    case 4: return launch_c<PREC, ORDER, 4, LIM>(wpb, nblocks, s, a);
    case 8: return launch_c<PREC, ORDER, 8, LIM>(wpb, nblocks, s, a);
    default: return hipErrorInvalidValue;
    }
}

// first order; second order with minmod; second order with the MC limiter
template <int PREC>
static hipError_t launch_p(int order, int lim, int C, int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    if (order == 1) return lim == LIM_MINMOD ? launch_o<PREC, 1, LIM_MINMOD>(C, wpb, nblocks, s, a) : hipErrorInvalidValue;
    if (order != 2) return hipErrorInvalidValue;
    if (lim == LIM_MINMOD) return launch_o<PREC, 2, LIM_MINMOD>(C, wpb, nblocks, s, a);
    if (lim == LIM_MC) return launch_o<PREC, 2, LIM_MC>(C, wpb, nblocks, s, a);
    return hipErrorInvalidValue;
}

hipError_t launch_step1d(int prec, int order, int lim, int C, int wpb, int n_items, hipStream_t s,
                         const StepArgs& a) {
    const int nblocks = (n_items + wpb - 1) / wpb;
    if (nblocks <= 0) return hipSuccess;
    switch (prec) {
    case PREC_F64: return launch_p<PREC_F64>(order, lim, C, wpb, nblocks, s, a);
    case PREC_REF32: return launch_p<PREC_REF32>(order, lim, C, wpb, nblocks, s, a);
    case PREC_C32: return launch_p<PREC_C32>(order, lim, C, wpb, nblocks, s, a);
    case PREC_C32K: return launch_p<PREC_C32K>(order, lim, C, wpb, nblocks, s, a);
    default: return hipErrorInvalidValue;
    }
}

bool step1d_supported(int C, int wpb) { return (C == 1 || C == 2 || C == 4 || C == 8) && (wpb == 1 || wpb == 4); }

}  // namespace shll
