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
// With 4 wavefronts per block and at least 2 cells per lane the block's wavefronts
// instead hold consecutive segments of one item and swap their edge fluxes through LDS
// once a step (one barrier), so only the block's ends carry halo.
//
// Precisions: F64 / REF32 in the reference's association (cell by cell), C32 / C32K and
// F64F in face-term form (one term per face, shared by both cells).
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

// window of a per-lane array q[C]: W[p+2] = value at lane position p, p in [-2, C+1].
// The segment's end lanes get don't-care values for their missing neighbours (recomputed
// halo): with ZERO, zeros by shifts without an "old" operand, kept as register moves (an
// empty asm statement stops their folding into the instructions that use them); without,
// the lane's own values.  Both only change the instruction schedule: the exact TwoSum
// form measured 7% faster with ZERO, the Kahan form 4% slower.
template <class RT, int C, bool ZERO>
__device__ __forceinline__ void make_window(const RT (&q)[C], RT (&W)[C + 4]) {
#pragma unroll
    for (int m = 0; m < C; ++m) W[m + 2] = q[m];
    if constexpr (C >= 2 && ZERO) {
        W[0] = from_left_z(q[C - 2]);
        W[1] = from_left_z(q[C - 1]);
        W[C + 2] = from_right_z(q[0]);
        W[C + 3] = from_right_z(q[1]);
        asm volatile("" : "+v"(W[0]), "+v"(W[1]), "+v"(W[C + 2]), "+v"(W[C + 3]));
    } else if constexpr (C >= 2) {
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

// Cooperative blocks: the XW wavefronts of a block hold consecutive segments of one item.
// Each step they swap the split fluxes of their two edge cells on each side (every
// variable, F+ and F-) through LDS, so a segment edge inside the block sees its true
// neighbours and only the block's two ends carry recomputed halo.  XW = 1: independent
// wavefronts.
//
// LDS, per step parity and wavefront, three slots of 12 values (variable v: F+ of two
// cells, then F- of the same two, index 4v + {0, 1, 2, 3}): slot 0 the first two cells (lane 0), slot 1 the
// last two (lane 63), slot 2 a sink for the other pair of each edge lane, so the writes
// need no per-lane choice of values.  The area is double-buffered by step parity, so one
// barrier per step orders the writes before the reads and the reads before the next
// writes to the same buffer.
template <class RT>
struct CoopSlots {
    static constexpr int NV = 16 / sizeof(RT);  // values per 16-byte access
    typedef RT V __attribute__((ext_vector_type(16 / sizeof(RT))));
    static constexpr int SLOT = 12, WAVE = 3 * SLOT;
};

template <class RT, int XW>
struct Coop {
    RT* lds;   // [2 parity][XW waves][3 slots][12]
    int wib;   // wavefront in block
    int lane;
};

template <class RT, int NQ>
__device__ __forceinline__ void slot_store(RT* p, const RT (&q)[NQ]) {
    using CS = CoopSlots<RT>;
#pragma unroll
    for (int i = 0; i < CS::SLOT; i += CS::NV) {
        typename CS::V w;
#pragma unroll
        for (int j = 0; j < CS::NV; ++j) w[j] = q[i + j];
        *(typename CS::V*)(p + i) = w;
    }
}

// publish the split fluxes of the wavefront's first two and last two cells, then wait for
// the block
template <class RT, int C, int XW>
__device__ __forceinline__ void coop_put(const Coop<RT, XW>& x, int parity, const RT (&fp)[3][C],
                                         const RT (&fm)[3][C]) {
    using CS = CoopSlots<RT>;
    RT* w = x.lds + (parity * XW + x.wib) * CS::WAVE;
    if (x.lane == 0 || x.lane == 63) {
        RT first[12], last[12];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            first[4 * v] = fp[v][0];
            first[4 * v + 1] = fp[v][1];
            first[4 * v + 2] = fm[v][0];
            first[4 * v + 3] = fm[v][1];
            last[4 * v] = fp[v][C - 2];
            last[4 * v + 1] = fp[v][C - 1];
            last[4 * v + 2] = fm[v][C - 2];
            last[4 * v + 3] = fm[v][C - 1];
        }
        const bool l0 = x.lane == 0;
        slot_store(w + (l0 ? 0 : 2) * CS::SLOT, first);
        slot_store(w + (l0 ? 2 : 1) * CS::SLOT, last);
    }
    __syncthreads();
}

// what lane 0 takes for its left neighbour cells (the previous wavefront's last two) and
// lane 63 for its right ones (the next wavefront's first two), variable v: e = F+ of the
// two cells, then F-; the block's end lanes read their own wavefront's values, which only
// reach the recomputed halo
template <class RT, int XW>
__device__ __forceinline__ void coop_get(const Coop<RT, XW>& x, int parity, int v, RT (&e)[4]) {
    using CS = CoopSlots<RT>;
    const int src = x.lane == 0 ? (x.wib > 0 ? x.wib - 1 : 0) * CS::WAVE + CS::SLOT
                                : (x.wib + 1 < XW ? x.wib + 1 : x.wib) * CS::WAVE;
    const RT* p = x.lds + parity * XW * CS::WAVE + src + 4 * v;
#pragma unroll
    for (int i = 0; i < 4; i += CS::NV) {
        const typename CS::V w = *(const typename CS::V*)(p + i);
#pragma unroll
        for (int j = 0; j < CS::NV; ++j) e[i + j] = w[j];
    }
}

// make_window with lane 0's two left cells and lane 63's two right cells taken from e0, e1
template <class RT, int C>
__device__ __forceinline__ void make_window_or(const RT (&q)[C], RT (&W)[C + 4], RT e0, RT e1) {
    static_assert(C >= 2, "cooperative blocks hold at least 2 cells per lane");
#pragma unroll
    for (int m = 0; m < C; ++m) W[m + 2] = q[m];
    W[0] = from_left_or(q[C - 2], e0);
    W[1] = from_left_or(q[C - 1], e1);
    W[C + 2] = from_right_or(q[0], e0);
    W[C + 3] = from_right_or(q[1], e1);
}

// Second order with independent wavefronts (XW = 1, C >= 2): windows of one neighbour cell
// per side, W[p+1] = value at lane position p in [-1, C], and each limited slope computed
// once.  A lane limits its own C cells; the F+ slope of cell -1 and the F- slope of cell C
// are the neighbour lanes' slopes of those same cells, from the same window values (so the
// same bits), moved over by a lane shift.  sp(f) / sm(f): the F+ slope of cell f-1 and the
// F- slope of cell f, the two slopes face f uses; the interior ones are computed where they
// are used, so only the two edge slopes stay live across the faces.  Edge lanes of the
// segment get don't-care values, as the wider windows give them.
template <class A, int WM, int LIM, class RT, int C, bool ZERO>
struct SharedSlopes {
    static_assert(C >= 2, "a lane's own slopes need two cells");
    RT FP[C + 2], FM[C + 2];
    RT sP_in, sP_last, sM_first, sM_in;
    int gl, n;
    RT th;
    __device__ __forceinline__ SharedSlopes(const RT (&fp)[C], const RT (&fm)[C], int gl_, int n_, RT th_)
        : gl(gl_), n(n_), th(th_) {
#pragma unroll
        for (int m = 0; m < C; ++m) {
            FP[m + 1] = fp[m];
            FM[m + 1] = fm[m];
        }
        if constexpr (ZERO) {
            FP[0] = from_left_z(fp[C - 1]);
            FP[C + 1] = from_right_z(fp[0]);
            FM[0] = from_left_z(fm[C - 1]);
            FM[C + 1] = from_right_z(fm[0]);
            if constexpr (sizeof(RT) == 4)  // FP32: keep the shifts as moves (see make_window)
                asm volatile("" : "+v"(FP[0]), "+v"(FP[C + 1]), "+v"(FM[0]), "+v"(FM[C + 1]));
        } else {
            FP[0] = from_left(fp[C - 1]);
            FP[C + 1] = from_right(fp[0]);
            FM[0] = from_left(fm[C - 1]);
            FM[C + 1] = from_right(fm[0]);
        }
        sP_last = slope_at<A, WM, LIM>(FP, C - 2, gl + C - 1, n, th);  // F+ slope of cell C-1
        sM_first = slope_at<A, WM, LIM>(FM, -1, gl, n, th);            // F- slope of cell 0
        if constexpr (ZERO) {
            sP_in = from_left_z(sP_last);
            sM_in = from_right_z(sM_first);
        } else {
            sP_in = from_left(sP_last);
            sM_in = from_right(sM_first);
        }
    }
    __device__ __forceinline__ RT sp(int f) const {
        return f == 0 ? sP_in : f == C ? sP_last : slope_at<A, WM, LIM>(FP, f - 2, gl + f - 1, n, th);
    }
    __device__ __forceinline__ RT sm(int f) const {
        return f == C ? sM_in : f == 0 ? sM_first : slope_at<A, WM, LIM>(FM, f - 1, gl + f, n, th);
    }
};

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
template <int PREC, int ORDER, int C, int WM, int LIM, int XW>
__device__ __forceinline__ void step_ref(typename Real<PREC>::T (&u)[3][C], const Inst& in, int gl,
                                         const Coop<typename Real<PREC>::T, XW>& x, int parity) {
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
    if constexpr (ORDER == 2 && XW == 1 && C >= 2) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            // FP64: zero-filled edge lanes, no copies of the shifted values' old contents
            const SharedSlopes<A, WM, LIM, RT, C, sizeof(RT) == 8> w(fp[v], fm[v], gl, n, th);
#pragma unroll
            for (int m = 0; m < C; ++m) {
                // F+ of the left neighbour, F- of the right one; own slopes, the left
                // neighbour's F+ slope and the right neighbour's F- slope
                RT uu = u[v][m] - lam * (fp[v][m] - fm[v][m] + w.FM[m + 2] - w.FP[m]);
                uu = (RT)((double)uu - 0.5 * (double)lam * (double)(w.sp(m + 1) + w.sm(m) - w.sm(m + 1) - w.sp(m)));
                u[v][m] = uu;
            }
        }
        return;
    }
    if constexpr (XW > 1) coop_put<RT, C, XW>(x, parity, fp, fm);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        RT FP[C + 4], FM[C + 4];
        if constexpr (XW > 1) {
            RT e[4];
            coop_get<RT, XW>(x, parity, v, e);
            make_window_or<RT, C>(fp[v], FP, e[0], e[1]);
            make_window_or<RT, C>(fm[v], FM, e[2], e[3]);
        } else {
            make_window<RT, C, false>(fp[v], FP);
            make_window<RT, C, false>(fm[v], FM);
        }
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
template <bool KAHAN>
__device__ __forceinline__ void c32_update(float& hi, float& lo, float tl, float tr) {
    if constexpr (KAHAN) {
        const float y = (tl - tr) + lo;
        const float s = hi + y;
        lo = y - (s - hi);
        hi = s;
    } else {
        float a = hi, b = lo, e;
        two_sum(a, tl, a, e);
        b = b + e;
        two_sum(a, -tr, a, e);
        b = b + e;
        fast_two_sum(a, b, hi, lo);
    }
}

template <int ORDER, int C, int WM, bool KAHAN, int LIM, int XW>
__device__ __forceinline__ void step_c32(float (&hi)[3][C], float (&lo)[3][C], const Inst& in, int gl,
                                         const Coop<float, XW>& x, int parity) {
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
    // shared slopes up to 4 cells per lane: at 8 the FP32 kernels ran 4% slower with them
    // (their schedule needed 168 registers instead of 164) than they gain in instructions
    if constexpr (ORDER == 2 && XW == 1 && C >= 2 && C <= 4) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const SharedSlopes<A, WM, LIM, float, C, !KAHAN> w(fp[v], fm[v], gl, n, th);
            float t[C + 1];
#pragma unroll
            for (int f = 0; f <= C; ++f)  // face f lies between cells gl+f-1 and gl+f
                t[f] = __builtin_fmaf(hlam, w.sp(f) - w.sm(f), lam * (w.FP[f] + w.FM[f + 1]));
#pragma unroll
            for (int m = 0; m < C; ++m) c32_update<KAHAN>(hi[v][m], lo[v][m], t[m], t[m + 1]);
        }
        return;
    }
    if constexpr (XW > 1) coop_put<float, C, XW>(x, parity, fp, fm);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float FP[C + 4], FM[C + 4], t[C + 1];
        if constexpr (XW > 1) {
            float e[4];
            coop_get<float, XW>(x, parity, v, e);
            make_window_or<float, C>(fp[v], FP, e[0], e[1]);
            make_window_or<float, C>(fm[v], FM, e[2], e[3]);
        } else {
            make_window<float, C, !KAHAN>(fp[v], FP);
            make_window<float, C, !KAHAN>(fm[v], FM);
        }
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
        for (int m = 0; m < C; ++m) c32_update<KAHAN>(hi[v][m], lo[v][m], t[m], t[m + 1]);
    }
}

// ---------------------------------------------------------------------------
// F64F: the C32 face-term arrangement with every value a double (no compensation: the
// face terms are added to the FP64 state directly, u += t[m] - t[m+1]).  The split fluxes
// come scaled by DT/DX, so a face term is G + d/2 (G = F+ of the left cell + F- of the
// right cell, d = the difference of their limited slopes).
// ---------------------------------------------------------------------------
template <int ORDER, int C, int WM, int LIM, int XW>
__device__ __forceinline__ void step_f64f(double (&u)[3][C], const Inst& in, int gl, const Coop<double, XW>& x,
                                          int parity) {
    using A = F64FArith;
    const double lam = in.k.d_lamx, th = in.k.d_theta;
    const int n = in.nx;
    if constexpr (WM != WM_NONE) ghost_cells<C>(u, gl, n, in.bc == BC_REFLECT);
    double fp[3][C], fm[3][C];
#pragma unroll
    for (int m = 0; m < C; ++m) {
        double p[3], q[3], a;
        flux1_f64f(u[0][m], u[1][m], u[2][m], in, lam, p, q, a);  // scaled by DT/DX
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            fp[v][m] = p[v];
            fm[v][m] = q[v];
        }
    }
    if constexpr (ORDER == 2 && XW == 1 && C >= 2) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const SharedSlopes<A, WM, LIM, double, C, true> w(fp[v], fm[v], gl, n, th);
            double t[C + 1];
#pragma unroll
            for (int f = 0; f <= C; ++f)  // face f lies between cells gl+f-1 and gl+f
                t[f] = __builtin_fma(0.5, w.sp(f) - w.sm(f), w.FP[f] + w.FM[f + 1]);  // G + d/2
#pragma unroll
            for (int m = 0; m < C; ++m) u[v][m] = u[v][m] + (t[m] - t[m + 1]);
        }
        return;
    }
    if constexpr (XW > 1) coop_put<double, C, XW>(x, parity, fp, fm);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        double FP[C + 4], FM[C + 4], t[C + 1];
        if constexpr (XW > 1) {
            double e[4];
            coop_get<double, XW>(x, parity, v, e);
            make_window_or<double, C>(fp[v], FP, e[0], e[1]);
            make_window_or<double, C>(fm[v], FM, e[2], e[3]);
        } else {
            make_window<double, C, false>(fp[v], FP);
            make_window<double, C, false>(fm[v], FM);
        }
#pragma unroll
        for (int f = 0; f <= C; ++f) {
            const int gf = gl + f;  // face gf lies between cells gf-1 and gf
            const double G = FP[f + 1] + FM[f + 2];
            if constexpr (ORDER == 2) {
                const double d =
                    slope_at<A, WM, LIM>(FP, f - 1, gf - 1, n, th) - slope_at<A, WM, LIM>(FM, f, gf, n, th);
                t[f] = __builtin_fma(0.5, d, G);  // G + d/2
            } else {
                t[f] = G;
            }
        }
#pragma unroll
        for (int m = 0; m < C; ++m) u[v][m] = u[v][m] + (t[m] - t[m + 1]);
    }
}

// ---------------------------------------------------------------------------
// ns steps of one segment from registers, then the owned cells back to HBM.  Each
// wall mode gets its own copy of loop and store, so the loop-carried state never
// has to be moved between the copies' registers.
// ---------------------------------------------------------------------------
template <int ORDER, int C, int WM, bool KAHAN, int LIM, int XW>
__device__ __forceinline__ void advance_c32(float (&hi)[3][C], float (&lo)[3][C], const Inst& in, const Item1& it,
                                            int gl, int ns, const StepArgs& a, const Coop<float, XW>& x) {
    for (int s = 0; s < ns; ++s) step_c32<ORDER, C, WM, KAHAN, LIM, XW>(hi, lo, in, gl, x, s & 1);
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

template <int ORDER, int C, int WM, int LIM, int XW>
__device__ __forceinline__ void advance_f64f(double (&u)[3][C], const Inst& in, const Item1& it, int gl, int ns,
                                             const StepArgs& a, const Coop<double, XW>& x) {
    for (int s = 0; s < ns; ++s) step_f64f<ORDER, C, WM, LIM, XW>(u, in, gl, x, s & 1);
    double* __restrict__ ou = (double*)a.out0 + in.off;
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

template <int PREC, int ORDER, int C, int WM, int LIM, int XW>
__device__ __forceinline__ void advance_ref(typename Real<PREC>::T (&u)[3][C], const Inst& in, const Item1& it,
                                            int gl, int ns, const StepArgs& a,
                                            const Coop<typename Real<PREC>::T, XW>& x) {
    using RT = typename Real<PREC>::T;
    for (int s = 0; s < ns; ++s) step_ref<PREC, ORDER, C, WM, LIM, XW>(u, in, gl, x, s & 1);
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
// Wavefronts per SIMD the register allocator must leave room for (1: no bound).  Left to
// itself it spends registers on scheduling freedom in the second-order F64F kernel with
// independent wavefronts at 8 cells per lane and drops it to one wavefront per SIMD; bound
// to 2, it spills a few values outside the step loop.
template <int PREC, int ORDER, int C, int WPB>
constexpr int step1d_min_waves() {
    return PREC == PREC_F64F && ORDER == 2 && C == 8 && WPB == 1 ? 2 : 1;
}

template <int PREC, int ORDER, int C, int WPB, int LIM>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(step1d_min_waves<PREC, ORDER, C, WPB>())))
void k_step1d(StepArgs a) {
    using RT = typename Real<PREC>::T;
    // blocks of several wavefronts with at least 2 cells per lane are cooperative: one item
    // per block, one segment of 64*C cells per wavefront
    constexpr int XW = WPB > 1 && C >= 2 ? WPB : 1;
    __shared__ __align__(16) RT xlds[XW > 1 ? 2 * XW * CoopSlots<RT>::WAVE : 1];
    const int wib = uniform((int)threadIdx.x >> 6);
    const LaunchView lv = launch_view(a);
    constexpr int IPB = XW > 1 ? 1 : WPB;  // items per block
    const int blk = xcd_work((int)blockIdx.x, (lv.n_items + IPB - 1) / IPB);
    const int item = XW > 1 ? blk : blk * WPB + wib;
    // every wavefront of a cooperative block has the same item and step count, so all of
    // them return here together or reach the same number of barriers
    if (blk < 0 || item >= lv.n_items) return;
    const int lane = (int)threadIdx.x & 63;
    const Item1 it = ((const Item1*)a.items)[item];
    const Inst in = a.inst[it.inst];
    const int64_t left = in.nsteps - lv.step_base;
    const int ns = left < (int64_t)a.ksteps ? (int)left : a.ksteps;
    if (ns <= 0) return;
    const int n = in.nx;
    const int gw = it.g0 + (XW > 1 ? wib * 64 * C : 0);  // this wavefront's first cell
    const int gl = gw + lane * C;
    const int64_t P = a.plane;
    const Coop<RT, XW> x{xlds, wib, lane};
    // wall rules can only apply if the segment reaches within 3 cells of an end
    const bool walls = gw <= 2 || gw + 64 * C >= n - 3;
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
        if (wm == WM_NONE) advance_c32<ORDER, C, WM_NONE, K, LIM, XW>(hi, lo, in, it, gl, ns, a, x);
        else if (ORDER == 1 || wm == WM_GHOST) advance_c32<ORDER, C, WM_GHOST, K, LIM, XW>(hi, lo, in, it, gl, ns, a, x);
        else advance_c32<ORDER, C, WR, K, LIM, XW>(hi, lo, in, it, gl, ns, a, x);
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
        if constexpr (PREC == PREC_F64F) {
            if (wm == WM_NONE) advance_f64f<ORDER, C, WM_NONE, LIM, XW>(u, in, it, gl, ns, a, x);
            else if (ORDER == 1 || wm == WM_GHOST) advance_f64f<ORDER, C, WM_GHOST, LIM, XW>(u, in, it, gl, ns, a, x);
            else advance_f64f<ORDER, C, WR, LIM, XW>(u, in, it, gl, ns, a, x);
        } else {
            if (wm == WM_NONE) advance_ref<PREC, ORDER, C, WM_NONE, LIM, XW>(u, in, it, gl, ns, a, x);
            else if (ORDER == 1 || wm == WM_GHOST) advance_ref<PREC, ORDER, C, WM_GHOST, LIM, XW>(u, in, it, gl, ns, a, x);
            else advance_ref<PREC, ORDER, C, WR, LIM, XW>(u, in, it, gl, ns, a, x);
        }
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

// items (segments) per block: one for cooperative blocks (several waves, C >= 2), else one per wave
int step1d_items_per_block(int C, int wpb) { return wpb > 1 && C >= 2 ? 1 : wpb; }

hipError_t launch_step1d(int prec, int order, int lim, int C, int wpb, int n_items, hipStream_t s,
                         const StepArgs& a) {
    const int ipb = step1d_items_per_block(C, wpb);
    const int nblocks = xcd_grid((n_items + ipb - 1) / ipb);
    if (nblocks <= 0) return hipSuccess;
    switch (prec) {
    case PREC_F64: return launch_p<PREC_F64>(order, lim, C, wpb, nblocks, s, a);
    case PREC_REF32: return launch_p<PREC_REF32>(order, lim, C, wpb, nblocks, s, a);
    case PREC_C32: return launch_p<PREC_C32>(order, lim, C, wpb, nblocks, s, a);
    case PREC_C32K: return launch_p<PREC_C32K>(order, lim, C, wpb, nblocks, s, a);
    case PREC_F64F: return launch_p<PREC_F64F>(order, lim, C, wpb, nblocks, s, a);
    default: return hipErrorInvalidValue;
    }
}

bool step1d_supported(int C, int wpb) { return (C == 1 || C == 2 || C == 4 || C == 8) && (wpb == 1 || wpb == 4); }

}  // namespace shll
