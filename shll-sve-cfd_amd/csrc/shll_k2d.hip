// shll_k2d.hip -- fused 2D SHLL step kernels (first and second order) for gfx950.
//
// Layout: SoA planes, cell (i, j) at i*ny + j, y contiguous (base_shll_2d.c:146-221).
// One wavefront takes a strip of 64 consecutive columns j (lane = j - j0) and streams
// down the rows i.  For each row it loads the state once (64 consecutive cells per
// variable: one coalesced access), recovers the primitives, forms the x and y split
// fluxes, and keeps a rolling window of 2*ORDER+1 rows of x fluxes in registers.
// Row c = t - ORDER is updated once row t is in: its x neighbours come from the
// window, its y neighbours from the adjacent lanes (DPP wave shifts).  The outer
// ORDER lanes of a strip are recomputed halo.  No LDS, no barriers.
//
// The update keeps the reference's unsplit order: x term, (second-order x term),
// y term, (second-order y term), all from the same state (base_shll_2d.c:225-236,
// 2nd_order_base_shll.c:436-458).  Walls: reflective with per-component signs
// (base_shll_2d.c:150-216) or outflow (2nd_order_base_shll.c:214-323).  Second-order
// slopes: minmod or the MC limiter of base-omp/2nd_order_base_shll.c:317-325,385-392.
#include "shll_device.h"

#ifndef PF_DEPTH
#define PF_DEPTH 2
#endif

namespace shll {

template <int PREC> struct Real2 { using T = double; };
template <> struct Real2<PREC_REF32> { using T = float; };
template <> struct Real2<PREC_C32> { using T = float; };
template <> struct Real2<PREC_C32K> { using T = float; };

// x and y split fluxes of one cell.  U: conserved (F64/REF32: the state; C32: hi).
template <int PREC, class RT>
__device__ __forceinline__ void fluxes2(const RT (&U)[4], const Inst& in, RT (&xp)[4], RT (&xm)[4], RT (&yp)[4],
                                        RT (&ym)[4]) {
    RT vx, vy, T, a, z1, z2, z3;
    if constexpr (PREC == PREC_C32 || PREC == PREC_C32K) {
        C32Arith::prim2(U[0], U[1], U[2], U[3], in.k.c_rcv, in.k.c_gR, vx, vy, T, a);
    } else {
        const RT cv = PREC == PREC_F64 ? (RT)in.k.d_cv : (RT)in.k.f_cv;
        const RT gR = PREC == PREC_F64 ? (RT)in.k.d_gR : (RT)in.k.f_gR;
        RefArith<RT>::prim2(U[0], U[1], U[2], U[3], cv, gR, vx, vy, T, a);
    }
    const RT P = U[0] * T;  // (rho*R)*T with R = 1 (base_shll_2d.c:24, :246)
    if constexpr (PREC == PREC_C32 || PREC == PREC_C32K) {
        // the compensated scheme's float form: fused multiply-adds (split_c32)
        C32Arith::zinv(div_fast(vx, a), a, z1, z2, z3);
        split_c32(U[1], U[0], z1, z2, z3, xp[0], xm[0]);
        split_c32(__builtin_fmaf(U[1], vx, P), U[1], z1, z2, z3, xp[1], xm[1]);
        split_c32(U[1] * vy, U[2], z1, z2, z3, xp[2], xm[2]);
        split_c32(vx * (U[3] + P), U[3], z1, z2, z3, xp[3], xm[3]);
        C32Arith::zinv(div_fast(vy, a), a, z1, z2, z3);
        split_c32(U[2], U[0], z1, z2, z3, yp[0], ym[0]);
        split_c32(U[2] * vx, U[1], z1, z2, z3, yp[1], ym[1]);
        split_c32(__builtin_fmaf(U[2], vy, P), U[2], z1, z2, z3, yp[2], ym[2]);
        split_c32(vy * (U[3] + P), U[3], z1, z2, z3, yp[3], ym[3]);
    } else {
        // x direction (base_shll_2d.c:249-272)
        RefArith<RT>::zinv(div_rn(vx, a), a, z1, z2, z3);
        split<RT>(U[1], U[0], z1, z2, z3, xp[0], xm[0]);
        split<RT>(U[1] * vx + P, U[1], z1, z2, z3, xp[1], xm[1]);
        split<RT>(U[1] * vy, U[2], z1, z2, z3, xp[2], xm[2]);
        split<RT>(vx * (U[3] + P), U[3], z1, z2, z3, xp[3], xm[3]);
        // y direction (base_shll_2d.c:275-298)
        RefArith<RT>::zinv(div_rn(vy, a), a, z1, z2, z3);
        split<RT>(U[2], U[0], z1, z2, z3, yp[0], ym[0]);
        split<RT>(U[2] * vx, U[1], z1, z2, z3, yp[1], ym[1]);
        split<RT>(U[2] * vy + P, U[2], z1, z2, z3, yp[2], ym[2]);
        split<RT>(vy * (U[3] + P), U[3], z1, z2, z3, yp[3], ym[3]);
    }
}

// reflective wall signs: x walls (-,+,-,-), y walls (-,-,+,-)
__device__ __forceinline__ int sgn_x(int v) { return v == 1 ? 1 : -1; }
__device__ __forceinline__ int sgn_y(int v) { return v == 2 ? 1 : -1; }

template <int PREC, int ORDER, int WPB, int LIM>
__global__ __launch_bounds__(64 * WPB) void k_step2d(StepArgs a) {
    using RT = typename Real2<PREC>::T;
    using A = typename std::conditional<PREC == PREC_C32 || PREC == PREC_C32K, C32Arith, RefArith<RT>>::type;
    constexpr int NW = 2 * ORDER + 1;  // rows of x fluxes kept: c-ORDER .. c+ORDER
    constexpr int NS = ORDER + 1;      // rows of state / y fluxes kept: c .. c+ORDER
    constexpr bool CMP = PREC == PREC_C32 || PREC == PREC_C32K;  // (hi, lo) state

    const int wave = (int)blockIdx.x * WPB + uniform((int)threadIdx.x >> 6);
    const LaunchView lv = launch_view(a);
    if (wave >= lv.n_items) return;
    const int lane = (int)threadIdx.x & 63;
    const Item2 it = ((const Item2*)a.items)[wave];
    const Inst in = a.inst[it.inst];
    if (in.nsteps - lv.step_base <= 0) return;
    const int nx = in.nx, ny = in.ny;
    const int64_t P = a.plane;
    const int j = it.j0 + lane;
    const bool jok = j >= 0 && j < ny;
    const bool jown = j >= it.jx0 && j < it.jx0 + it.jw;
    const bool refl = in.bc == BC_REFLECT;
    const int jc = jok ? j : (j < 0 ? 0 : ny - 1);  // clamped column for addressing

    const RT* __restrict__ i0p = (const RT*)a.in0 + in.off;
    const RT* __restrict__ i1p = CMP ? (const RT*)a.in1 + in.off : nullptr;
    RT* __restrict__ o0p = (RT*)a.out0 + in.off;
    RT* __restrict__ o1p = CMP ? (RT*)a.out1 + in.off : nullptr;

    RT xp[NW][4], xm[NW][4];   // x split fluxes, window rows
    RT yp[NS][4], ym[NS][4];   // y split fluxes
    RT su[NS][4], sl[NS][4];   // state (C32: hi, lo)

#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int v = 0; v < 4; ++v) xp[w][v] = xm[w][v] = (RT)0;
#pragma unroll
    for (int w = 0; w < NS; ++w)
#pragma unroll
        for (int v = 0; v < 4; ++v) yp[w][v] = ym[w][v] = su[w][v] = sl[w][v] = (RT)0;

    const RT lamx = PREC == PREC_F64 ? (RT)in.k.d_lamx : (PREC == PREC_REF32 ? (RT)in.k.f_lamx : (RT)in.k.c_lamx);
    const RT lamy = PREC == PREC_F64 ? (RT)in.k.d_lamy : (PREC == PREC_REF32 ? (RT)in.k.f_lamy : (RT)in.k.c_lamy);
    const RT th = PREC == PREC_F64 ? (RT)in.k.d_theta : (PREC == PREC_REF32 ? (RT)in.k.f_theta : (RT)in.k.c_theta);

    // The state of row t+D is loaded while row t is processed (before row t's stores:
    // one counter tracks loads and stores, so a row's wait never covers the stores of
    // the rows before it); D rows of loads in flight per wave keep HBM busy.
    constexpr int D = PF_DEPTH;
    RT nu[D][4], nl[D][4];
    auto load_row = [&](int t, RT (&U)[4], RT (&Ul)[4]) {
        if (t >= 0 && t < nx) {
            const int64_t o = (int64_t)t * ny + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                U[v] = jok ? i0p[v * P + o] : (RT)(v == 0 ? 1.0 : (v == 3 ? 2.5 : 0.0));
                Ul[v] = CMP && jok ? i1p[v * P + o] : (RT)0;
            }
        }
    };
    const int tend = it.ix0 + it.iw + ORDER;
#pragma unroll
    for (int d = 0; d < D; ++d) load_row(it.ix0 - ORDER + d, nu[d], nl[d]);
    for (int t = it.ix0 - ORDER; t < tend; ++t) {
        // rotate the windows by one row
#pragma unroll
        for (int w = 0; w + 1 < NW; ++w)
#pragma unroll
            for (int v = 0; v < 4; ++v) { xp[w][v] = xp[w + 1][v]; xm[w][v] = xm[w + 1][v]; }
#pragma unroll
        for (int w = 0; w + 1 < NS; ++w)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                yp[w][v] = yp[w + 1][v]; ym[w][v] = ym[w + 1][v];
                su[w][v] = su[w + 1][v]; sl[w][v] = sl[w + 1][v];
            }
        const bool tin = t >= 0 && t < nx;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            su[NS - 1][v] = nu[0][v];
            sl[NS - 1][v] = nl[0][v];
        }
#pragma unroll
        for (int d = 0; d + 1 < D; ++d)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                nu[d][v] = nu[d + 1][v];
                nl[d][v] = nl[d + 1][v];
            }
        if (t + D < tend) load_row(t + D, nu[D - 1], nl[D - 1]);
        if (tin) fluxes2<PREC, RT>(su[NS - 1], in, xp[NW - 1], xm[NW - 1], yp[NS - 1], ym[NS - 1]);
        const int c = t - ORDER;  // row to update
        if (c < it.ix0 || c >= nx) continue;
        const int64_t o = (int64_t)c * ny + jc;

#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const RT sx = (RT)sgn_x(v), sy = (RT)sgn_y(v);
            const RT fpc = xp[ORDER][v], fmc = xm[ORDER][v];
            const RT hpc = yp[0][v], hmc = ym[0][v];
            const RT hpl = from_left(hpc), hml = from_left(hmc);
            const RT hpr = from_right(hpc), hmr = from_right(hmc);
            // slopes (second order); zero on boundary rows / columns
            RT dfp = 0, dfm = 0, dfpl = 0, dfmr = 0, dhp = 0, dhm = 0, dhpl = 0, dhmr = 0;
            if constexpr (ORDER == 2) {
                dfp = (c <= 0 || c >= nx - 1) ? (RT)0 : limited_slope<A, LIM>(xp[1][v], xp[2][v], xp[3][v], th);
                dfm = (c <= 0 || c >= nx - 1) ? (RT)0 : limited_slope<A, LIM>(xm[1][v], xm[2][v], xm[3][v], th);
                dfpl = (c - 1 <= 0 || c - 1 >= nx - 1) ? (RT)0
                                                       : limited_slope<A, LIM>(xp[0][v], xp[1][v], xp[2][v], th);
                dfmr = (c + 1 <= 0 || c + 1 >= nx - 1) ? (RT)0
                                                       : limited_slope<A, LIM>(xm[2][v], xm[3][v], xm[4][v], th);
                dhp = (j <= 0 || j >= ny - 1) ? (RT)0 : limited_slope<A, LIM>(hpl, hpc, hpr, th);
                dhm = (j <= 0 || j >= ny - 1) ? (RT)0 : limited_slope<A, LIM>(hml, hmc, hmr, th);
                dhpl = from_left(dhp);
                dhmr = from_right(dhm);
            }
            if constexpr (!CMP) {
                // reference association
                RT Lv = c == 0 ? (refl ? sx * fmc : fpc) : xp[ORDER - 1][v];
                RT Rv = c == nx - 1 ? (refl ? sx * fpc : fmc) : xm[ORDER + 1][v];
                RT Bv = j == 0 ? (refl ? sy * hmc : hpc) : hpl;
                RT Tv = j == ny - 1 ? (refl ? sy * hpc : hmc) : hmr;
                RT uu = su[0][v] - lamx * (fpc - fmc + Rv - Lv);
                if constexpr (ORDER == 2) {
                    RT Ldf = c == 0 ? (RT)0 : dfpl;
                    RT Rdf = c == nx - 1 ? (RT)0 : dfmr;
                    uu = (RT)((double)uu - 0.5 * (double)lamx * (double)(dfp + dfm - Rdf - Ldf));
                }
                uu = uu - lamy * (hpc - hmc + Tv - Bv);
                if constexpr (ORDER == 2) {
                    RT Bdf = j == 0 ? (RT)0 : dhpl;
                    RT Tdf = j == ny - 1 ? (RT)0 : dhmr;
                    uu = (RT)((double)uu - 0.5 * (double)lamy * (double)(dhp + dhm - Tdf - Bdf));
                }
                if (jown && c < it.ix0 + it.iw) o0p[v * P + o] = uu;
            } else {
                // face terms t = lam*(G + d/2) (regrouped reference update; d = slope
                // difference, zero at walls and in first order), TwoSum into (hi, lo)
                const RT fpl = xp[ORDER - 1][v], fmr = xm[ORDER + 1][v];
                RT Gl = fpl + fmc, Gr = fpc + fmr;
                RT Gb = hpl + hmc;
                Gl = c == 0 ? (refl ? fmc + sx * fmc : fpc + fmc) : Gl;
                Gr = c == nx - 1 ? (refl ? fpc + sx * fpc : fpc + fmc) : Gr;
                Gb = j == 0 ? (refl ? hmc + sy * hmc : hpc + hmc) : Gb;
                RT tl = lamx * Gl, tr = lamx * Gr, tb = lamy * Gb;
                if constexpr (ORDER == 2) {
                    const RT dl = c == 0 ? (RT)0 : dfpl - dfm;
                    const RT dr = c == nx - 1 ? (RT)0 : dfp - dfmr;
                    const RT db = j == 0 ? (RT)0 : dhpl - dhm;
                    tl = __builtin_fmaf(0.5f * lamx, dl, tl);
                    tr = __builtin_fmaf(0.5f * lamx, dr, tr);
                    tb = __builtin_fmaf(0.5f * lamy, db, tb);
                }
                RT tt = from_right(tb);
                if constexpr (ORDER == 2)
                    tt = j == ny - 1 ? __builtin_fmaf(0.5f * lamy, (RT)0, lamy * (refl ? hpc + sy * hpc : hpc + hmc)) : tt;
                else
                    tt = j == ny - 1 ? lamy * (refl ? hpc + sy * hpc : hpc + hmc) : tt;
                RT aa = su[0][v], bb = sl[0][v], e;
                if constexpr (PREC == PREC_C32K) {
                    // one Kahan step: lo (bb) is the part of the value not yet in hi
                    const RT y = ((tl - tr) + (tb - tt)) + bb;
                    const RT s2 = aa + y;
                    e = y - (s2 - aa);
                    aa = s2;
                } else {
                    two_sum(aa, tl, aa, e);
                    bb = bb + e;
                    two_sum(aa, -tr, aa, e);
                    bb = bb + e;
                    two_sum(aa, tb, aa, e);
                    bb = bb + e;
                    two_sum(aa, -tt, aa, e);
                    bb = bb + e;
                    fast_two_sum(aa, bb, aa, e);
                }
                if (jown && c < it.ix0 + it.iw) {
                    o0p[v * P + o] = aa;
                    o1p[v * P + o] = e;
                }
            }
        }
    }
}

template <int PREC, int ORDER, int LIM>
static hipError_t launch2_o(int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    if (wpb == 1) hipLaunchKernelGGL((k_step2d<PREC, ORDER, 1, LIM>), dim3(nblocks), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_step2d<PREC, ORDER, 4, LIM>), dim3(nblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// first order; second order with minmod; second order with the MC limiter
template <int PREC>
static hipError_t launch2_p(int order, int lim, int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    if (order == 1) return lim == LIM_MINMOD ? launch2_o<PREC, 1, LIM_MINMOD>(wpb, nblocks, s, a) : hipErrorInvalidValue;
    if (order != 2) return hipErrorInvalidValue;
    if (lim == LIM_MINMOD) return launch2_o<PREC, 2, LIM_MINMOD>(wpb, nblocks, s, a);
    if (lim == LIM_MC) return launch2_o<PREC, 2, LIM_MC>(wpb, nblocks, s, a);
    return hipErrorInvalidValue;
}

hipError_t launch_step2d(int prec, int order, int lim, int wpb, int n_items, hipStream_t s, const StepArgs& a) {
    const int nblocks = (n_items + wpb - 1) / wpb;
    if (nblocks <= 0) return hipSuccess;
    switch (prec) {
    case PREC_F64: return launch2_p<PREC_F64>(order, lim, wpb, nblocks, s, a);
    case PREC_REF32: return launch2_p<PREC_REF32>(order, lim, wpb, nblocks, s, a);
    case PREC_C32: return launch2_p<PREC_C32>(order, lim, wpb, nblocks, s, a);
    case PREC_C32K: return launch2_p<PREC_C32K>(order, lim, wpb, nblocks, s, a);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace shll
