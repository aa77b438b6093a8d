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
#include "shll_flux.h"

#ifndef PF_DEPTH
#define PF_DEPTH 2
#endif

namespace shll {

// reflective wall signs: x walls (-,+,-,-), y walls (-,-,+,-)
__device__ __forceinline__ int sgn_x(int v) { return v == 1 ? 1 : -1; }
__device__ __forceinline__ int sgn_y(int v) { return v == 2 ? 1 : -1; }

template <int PREC, int ORDER, int WPB, int LIM>
__global__ __launch_bounds__(64 * WPB) void k_step2d(StepArgs a) {
    using RT = typename Real<PREC>::T;
    using A = typename std::conditional<PREC == PREC_C32 || PREC == PREC_C32K, C32Arith, RefArith<RT>>::type;
    constexpr int NW = 2 * ORDER + 1;  // rows of x fluxes kept: c-ORDER .. c+ORDER
    constexpr int NS = ORDER + 1;      // rows of state / y fluxes kept: c .. c+ORDER
    constexpr bool CMP = PREC == PREC_C32 || PREC == PREC_C32K;  // (hi, lo) state

    const LaunchView lv = launch_view(a);
    const int blk = xcd_work((int)blockIdx.x, (lv.n_items + WPB - 1) / WPB);
    const int wave = blk * WPB + uniform((int)threadIdx.x >> 6);
    if (blk < 0 || wave >= lv.n_items) return;
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
            const RT hpl = from_left_z(hpc), hml = from_left_z(hmc);
            const RT hpr = from_right_z(hpc), hmr = from_right_z(hmc);
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
                dhpl = from_left_z(dhp);
                dhmr = from_right_z(dhm);
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
                RT tt = from_right_z(tb);
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

// ---------------------------------------------------------------------------
// First order, K steps per launch: a pipeline of K stages along the rows.
//
// Stage s takes the rows of its input level one at a time (stage 0 from HBM, stage s
// from stage s-1) and holds the last row it took.  When row t arrives, the x face
// between rows t-1 and t is known, so the held row t-1 is advanced one step and
// passed on; row t is held in its place.  After the last row a stage drains its held
// row with the wall face.  Each stage keeps only one row (state, x split fluxes and
// the row's face and y terms), so K steps cost K times the arithmetic of one step but
// one read and one write of the state.  The row next to a cut (an item edge that is
// not a wall) gets a made-up face and a strip's edge lanes get their own values from
// the DPP shifts, so each stage spoils one more row and column next to every cut: the
// items' halos are K rows and K columns wide and only owned cells are stored.  Stages
// at or past the instance's remaining step count pass rows through.
// ---------------------------------------------------------------------------
template <int PREC>
struct Held2 {
    using RT = typename Real<PREC>::T;
    RT u[4], l[4];    // state (C32: hi, lo)
    RT fp[4];         // x split flux F+ (the next row's lower face)
    RT d[4];          // F64/REF32: F+ - F- of the row
    RT tl[4];         // C32: lam_x * G of the lower x face; F64/REF32: L (F+ of the row below)
    RT yb[4], yt[4];  // C32: lam_y * G of the two y faces (C32K: yb = their difference);
                      // F64/REF32: yb = (hp - hm) + T - B
};

// A lane's y walls, as per-lane flags (lane-varying, so the wall choices are selects
// rather than branches): lo / hi = the lane is column 0 / ny-1, *_refl = and reflective
struct LaneWalls {
    bool lo, hi, lo_refl, hi_refl, refl;
};

enum { ST_TAKE_FIRST = 0, ST_TAKE = 1, ST_DRAIN = 2 };

// One stage, one row.  io: the incoming row (ST_TAKE_FIRST / ST_TAKE); on return the
// outgoing row (ST_TAKE / ST_DRAIN).  act: the stage advances (else it only delays).
template <int PREC, bool YW>
__device__ __forceinline__ void stage2(Held2<PREC>& h, typename Real<PREC>::T (&iou)[4],
                                       typename Real<PREC>::T (&iol)[4], int mode, bool act, const Inst& in,
                                       const LaneWalls& w, typename Real<PREC>::T lamx,
                                       typename Real<PREC>::T lamy) {
    using RT = typename Real<PREC>::T;
    constexpr bool CMP = PREC == PREC_C32 || PREC == PREC_C32K;
    const bool refl = w.refl;
    if (!act) {
        // pass-through with the stage's one-row delay
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const RT u = h.u[v], l = h.l[v];
            if (mode != ST_DRAIN) { h.u[v] = iou[v]; h.l[v] = iol[v]; }
            iou[v] = u;
            iol[v] = l;
        }
        return;
    }
    RT xp[4], xm[4], yp[4], ym[4], face[4], nu[4], nl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        nu[v] = iou[v];  // the incoming row (iou is overwritten by the outgoing one)
        nl[v] = iol[v];
    }
    // the x fluxes of the row that meets the held row's upper face: the incoming row, or
    // at a drain the held row itself (its F- for the outflow wall)
    fluxes2<PREC, RT>(mode == ST_DRAIN ? h.u : nu, in, xp, xm, yp, ym);
    if (mode != ST_TAKE_FIRST) {
        // advance the held row
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const RT sx = (RT)sgn_x(v);
            const RT fpc = h.fp[v];
            if constexpr (!CMP) {
                const RT Rv = mode == ST_DRAIN ? (refl ? sx * fpc : xm[v]) : xm[v];
                RT uu = h.u[v] - lamx * (h.d[v] + Rv - h.tl[v]);
                uu = uu - lamy * h.yb[v];
                iou[v] = uu;
            } else {
                // the face between the held and the incoming row (kept: the incoming row's lower face)
                const RT Gr = mode == ST_DRAIN ? (refl ? fpc + sx * fpc : fpc + xm[v]) : fpc + xm[v];
                const RT tl = h.tl[v], tr = lamx * Gr;
                face[v] = tr;
                RT aa = h.u[v], bb = h.l[v], e;
                if constexpr (PREC == PREC_C32K) {
                    const RT y = ((tl - tr) + h.yb[v]) + bb;
                    const RT s2 = aa + y;
                    e = y - (s2 - aa);
                    aa = s2;
                } else {
                    two_sum(aa, tl, aa, e);
                    bb = bb + e;
                    two_sum(aa, -tr, aa, e);
                    bb = bb + e;
                    two_sum(aa, h.yb[v], aa, e);
                    bb = bb + e;
                    two_sum(aa, -h.yt[v], aa, e);
                    bb = bb + e;
                    fast_two_sum(aa, bb, aa, e);
                }
                iou[v] = aa;
                iol[v] = e;
            }
        }
    }
    if (mode == ST_DRAIN) return;
    // hold the incoming row: its lower x face (the wall rule on a stage's first row), y terms
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const RT sx = (RT)sgn_x(v), sy = (RT)sgn_y(v);
        const RT fpc = xp[v], fmc = xm[v], hpc = yp[v], hmc = ym[v];
        const RT hpl = from_left_z(hpc), hmr = from_right_z(hmc);
        if constexpr (!CMP) {
            h.tl[v] = mode == ST_TAKE_FIRST ? (refl ? sx * fmc : fpc) : h.fp[v];
            const RT Bv = YW ? (w.lo_refl ? sy * hmc : (w.lo ? hpc : hpl)) : hpl;
            const RT Tv = YW ? (w.hi_refl ? sy * hpc : (w.hi ? hmc : hmr)) : hmr;
            h.yb[v] = hpc - hmc + Tv - Bv;
            h.d[v] = fpc - fmc;
        } else {
            h.tl[v] = mode == ST_TAKE_FIRST ? lamx * (refl ? fmc + sx * fmc : fpc + fmc) : face[v];
            // wall faces: (mirrored or own) flux + own flux; the sum commutes exactly
            const RT Gb = (YW ? (w.lo_refl ? sy * hmc : (w.lo ? hpc : hpl)) : hpl) + hmc;
            const RT tb = lamy * Gb;
            RT tt = from_right_z(tb);
            if constexpr (YW) tt = w.hi ? lamy * (hpc + (w.hi_refl ? sy * hpc : hmc)) : tt;
            if constexpr (PREC == PREC_C32K) {
                h.yb[v] = tb - tt;
            } else {
                h.yb[v] = tb;
                h.yt[v] = tt;
            }
        }
        h.u[v] = nu[v];
        h.l[v] = nl[v];
        h.fp[v] = fpc;
    }
}

template <int PREC, int K, bool YW>
__device__ __forceinline__ void pipeline2(const StepArgs& a, const Item2& it, const Inst& in, int lane, int ns) {
    using RT = typename Real<PREC>::T;
    constexpr bool CMP = PREC == PREC_C32 || PREC == PREC_C32K;
    const int nx = in.nx, ny = in.ny;
    const int64_t P = a.plane;
    const int j = it.j0 + lane;
    const bool jok = j >= 0 && j < ny;
    const bool jown = j >= it.jx0 && j < it.jx0 + it.jw;
    const int jc = jok ? j : (j < 0 ? 0 : ny - 1);
    const RT* __restrict__ i0p = (const RT*)a.in0 + in.off;
    const RT* __restrict__ i1p = CMP ? (const RT*)a.in1 + in.off : nullptr;
    RT* __restrict__ o0p = (RT*)a.out0 + in.off;
    RT* __restrict__ o1p = CMP ? (RT*)a.out1 + in.off : nullptr;
    const RT lamx = PREC == PREC_F64 ? (RT)in.k.d_lamx : (PREC == PREC_REF32 ? (RT)in.k.f_lamx : (RT)in.k.c_lamx);
    const RT lamy = PREC == PREC_F64 ? (RT)in.k.d_lamy : (PREC == PREC_REF32 ? (RT)in.k.f_lamy : (RT)in.k.c_lamy);
    const bool refl = in.bc == BC_REFLECT;
    const LaneWalls w{j == 0, j == ny - 1, j == 0 && refl, j == ny - 1 && refl, refl};

    // rows taken by every stage: [r0, r1), then a drain
    const int r0 = it.ix0 - K < 0 ? 0 : it.ix0 - K;
    const int r1 = it.ix0 + it.iw + K > nx ? nx : it.ix0 + it.iw + K;
    Held2<PREC> h[K];
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            h[s].u[v] = h[s].l[v] = h[s].fp[v] = h[s].d[v] = h[s].tl[v] = h[s].yb[v] = h[s].yt[v] = (RT)0;

    constexpr int D = PF_DEPTH;
    RT nu[D][4], nl[D][4];
    // jc is clamped into the domain: every lane loads, lanes off the domain keep a benign state
    auto load_row = [&](int t, RT (&U)[4], RT (&Ul)[4]) {
        if (t < r1) {
            const int64_t o = (int64_t)t * ny + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const RT x = i0p[v * P + o];
                U[v] = jok ? x : (RT)(v == 0 ? 1.0 : (v == 3 ? 2.5 : 0.0));
                if constexpr (CMP) {
                    const RT y = i1p[v * P + o];
                    Ul[v] = jok ? y : (RT)0;
                } else {
                    Ul[v] = (RT)0;
                }
            }
        }
    };
#pragma unroll
    for (int d = 0; d < D; ++d) load_row(r0 + d, nu[d], nl[d]);
    // Stage s is at row t - s in iteration t: it takes rows [r0, r1) (from stage s-1,
    // which gives up its held row when it takes the next one or drains) and drains at r1.
    // Stage K-1 gives row t - K.
    const int tend = r1 + K;
    auto iter = [&](int t, bool steady) {
        RT U[4], Ul[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            U[v] = nu[0][v];
            Ul[v] = nl[0][v];
        }
#pragma unroll
        for (int d = 0; d + 1 < D; ++d)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                nu[d][v] = nu[d + 1][v];
                nl[d][v] = nl[d + 1][v];
            }
        load_row(t + D, nu[D - 1], nl[D - 1]);
        // live: U holds a row at the current stage's input level
        bool live = true;
        if (steady) {
            // every stage takes an inner row and advances: one straight-line body
#pragma unroll
            for (int s = 0; s < K; ++s) stage2<PREC, YW>(h[s], U, Ul, ST_TAKE, true, in, w, lamx, lamy);
        } else {
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const int ts = t - s;
                if (ts < r0 || ts > r1) {
                    live = false;
                } else if (ts == r1) {
                    stage2<PREC, YW>(h[s], U, Ul, ST_DRAIN, s < ns, in, w, lamx, lamy);
                    live = true;
                } else if (live) {
                    const int mode = ts == r0 ? ST_TAKE_FIRST : ST_TAKE;
                    stage2<PREC, YW>(h[s], U, Ul, mode, s < ns, in, w, lamx, lamy);
                    live = mode == ST_TAKE;
                }
            }
        }
        const int c = t - K;
        if (live && jown && c >= it.ix0 && c < it.ix0 + it.iw) {
            const int64_t o = (int64_t)c * ny + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                o0p[v * P + o] = U[v];
                if constexpr (CMP) o1p[v * P + o] = Ul[v];
            }
        }
    };
    // iterations t in [r0 + K, r1) are steady when every stage advances (ns == K): they run
    // in a loop of their own, K at a time, apart from the ramp-up and drain iterations
    int t = r0;
    if (ns == K && r1 - (r0 + K) >= K) {
        for (; t < r0 + K; ++t) iter(t, false);
        for (; t + K <= r1; t += K)
#pragma unroll
            for (int q = 0; q < K; ++q) iter(t + q, true);
    }
    for (; t < tend; ++t) iter(t, ns == K && t - (K - 1) > r0 && t < r1);
}

template <int PREC, int K, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_step2d_tb(StepArgs a) {
    const LaunchView lv = launch_view(a);
    const int blk = xcd_work((int)blockIdx.x, (lv.n_items + WPB - 1) / WPB);
    const int wave = blk * WPB + uniform((int)threadIdx.x >> 6);
    if (blk < 0 || wave >= lv.n_items) return;
    const int lane = (int)threadIdx.x & 63;
    const Item2 it = ((const Item2*)a.items)[wave];
    const Inst in = a.inst[it.inst];
    const int64_t left = in.nsteps - lv.step_base;
    if (left <= 0) return;
    const int kk = a.ksteps < K ? a.ksteps : K;
    const int ns = left < (int64_t)kk ? (int)left : kk;  // stages that advance
    // strips away from both y walls take the copy without wall choices
    const bool yw = it.j0 <= 0 || it.j0 + 64 >= in.ny;
    if (yw) pipeline2<PREC, K, true>(a, it, in, lane, ns);
    else pipeline2<PREC, K, false>(a, it, in, lane, ns);
}

// ---------------------------------------------------------------------------
// Second order (FP32-compensated forms), K steps per launch: the same row pipeline with a
// delay of two rows per stage.  The face between rows c and c+1 needs the limited slopes
// of rows c and c+1, and the slope of row c+1 needs row c+2, so a stage advances row t-2
// when row t arrives.  It keeps two pending rows (state, x split fluxes, y terms), the F+
// slope of the first one and that row's lower face term.  Slopes of the domain's first
// and last rows are zero (2nd_order_base_shll.c:227-276); at a cut the first row's slope
// and the last row's F- slope are taken as zero, which spoils two rows per stage, like
// the radius-2 y stencil spoils two lanes at a strip edge: items carry 2K halo rows and
// columns.  Bitwise equal to K launches of k_step2d<prec, 2, ., LIM>.
// ---------------------------------------------------------------------------
template <int PREC>
struct Held2o2 {
    using RT = typename Real<PREC>::T;
    RT u0[4], l0[4], u1[4], l1[4];  // pending rows: the older (0) and the newer (1)
    RT xp0[4], xm0[4], xp1[4], xm1[4];
    RT dfp0[4];                     // limited F+ slope of row 0
    RT tl0[4];                      // lower face term of row 0
    RT yb0[4], yt0[4], yb1[4], yt1[4];  // y face terms (C32K: yb = tb - tt)
};

enum { S2_FIRST = 0, S2_SECOND = 1, S2_TAKE = 2, S2_DRAIN1 = 3, S2_DRAIN2 = 4 };

template <int PREC, bool YW, int LIM>
__device__ __forceinline__ void stage2o2(Held2o2<PREC>& h, float (&iou)[4], float (&iol)[4], int mode, bool act,
                                         int ts, const Inst& in, const LaneWalls& w, bool yedge, float lamx,
                                         float lamy, float th) {
    using A = C32Arith;
    const int nx = in.nx;
    const bool refl = w.refl;
    const float hl = 0.5f * lamx, hly = 0.5f * lamy;
    float nu[4], nl[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        nu[v] = iou[v];
        nl[v] = iol[v];
    }
    if (!act) {
        // pass-through with the stage's two-row delay
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float u = h.u0[v], l = h.l0[v];
            if (mode == S2_FIRST) {
                h.u0[v] = nu[v]; h.l0[v] = nl[v];
            } else if (mode == S2_SECOND) {
                h.u1[v] = nu[v]; h.l1[v] = nl[v];
            } else if (mode == S2_TAKE) {
                h.u0[v] = h.u1[v]; h.l0[v] = h.l1[v]; h.u1[v] = nu[v]; h.l1[v] = nl[v];
            } else if (mode == S2_DRAIN1) {
                h.u0[v] = h.u1[v]; h.l0[v] = h.l1[v];
            }
            iou[v] = u;
            iol[v] = l;
        }
        return;
    }
    float xp[4], xm[4], yp[4], ym[4], yb[4], yt[4];
    const bool take = mode == S2_FIRST || mode == S2_SECOND || mode == S2_TAKE;
    if (take) {
        fluxes2<PREC, float>(nu, in, xp, xm, yp, ym);
        // y face terms of the incoming row: lower face tb, upper face tt (lane above)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float sy = (float)sgn_y(v);
            const float hpc = yp[v], hmc = ym[v];
            const float hpl = from_left_z(hpc), hpr = from_right_z(hpc);
            const float hml = from_left_z(hmc), hmr = from_right_z(hmc);
            float dhp = limited_slope<A, LIM>(hpl, hpc, hpr, th);
            float dhm = limited_slope<A, LIM>(hml, hmc, hmr, th);
            if constexpr (YW) {
                dhp = yedge ? 0.0f : dhp;
                dhm = yedge ? 0.0f : dhm;
            }
            const float dhpl = from_left_z(dhp);
            float Gb = hpl + hmc, db = dhpl - dhm;
            if constexpr (YW) {
                Gb = w.lo ? (w.lo_refl ? hmc + sy * hmc : hpc + hmc) : Gb;
                db = w.lo ? 0.0f : db;
            }
            const float tb = __builtin_fmaf(hly, db, lamy * Gb);
            float tt = from_right_z(tb);
            if constexpr (YW)
                tt = w.hi ? __builtin_fmaf(hly, 0.0f, lamy * (w.hi_refl ? hpc + sy * hpc : hpc + hmc)) : tt;
            if constexpr (PREC == PREC_C32K) {
                yb[v] = tb - tt;
                yt[v] = 0.0f;
            } else {
                yb[v] = tb;
                yt[v] = tt;
            }
        }
    }
    if (mode == S2_FIRST) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float sx = (float)sgn_x(v);
            h.u0[v] = nu[v]; h.l0[v] = nl[v];
            h.xp0[v] = xp[v]; h.xm0[v] = xm[v];
            h.yb0[v] = yb[v]; h.yt0[v] = yt[v];
            // the wall face below the first row (made up at a cut)
            h.tl0[v] = __builtin_fmaf(hl, 0.0f, lamx * (refl ? xm[v] + sx * xm[v] : xp[v] + xm[v]));
        }
        return;
    }
    if (mode == S2_SECOND) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            h.u1[v] = nu[v]; h.l1[v] = nl[v];
            h.xp1[v] = xp[v]; h.xm1[v] = xm[v];
            h.yb1[v] = yb[v]; h.yt1[v] = yt[v];
            h.dfp0[v] = 0.0f;  // the first row: a wall row, or a cut
        }
        return;
    }
    // advance row 0 with its upper face
    const int r = ts - 1;  // the row whose slopes complete now (TAKE)
    const bool bnd = r <= 0 || r >= nx - 1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float sx = (float)sgn_x(v);
        float face, dfp1 = 0.0f;
        if (mode == S2_TAKE) {
            const float dfm1 = bnd ? 0.0f : limited_slope<A, LIM>(h.xm0[v], h.xm1[v], xm[v], th);
            dfp1 = bnd ? 0.0f : limited_slope<A, LIM>(h.xp0[v], h.xp1[v], xp[v], th);
            face = __builtin_fmaf(hl, h.dfp0[v] - dfm1, lamx * (h.xp0[v] + h.xm1[v]));
        } else if (mode == S2_DRAIN1) {
            face = __builtin_fmaf(hl, h.dfp0[v] - 0.0f, lamx * (h.xp0[v] + h.xm1[v]));
        } else {
            const float fpc = h.xp0[v], fmc = h.xm0[v];
            face = __builtin_fmaf(hl, 0.0f, lamx * (refl ? fpc + sx * fpc : fpc + fmc));
        }
        const float tl = h.tl0[v];
        float aa = h.u0[v], bb = h.l0[v], e;
        if constexpr (PREC == PREC_C32K) {
            const float y = ((tl - face) + h.yb0[v]) + bb;
            const float s2 = aa + y;
            e = y - (s2 - aa);
            aa = s2;
        } else {
            two_sum(aa, tl, aa, e);
            bb = bb + e;
            two_sum(aa, -face, aa, e);
            bb = bb + e;
            two_sum(aa, h.yb0[v], aa, e);
            bb = bb + e;
            two_sum(aa, -h.yt0[v], aa, e);
            bb = bb + e;
            fast_two_sum(aa, bb, aa, e);
        }
        iou[v] = aa;
        iol[v] = e;
        if (mode != S2_DRAIN2) {
            // row 1 becomes row 0
            h.u0[v] = h.u1[v]; h.l0[v] = h.l1[v];
            h.xp0[v] = h.xp1[v]; h.xm0[v] = h.xm1[v];
            h.yb0[v] = h.yb1[v]; h.yt0[v] = h.yt1[v];
            h.tl0[v] = face;
            h.dfp0[v] = dfp1;
        }
        if (mode == S2_TAKE) {
            h.u1[v] = nu[v]; h.l1[v] = nl[v];
            h.xp1[v] = xp[v]; h.xm1[v] = xm[v];
            h.yb1[v] = yb[v]; h.yt1[v] = yt[v];
        }
    }
}

template <int PREC, int K, bool YW, int LIM>
__device__ __forceinline__ void pipeline2o2(const StepArgs& a, const Item2& it, const Inst& in, int lane, int ns) {
    const int nx = in.nx, ny = in.ny;
    const int64_t P = a.plane;
    const int j = it.j0 + lane;
    const bool jok = j >= 0 && j < ny;
    const bool jown = j >= it.jx0 && j < it.jx0 + it.jw;
    const int jc = jok ? j : (j < 0 ? 0 : ny - 1);
    const float* __restrict__ i0p = (const float*)a.in0 + in.off;
    const float* __restrict__ i1p = (const float*)a.in1 + in.off;
    float* __restrict__ o0p = (float*)a.out0 + in.off;
    float* __restrict__ o1p = (float*)a.out1 + in.off;
    const float lamx = in.k.c_lamx, lamy = in.k.c_lamy, th = in.k.c_theta;
    const bool refl = in.bc == BC_REFLECT;
    const LaneWalls w{j == 0, j == ny - 1, j == 0 && refl, j == ny - 1 && refl, refl};
    const bool yedge = j <= 0 || j >= ny - 1;  // zero y slopes (2nd_order_base_shll.c:293-344)

    constexpr int H = 2 * K;
    const int r0 = it.ix0 - H < 0 ? 0 : it.ix0 - H;
    const int r1 = it.ix0 + it.iw + H > nx ? nx : it.ix0 + it.iw + H;
    Held2o2<PREC> h[K];
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            h[s].u0[v] = h[s].l0[v] = h[s].u1[v] = h[s].l1[v] = 0.0f;
            h[s].xp0[v] = h[s].xm0[v] = h[s].xp1[v] = h[s].xm1[v] = h[s].dfp0[v] = h[s].tl0[v] = 0.0f;
            h[s].yb0[v] = h[s].yt0[v] = h[s].yb1[v] = h[s].yt1[v] = 0.0f;
        }
    constexpr int D = PF_DEPTH;
    float nu[D][4], nl[D][4];  // rows t+1 .. t+D, loaded ahead
    auto load_row = [&](int t, float (&U)[4], float (&Ul)[4]) {
        if (t < r1) {
            const int64_t o = (int64_t)t * ny + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float x = i0p[v * P + o], y = i1p[v * P + o];
                U[v] = jok ? x : (v == 0 ? 1.0f : (v == 3 ? 2.5f : 0.0f));
                Ul[v] = jok ? y : 0.0f;
            }
        }
    };
#pragma unroll
    for (int d = 0; d < D; ++d) load_row(r0 + d, nu[d], nl[d]);
    // Stage s is at row t - 2s in iteration t: it takes rows [r0, r1) and drains at r1 and
    // r1 + 1, giving row t - 2s - 2 from its third row on.  Stage K-1 gives row t - 2K.
    const int tend = r1 + 2 * K;
    auto iter = [&](int t, bool steady) {
        float U[4], Ul[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            U[v] = nu[0][v];
            Ul[v] = nl[0][v];
        }
#pragma unroll
        for (int d = 0; d + 1 < D; ++d)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                nu[d][v] = nu[d + 1][v];
                nl[d][v] = nl[d + 1][v];
            }
        load_row(t + D, nu[D - 1], nl[D - 1]);
        bool live = true;  // U holds a row at the current stage's input level
        if (steady) {
#pragma unroll
            for (int s = 0; s < K; ++s)
                stage2o2<PREC, YW, LIM>(h[s], U, Ul, S2_TAKE, true, t - 2 * s, in, w, yedge, lamx, lamy, th);
        } else {
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const int ts = t - 2 * s;
                if (ts < r0 || ts > r1 + 1) {
                    live = false;
                } else if (ts >= r1) {
                    stage2o2<PREC, YW, LIM>(h[s], U, Ul, ts == r1 ? S2_DRAIN1 : S2_DRAIN2, s < ns, ts, in, w, yedge,
                                            lamx, lamy, th);
                    live = true;
                } else if (live) {
                    const int mode = ts == r0 ? S2_FIRST : (ts == r0 + 1 ? S2_SECOND : S2_TAKE);
                    stage2o2<PREC, YW, LIM>(h[s], U, Ul, mode, s < ns, ts, in, w, yedge, lamx, lamy, th);
                    live = mode == S2_TAKE;
                }
            }
        }
        const int c = t - 2 * K;
        if (live && jown && c >= it.ix0 && c < it.ix0 + it.iw) {
            const int64_t o = (int64_t)c * ny + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                o0p[v * P + o] = U[v];
                o1p[v * P + o] = Ul[v];
            }
        }
    };
    // iterations t in [r0 + 2K, r1) are steady when every stage advances (ns == K): they run
    // in a loop of their own, apart from the ramp-up and drain iterations
    int t = r0;
    if (ns == K && r1 - (r0 + 2 * K) >= 2 * K) {
        for (; t < r0 + 2 * K; ++t) iter(t, false);
        for (; t + 2 * K <= r1; t += 2 * K)
#pragma unroll
            for (int q = 0; q < 2 * K; ++q) iter(t + q, true);
    }
    for (; t < tend; ++t) iter(t, ns == K && t - 2 * (K - 1) >= r0 + 2 && t < r1);
}

template <int PREC, int K, int WPB, int LIM>
__global__ __launch_bounds__(64 * WPB) void k_step2d_tb2(StepArgs a) {
    const LaunchView lv = launch_view(a);
    const int blk = xcd_work((int)blockIdx.x, (lv.n_items + WPB - 1) / WPB);
    const int wave = blk * WPB + uniform((int)threadIdx.x >> 6);
    if (blk < 0 || wave >= lv.n_items) return;
    const int lane = (int)threadIdx.x & 63;
    const Item2 it = ((const Item2*)a.items)[wave];
    const Inst in = a.inst[it.inst];
    const int64_t left = in.nsteps - lv.step_base;
    if (left <= 0) return;
    const int kk = a.ksteps < K ? a.ksteps : K;
    const int ns = left < (int64_t)kk ? (int)left : kk;
    const bool yw = it.j0 <= 0 || it.j0 + 64 >= in.ny;
    if (yw) pipeline2o2<PREC, K, true, LIM>(a, it, in, lane, ns);
    else pipeline2o2<PREC, K, false, LIM>(a, it, in, lane, ns);
}

template <int PREC, int WPB, int LIM>
static hipError_t launch2_tb2_w(int k, int nblocks, hipStream_t s, const StepArgs& a) {
    const dim3 grid(nblocks), block(64 * WPB);
    switch (k) {
    case 2: hipLaunchKernelGGL((k_step2d_tb2<PREC, 2, WPB, LIM>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int PREC>
static hipError_t launch2_tb2(int k, int lim, int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    if (lim == LIM_MC)
        return wpb == 1 ? launch2_tb2_w<PREC, 1, LIM_MC>(k, nblocks, s, a) : launch2_tb2_w<PREC, 4, LIM_MC>(k, nblocks, s, a);
    return wpb == 1 ? launch2_tb2_w<PREC, 1, LIM_MINMOD>(k, nblocks, s, a)
                    : launch2_tb2_w<PREC, 4, LIM_MINMOD>(k, nblocks, s, a);
}

template <int PREC, int WPB>
static hipError_t launch2_tb_w(int k, int nblocks, hipStream_t s, const StepArgs& a) {
    const dim3 grid(nblocks), block(64 * WPB);
    switch (k) {
    case 2: hipLaunchKernelGGL((k_step2d_tb<PREC, 2, WPB>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_step2d_tb<PREC, 3, WPB>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_step2d_tb<PREC, 4, WPB>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_step2d_tb<PREC, 8, WPB>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int PREC>
static hipError_t launch2_tb(int k, int wpb, int nblocks, hipStream_t s, const StepArgs& a) {
    return wpb == 1 ? launch2_tb_w<PREC, 1>(k, nblocks, s, a) : launch2_tb_w<PREC, 4>(k, nblocks, s, a);
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

bool step2d_supported(int prec, int order, int k) {
    if (k == 1) return true;
    if (order == 1) return k == 2 || k == 3 || k == 4 || k == 8;
    return (prec == PREC_C32 || prec == PREC_C32K) && k == 2;
}

hipError_t launch_step2d(int prec, int order, int lim, int k, int wpb, int n_items, hipStream_t s,
                         const StepArgs& a) {
    const int nblocks = xcd_grid((n_items + wpb - 1) / wpb);
    if (nblocks <= 0) return hipSuccess;
    if (k != 1 && order == 2) {
        if (prec == PREC_C32) return launch2_tb2<PREC_C32>(k, lim, wpb, nblocks, s, a);
        if (prec == PREC_C32K) return launch2_tb2<PREC_C32K>(k, lim, wpb, nblocks, s, a);
        return hipErrorInvalidValue;
    }
    if (k != 1) {
        if (order != 1 || lim != LIM_MINMOD) return hipErrorInvalidValue;
        switch (prec) {
        case PREC_F64: return launch2_tb<PREC_F64>(k, wpb, nblocks, s, a);
        case PREC_REF32: return launch2_tb<PREC_REF32>(k, wpb, nblocks, s, a);
        case PREC_C32: return launch2_tb<PREC_C32>(k, wpb, nblocks, s, a);
        case PREC_C32K: return launch2_tb<PREC_C32K>(k, wpb, nblocks, s, a);
        default: return hipErrorInvalidValue;
        }
    }
    switch (prec) {
    case PREC_F64: return launch2_p<PREC_F64>(order, lim, wpb, nblocks, s, a);
    case PREC_REF32: return launch2_p<PREC_REF32>(order, lim, wpb, nblocks, s, a);
    case PREC_C32: return launch2_p<PREC_C32>(order, lim, wpb, nblocks, s, a);
    case PREC_C32K: return launch2_p<PREC_C32K>(order, lim, wpb, nblocks, s, a);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace shll
