// shll_flux.h -- the split fluxes of one cell (Compute_P_from_U followed by
// Compute_F_from_P for that cell), shared by the step kernels and the stage kernel so
// that both evaluate exactly the same operations.
#pragma once

#include "shll_device.h"

namespace shll {

// storage / arithmetic type of a precision
template <int PREC> struct Real { using T = double; };
template <> struct Real<PREC_REF32> { using T = float; };
template <> struct Real<PREC_C32> { using T = float; };
template <> struct Real<PREC_C32K> { using T = float; };

// 1D, F64F: F+ and F- of the three variables scaled by sc, and the sound speed a
// (base_shll.c:134-157 in the compensated forms' arrangement).  The step kernels pass
// sc = DT/DX, so that their face term is G + d/2 with no product by DT/DX (the minmod and
// MC limiters commute with a positive scale); k_stage passes 1.  With R = 1:
//   p = rho R T = (E - (rho u) u / 2) R / cv       (no E/rho)
//   a^2 = gamma R T = gamma R p / rho,  y = 1/a     (rsq_nr2: no square root, no division)
//   M = u y,  a/2 = (a^2/2) y
//   Z1 = sc (M + 1)/2,  Z3 = sc (M - 1)/2,  Z2 = sc (a/2)(1 - M^2)
__device__ __forceinline__ void flux1_f64f(double u0, double u1, double u2, const Inst& in, double sc,
                                           double (&fp)[3], double (&fm)[3], double& a) {
    const double inv = rcp_nr2(u0);
    const double vx = u1 * inv;
    const double P = __builtin_fma(-0.5 * u1, vx, u2) * in.k.d_rcv;
    const double x = in.k.d_gR * (P * inv);
    const double y = rsq_nr2(x);
    const double M = vx * y, hs = 0.5 * sc;
    const double z1 = __builtin_fma(hs, M, hs), z3 = __builtin_fma(hs, M, -hs);
    const double z2 = ((hs * x) * y) * __builtin_fma(-M, M, 1.0);
    a = x * y;
    const double f1 = __builtin_fma(u1, vx, P), f2 = vx * (u2 + P);
    split_f64f(u1, u0, z1, z2, z3, fp[0], fm[0]);
    split_f64f(f1, u1, z1, z2, z3, fp[1], fm[1]);
    split_f64f(f2, u2, z1, z2, z3, fp[2], fm[2]);
}

// 1D: F+ and F- of the three variables and the sound speed a (base_shll.c:134-177;
// the second-order scheme's kinetic-energy association through in.ke).  U: the state
// (F64/REF32) or its hi word (C32/C32K).
template <int PREC, class RT>
__device__ __forceinline__ void flux1(RT u0, RT u1, RT u2, const Inst& in, RT (&fp)[3], RT (&fm)[3], RT& a) {
    RT vx, T, z1, z2, z3;
    if constexpr (PREC == PREC_F64F) {
        flux1_f64f(u0, u1, u2, in, 1.0, fp, fm, a);
    } else if constexpr (PREC == PREC_C32 || PREC == PREC_C32K) {
        C32Arith::prim1(u0, u1, u2, in.k.c_rcv, in.k.c_gR, vx, T, a);
        C32Arith::zinv(div_fast(vx, a), a, z1, z2, z3);
        const RT P = u0 * T;  // R = 1
        const RT f0 = u1, f1 = __builtin_fmaf(u1, vx, P), f2 = vx * (u2 + P);
        split_c32(f0, u0, z1, z2, z3, fp[0], fm[0]);
        split_c32(f1, u1, z1, z2, z3, fp[1], fm[1]);
        split_c32(f2, u2, z1, z2, z3, fp[2], fm[2]);
    } else {
        const RT cv = PREC == PREC_F64 ? (RT)in.k.d_cv : (RT)in.k.f_cv;
        const RT gR = PREC == PREC_F64 ? (RT)in.k.d_gR : (RT)in.k.f_gR;
        RefArith<RT>::prim1(u0, u1, u2, in.ke, cv, gR, vx, T, a);
        RefArith<RT>::zinv(div_rn(vx, a), a, z1, z2, z3);
        const RT P = u0 * T;  // (rho*R)*T with R = 1 (base_shll.c:17, :141)
        const RT f0 = u1, f1 = u1 * vx + P, f2 = vx * (u2 + P);
        split<RT>(f0, u0, z1, z2, z3, fp[0], fm[0]);
        split<RT>(f1, u1, z1, z2, z3, fp[1], fm[1]);
        split<RT>(f2, u2, z1, z2, z3, fp[2], fm[2]);
    }
}

// x and y split fluxes of one cell (Compute_P_from_U + Compute_F_from_P, base_shll_2d.c:239-319)
// and its sound speed a.  U: conserved (F64/REF32: the state; C32: hi).
template <int PREC, class RT>
__device__ __forceinline__ void fluxes2(const RT (&U)[4], const Inst& in, RT (&xp)[4], RT (&xm)[4], RT (&yp)[4],
                                        RT (&ym)[4], RT& a) {
    RT vx, vy, T, z1, z2, z3;
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

template <int PREC, class RT>
__device__ __forceinline__ void fluxes2(const RT (&U)[4], const Inst& in, RT (&xp)[4], RT (&xm)[4], RT (&yp)[4],
                                        RT (&ym)[4]) {
    RT a;
    fluxes2<PREC, RT>(U, in, xp, xm, yp, ym, a);
}

}  // namespace shll
