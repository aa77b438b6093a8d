// shll_device.h -- device arithmetic of the SHLL step for gfx950.
//
// The precisions share one set of templates:
//   PREC_F64   : every value double, the reference association (FP64 oracle form)
//   PREC_REF32 : float values with the reference's double promotions, so results are
//                bitwise those of the reference C on x86-64
//   PREC_C32   : float arithmetic, (hi, lo) float state, conservative face update
//                with TwoSum (the FP32-compensated form)
//   PREC_C32K  : PREC_C32 with one Kahan step per cell for the update
//   PREC_F64F  : FP64 values in the compensated forms' arrangement (fused multiply-adds,
//                1/rho and 1/a from faithful Newton sequences, face-term update); 1D
//
// The library is compiled with -ffp-contract=off: every + - * below is one IEEE-754
// operation in source order (fused multiply-adds only where written), and division
// and square root are correctly rounded (div_rn, sqrt_rn), so the F64 and REF32
// kernels reproduce the CPU results bit for bit (F64F alone uses faithful sequences).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shll_types.h"

namespace shll {

// ---------------------------------------------------------------------------
// cross-lane moves inside one wavefront (DPP wave_shr:1 / wave_shl:1)
// from_left(x)  : the value x of lane-1 (lane 0 keeps its own)
// from_right(x) : the value x of lane+1 (lane 63 keeps its own)
// Every lane of the wave executes these with the same control flow.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int dpp_from_left(int x) {
    return __builtin_amdgcn_update_dpp(x, x, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_from_right(int x) {
    return __builtin_amdgcn_update_dpp(x, x, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float from_left(float x) {
    return __int_as_float(dpp_from_left(__float_as_int(x)));
}
__device__ __forceinline__ float from_right(float x) {
    return __int_as_float(dpp_from_right(__float_as_int(x)));
}
__device__ __forceinline__ double from_left(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(dpp_from_left(hi), dpp_from_left(lo));
}
__device__ __forceinline__ double from_right(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(dpp_from_right(hi), dpp_from_right(lo));
}

// from_left_z(x) / from_right_z(x): as from_left / from_right, except that lane 0 (lane 63)
// reads 0 -- for values whose edge lanes are don't-cares; without an "old" operand the
// shift can be folded into the instruction that uses it
__device__ __forceinline__ float from_left_z(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right_z(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double from_left_z(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right_z(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// from_left_or(x, e) / from_right_or(x, e): as from_left / from_right, except that lane 0
// (lane 63) takes e, a value the lane brings from outside the wavefront
__device__ __forceinline__ float from_left_or(float x, float e) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(e), __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right_or(float x, float e) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(e), __float_as_int(x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double from_left_or(double x, double e) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(e), __double2loint(x), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(e), __double2hiint(x), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right_or(double x, double e) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(e), __double2loint(x), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(e), __double2hiint(x), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// wave-uniform value (moves it to a scalar register)
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// ---------------------------------------------------------------------------
// Correctly rounded division and square root.  The fast paths (div_fast, sqrt_fast)
// are the instruction sequences the compiler emits for IEEE '/', sqrtf and sqrt
// (-fhip-fp32-correctly-rounded-divide-sqrt; FP64 is always IEEE) minus their range
// handling -- the exponent pre-scaling of v_div_scale / v_div_fmas, the special-value
// pass of v_div_fixup, the square roots' pre-scaling of tiny radicands and their
// pass-through of 0 and inf.  With operands and results of moderate size (no
// intermediate leaves the normal range) they give the same bits as the operators.
// div_rn and sqrt_rn check the result or radicand (a quotient of at least 2^-64 in
// float or 2^-900 in double, a radicand of at least 2^-90 / 2^-760) and hand every
// other case -- zero, tiny, denormal, NaN -- to the operators themselves, so they
// are always bitwise '/' and sqrt.  The compensated scheme calls the fast paths
// directly: its divisors (density, sound speed) and radicand (gamma*R*T) are of order
// one, and its one possibly tiny quotient, the Mach number M, only enters as M/2 +-
// 1/2 and 1 - M^2, which absorb any last-place difference of a tiny M.
// ---------------------------------------------------------------------------
// rcp_fast(b): 1/b correctly rounded -- v_rcp_f32 (within 1 ulp) and one Newton step with
// fused multiply-adds.  Checked against IEEE 1.0f/b on the MI355X for every normal float b
// whose reciprocal is normal (2^23 mantissas x 252 exponents, no difference).
__device__ __forceinline__ float rcp_fast(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}
// a/b from the correctly rounded reciprocal y: q = a*y is within one ulp of a/b, the
// residual a - b*q is exact in one fma, and q + residual*y is the correctly rounded
// quotient (Markstein's theorem; also checked against IEEE a/b on 2^31 random pairs)
__device__ __forceinline__ float div_fast(float a, float b) {
    const float y = rcp_fast(b);
    const float q = a * y;
    const float d = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(d, y, q);
}
__device__ __forceinline__ double div_fast(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);            // two Newton steps on 1/b
    const double q = a * r;
    const double d = __builtin_fma(-b, q, a);
    return __builtin_fma(d, r, q);
}
__device__ __forceinline__ float sqrt_fast(float x) {
    float s = __builtin_amdgcn_sqrtf(x);   // within 1 ulp
    const float sdn = __int_as_float(__float_as_int(s) - 1);
    const float sup = __int_as_float(__float_as_int(s) + 1);
    const float rdn = __builtin_fmaf(-sdn, s, x);
    const float rup = __builtin_fmaf(-sup, s, x);
    s = rdn <= 0.0f ? sdn : s;
    return rup > 0.0f ? sup : s;
}
__device__ __forceinline__ double sqrt_fast(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;         // sqrt(x) and 1/(2 sqrt(x)) estimates
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float div_rn(float a, float b) {
    float q = div_fast(a, b);
    if (__builtin_expect(!(fabsf(q) >= 0x1p-64f), 0)) q = a / b;
    return q;
}
__device__ __forceinline__ double div_rn(double a, double b) {
    double q = div_fast(a, b);
    if (__builtin_expect(!(fabs(q) >= 0x1p-900), 0)) q = a / b;
    return q;
}
__device__ __forceinline__ double sqrt_rn(double x) {
    return __builtin_expect(!(x >= 0x1p-760), 0) ? sqrt(x) : sqrt_fast(x);
}

// Faithful FP64 reciprocal and reciprocal square root (the F64F form): the hardware seeds
// (v_rcp_f64, v_rsq_f64; relative error about 2^-24, measured) and two Newton steps each.
//   rcp_nr2(b): r <- r + r(1 - b r); equal to the correctly rounded 1/b on each of 8e8
//               random arguments with exponents -1000..1000 (MI355X)
//   rsq_nr2(x): y <- y (3/2 - (x/2) y^2); within 2 ulp of 1.0/sqrt(x) on the same arguments
// Neither is correctly rounded by construction, and the seeds are the hardware's, so the
// F64F results are reproduced on a CPU to within round-off, not bit for bit.  A zero argument
// gives NaN (not inf); negative and NaN arguments give NaN.
__device__ __forceinline__ double rcp_nr2(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double rsq_nr2(double x) {
    const double nhx = -0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * __builtin_fma(nhx, y * y, 1.5);
    return y * __builtin_fma(nhx, y * y, 1.5);
}

// ---------------------------------------------------------------------------
// Reference association (F64 and REF32).  RT is the storage/arithmetic type;
// (double) conversions sit where the reference's double literals and its double
// sqrt() put them (base_shll.c:135-139, :172-175; base_shll_2d.c:313-317).
// For RT = double every conversion is the identity.
// ---------------------------------------------------------------------------
template <class RT>
struct RefArith {
    // Compute_P_from_U, 1D.  ke = 0: (0.5*u)*u (base_shll.c:174); ke = 1: 0.5*(u*u)
    // (the 2D files' association, used by the derived 1D second-order scheme).
    __device__ static __forceinline__ void prim1(RT u0, RT u1, RT u2, int ke, RT cv, RT gR, RT& vx, RT& T,
                                                 RT& a) {
        vx = div_rn(u1, u0);
        double kin = ke == 0 ? (0.5 * (double)vx) * (double)vx : 0.5 * (double)(vx * vx);
        T = (RT)div_rn((double)div_rn(u2, u0) - kin, (double)cv);
        a = (RT)sqrt_rn((double)(gR * T));
    }
    // Compute_P_from_U, 2D (base_shll_2d.c:313-317)
    __device__ static __forceinline__ void prim2(RT u0, RT u1, RT u2, RT u3, RT cv, RT gR, RT& vx, RT& vy,
                                                 RT& T, RT& a) {
        vx = div_rn(u1, u0);
        vy = div_rn(u2, u0);
        T = (RT)div_rn((double)div_rn(u3, u0) - 0.5 * (double)(vx * vx + vy * vy), (double)cv);
        a = (RT)sqrt_rn((double)(gR * T));
    }
    // Z invariants (base_shll.c:137-139)
    __device__ static __forceinline__ void zinv(RT M, RT a, RT& z1, RT& z2, RT& z3) {
        z1 = (RT)(0.5 * ((double)M + 1.0));
        z2 = (RT)((0.5 * (double)a) * (1.0 - (double)(M * M)));
        z3 = (RT)(0.5 * ((double)M - 1.0));
    }
    // minmod (2nd_order_base_shll.c:191-201): product sign test, ties -> right
    __device__ static __forceinline__ RT mm(RT l, RT r) {
        RT pick = fabs((double)l) < fabs((double)r) ? l : r;
        return (double)(l * r) < 0.0 ? (RT)0 : pick;
    }
    // 0.5*(difference) with the double literal of base-omp/2nd_order_base_shll.c:317,
    // rounded to RT as minmod's argument
    __device__ static __forceinline__ RT half(RT x) { return (RT)(0.5 * (double)x); }
};

// FP32-compensated arithmetic: float, with explicit fused multiply-adds where a
// product feeds a sum (one rounding instead of two, and one instruction).
struct C32Arith {
    __device__ static __forceinline__ void prim1(float u0, float u1, float u2, float rcv, float gR, float& vx,
                                                 float& T, float& a) {
        float inv = rcp_fast(u0);
        vx = u1 * inv;
        float e = u2 * inv;
        T = __builtin_fmaf(-0.5f * vx, vx, e) * rcv;  // (e - vx^2/2) / cv
        a = sqrt_fast(gR * T);
    }
    __device__ static __forceinline__ void prim2(float u0, float u1, float u2, float u3, float rcv, float gR,
                                                 float& vx, float& vy, float& T, float& a) {
        float inv = rcp_fast(u0);
        vx = u1 * inv;
        vy = u2 * inv;
        float e = u3 * inv;
        T = __builtin_fmaf(-0.5f, __builtin_fmaf(vx, vx, vy * vy), e) * rcv;
        a = sqrt_fast(gR * T);
    }
    // z1 = (M+1)/2, z2 = (a/2)(1-M^2), z3 = (M-1)/2.  z1 and z3 are RN(M/2 +- 1/2): M/2 is
    // exact (or, below 2^-125, off by less than 2^-150, which the sum rounds away), so one
    // fused multiply-add gives the bits of the product-then-sum
    __device__ static __forceinline__ void zinv(float M, float a, float& z1, float& z2, float& z3) {
        z1 = __builtin_fmaf(0.5f, M, 0.5f);
        z2 = (0.5f * a) * __builtin_fmaf(-M, M, 1.0f);
        z3 = __builtin_fmaf(0.5f, M, -0.5f);
    }
    // minmod as the median of (l, r, 0): the smaller magnitude when the signs agree,
    // zero otherwise -- one v_med3_f32.  Unlike the reference's product test it also
    // clips slopes of opposite sign whose float product underflows, and the sign of a
    // zero result is unspecified (a zero slope only ever meets a nonzero or +0 face
    // flux, where the sign of the zero cannot change the sum).
    __device__ static __forceinline__ float mm(float l, float r) { return __builtin_amdgcn_fmed3f(l, r, 0.0f); }
    __device__ static __forceinline__ float half(float x) { return 0.5f * x; }
};

// FP64 in the compensated scheme's arrangement (PREC_F64F): the limiter of the C32 forms in
// double (the split fluxes are flux1_f64f in shll_flux.h)
struct F64FArith {
    // minmod as the median of (l, r, 0): the smaller magnitude when the signs agree, zero
    // otherwise (sign of a zero result unspecified, as in C32Arith::mm)
    __device__ static __forceinline__ double mm(double l, double r) {
        return __builtin_fmax(__builtin_fmin(l, r), __builtin_fmin(__builtin_fmax(l, r), 0.0));
    }
    __device__ static __forceinline__ double half(double x) { return 0.5 * x; }
};

// the same split in FP64: U*Z2 once, two fused multiply-adds
__device__ __forceinline__ void split_f64f(double f, double u, double z1, double z2, double z3, double& fp,
                                           double& fm) {
    const double uz = u * z2;
    fp = __builtin_fma(f, z1, uz);
    fm = __builtin_fma(-f, z3, -uz);
}

// Limited slope of a split flux from its values at c-1, c, c+1 (interior cells):
//   LIM_MINMOD  mm(qc - ql, qr - qc)                                 2nd_order_base_shll.c:268-276
//   LIM_MC      mm((qr - ql)/2, theta * mm(qc - ql, qr - qc))        base-omp/2nd_order_base_shll.c:317-325
template <class A, int LIM, class RT>
__device__ __forceinline__ RT limited_slope(RT ql, RT qc, RT qr, RT theta) {
    const RT s = A::mm(qc - ql, qr - qc);
    if constexpr (LIM == LIM_MINMOD) return s;
    else return A::mm(A::half(qr - ql), theta * s);
}

// TwoSum (Knuth, TAOCP vol. 2, 4.2.2): s + e == a + b exactly
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Fast2Sum (Dekker 1971): s + e == a + b exactly when |a| >= |b| (the renormalisation
// of a (hi, lo) pair, whose lo is below an ulp of hi)
__device__ __forceinline__ void fast_two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    e = b - (s - a);
}

// SHLL split of one flux component: F+ = F*Z1 + U*Z2, F- = -F*Z3 - U*Z2 (base_shll.c:149-157)
template <class RT>
__device__ __forceinline__ void split(RT f, RT u, RT z1, RT z2, RT z3, RT& fp, RT& fm) {
    fp = f * z1 + u * z2;
    fm = -f * z3 - u * z2;
}

// the same split in the compensated scheme's float arithmetic: U*Z2 once, two FMAs
__device__ __forceinline__ void split_c32(float f, float u, float z1, float z2, float z3, float& fp, float& fm) {
    const float uz = u * z2;
    fp = __builtin_fmaf(f, z1, uz);
    fm = __builtin_fmaf(-f, z3, -uz);
}

}  // namespace shll
