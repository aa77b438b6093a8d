/*
 * shll_oracle.h -- CPU restatement of the SHLL finite-volume time-step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library (libshll_amd.so).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * links against it or falls back to it.
 *
 * It restates, in independent code, the arithmetic of
 *   base-c/base_shll.c          (1D, first order, reflective walls)
 *   base-c/base_shll_2d.c       (2D, first order, reflective walls)
 *   base-c/2nd_order_base_shll.c(2D, second order minmod, outflow)
 *   base-omp/2nd_order_base_shll.c (the same with the MC limiter, alpha = 1.25)
 * and a derived 1D second-order scheme (the x sweep of the 2D second-order
 * file on the 3-variable system, i.e. the 2D file run on y-uniform data with
 * v = 0).  It keeps the reference's unfused pass structure (one full-array
 * pass per stage, temporaries for the neighbour gathers) so that it can also
 * be timed as the CPU baseline, serially or with one OpenMP worksharing loop
 * per pass (the base-omp pattern).
 *
 * Arithmetic modes:
 *   ORC_REF32 : float storage, with the double promotions the reference's
 *               double literals and double sqrt() cause.  Pinned bitwise to
 *               the compiled reference (tests/golden).
 *   ORC_F64   : the same association with every value in double.  This is
 *               the FP64 oracle.
 *   ORC_C32   : float flux arithmetic with a compensated (hi, lo) float state
 *               (TwoSum update).  Restates the product's FP32-compensated
 *               kernels so they can be checked bitwise; its accuracy is
 *               measured against ORC_F64.
 *   ORC_C32K  : ORC_C32 with the face terms added to each cell by one Kahan
 *               (compensated summation) step instead of the TwoSum chain.
 *   ORC_F64F  : 1D only: the compensated modes' face-term arrangement in double
 *               (split fluxes scaled by DT/DX, 1/sqrt(gamma R T) in place of a
 *               square root and a division), the face terms added to the FP64
 *               state directly.  Restates the product's fast FP64 kernel, which
 *               agrees with it to round-off (its 1/sqrt is faithful, not correctly
 *               rounded); measured against ORC_F64.
 *
 * Parity pinning status (which check fixes which function) is listed per
 * function in shll_oracle.c.
 */
#ifndef SHLL_ORACLE_H
#define SHLL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F64 = 0, ORC_REF32 = 1, ORC_C32 = 2, ORC_C32K = 3, ORC_F64F = 4 };
enum { ORC_BC_REFLECT = 0, ORC_BC_OUTFLOW = 1 };
/* Slope limiters of the second-order scheme:
 *   ORC_LIM_MINMOD : minmod of the one-sided differences (base-c/2nd_order_base_shll.c:268-276)
 *   ORC_LIM_MC     : minmod(central difference / 2, theta * minmod(one-sided differences)),
 *                    the monotonized-central family (base-omp/2nd_order_base_shll.c:317-325,385-392) */
enum { ORC_LIM_MINMOD = 0, ORC_LIM_MC = 1 };

typedef struct {
    int32_t dim;    /* 1 or 2 */
    int32_t order;  /* 1 or 2 */
    int32_t bc;     /* ORC_BC_REFLECT (first-order files) / ORC_BC_OUTFLOW (second-order file) */
    int32_t nx;     /* cells in x */
    int32_t ny;     /* cells in y (1 in 1D) */
    int32_t limiter;/* second order: ORC_LIM_MINMOD or ORC_LIM_MC */
    double lx;      /* domain length in x (reference: L = 1) */
    double ly;      /* domain length in y (reference: H = 1) */
    double theta;   /* ORC_LIM_MC weight (base-omp/2nd_order_base_shll.c:44: alpha = 1.25) */
} orc_problem;

/* Number of variables: 3 in 1D (rho, rho*u, E), 4 in 2D (rho, rho*u, rho*v, E). */
int orc_nvar(const orc_problem *pb);

/* Number of steps the reference's float time loop takes:
 *   float time = 0; while (time < t_end) { ...; time += DT; }
 * with DT = (CFL/(R+1))*DX computed in float (base_shll.c:83, :200-218). */
int64_t orc_steps_to_time(const orc_problem *pb, float t_end);

/* Time step constants of each mode, as doubles (diagnostics / tests). */
void orc_constants(const orc_problem *pb, int mode, double out[6] /* gamma,R,cv,dt,dt/dx,dt/dy */);

/*
 * Run nsteps of the scheme from primitive initial data.
 *   prim_in  : SoA doubles, nvar planes of nx*ny: rho, u, (v), T  (index i*ny+j)
 *              For ORC_REF32 every value must be exactly representable in float.
 *   prim_out : SoA doubles, same layout: rho, u, (v), T at the end
 *              (REF32: the float arrays widened; C32: FP64 primitives of hi+lo)
 *   cons_out : SoA doubles, nvar planes of conserved variables
 *              (C32: (double)hi + (double)lo), may be NULL
 *   nthreads : 1 = serial; >1 = one OpenMP worksharing loop per pass
 *   elapsed_s: wall time of the step loop only (CLOCK_MONOTONIC), may be NULL
 * Returns 0 on success, nonzero on invalid arguments / allocation failure.
 */
int orc_run(const orc_problem *pb, int mode, const double *prim_in, int64_t nsteps,
            double *prim_out, double *cons_out, int nthreads, double *elapsed_s);

/* C32 / C32K only: also return the raw hi and lo float planes (nvar*nx*ny each). */
int orc_run_c32_raw(const orc_problem *pb, int mode, const double *prim_in, int64_t nsteps,
                    float *hi_out, float *lo_out);

/* Stage evaluation on whole arrays: the reference's Compute_U_from_P, Compute_P_from_U and
 * Compute_F_from_P applied once to prim_in (base_shll.c:74-177).  Any output may be NULL.
 *   cons : nvar planes (C32: hi+lo),  prim : nvar planes after the P<-U pass,
 *   a    : sound speed (one plane; C32: of the hi state, in float),
 *   fp/fm: x split fluxes (nvar planes), hp/hm: y split fluxes (2D only). */
int orc_stage(const orc_problem *pb, int mode, const double *prim_in, double *cons, double *prim,
              double *a, double *fp, double *fm, double *hp, double *hm);

/* minmod of the second-order file, in float (REF32) or double (F64). */
double orc_minmod(int mode, double l, double r);

/* Test switch of the F64F mode (process-wide; 0, the default, leaves it as restated):
 * move its 1/sqrt(gamma R T) by a pseudo-random number of units in the last place between
 * -ulps and +ulps, chosen by a hash of the argument (so runs stay deterministic).  It
 * models the kernels' faithful reciprocal square root, which lies within 2 ulp of the
 * correctly rounded value, to bound the differences that causes. */
void orc_f64f_rsq_nudge(int ulps);
/* Limited slope of an interior cell from the values q at c-1, c, c+1 (ORC_LIM_*). */
double orc_slope(int mode, int limiter, double theta, double ql, double qc, double qr);

/* ---- initial conditions (primitive SoA doubles, exact floats) ---- */
/* Sod tube: cells with i < x0*N get (rhoL,uL,TL) else (rhoR,uR,TR).
 * The reference's test is  (i < 0.5*N)  in double (base_shll.c:55). */
void orc_ic_tube(int n, double x0, double rhoL, double uL, double TL,
                 double rhoR, double uR, double TR, double *prim);
/* Square implosion of base_shll_2d.c:93-103 (rho=1 inside (0.2,0.8)^2 by index test, 10 outside). */
void orc_ic_implosion(int nx, int ny, double *prim);
/* Four shocks of 2nd_order_base_shll.c:133-148. */
void orc_ic_four_shocks(int nx, int ny, double *prim);
/* Four contacts ("Configuration 6") of base-omp/2nd_order_base_shll.c:149-167. */
void orc_ic_four_contacts(int nx, int ny, double *prim);
/* Circle: rho_in for (x-cx)^2+(y-cy)^2 < r^2 at cell centres ((i+.5)/nx*lx,(j+.5)/ny*ly). */
void orc_ic_circle(int nx, int ny, double lx, double ly, double cx, double cy, double r,
                   double rho_in, double T_in, double rho_out, double T_out, double *prim);

#ifdef __cplusplus
}
#endif
#endif
