/*
 * shll.h -- C ABI of the MI355X SHLL time-step library (libshll_amd.so).
 *
 * Drop-in for the stage routines of the reference solvers
 *   base-c/base_shll.c          (1D, first order, reflective walls)
 *   base-c/base_shll_2d.c       (2D, first order, reflective walls)
 *   base-c/2nd_order_base_shll.c(2D, second order, outflow)
 *   base-omp/2nd_order_base_shll.c (the same with the MC limiter; its OpenMP worksharing
 *                                    is replaced by the GPU launch)
 * whose only caller is main() (base_shll.c:198-225).  The reference routines are
 * void functions over file-scope float arrays; here the state lives on the GPU
 * behind an opaque context, and each call maps to the routines it replaces (cited
 * per declaration).  Plain pointers and sizes only; no GPU or torch types.
 *
 * Conventions
 *   - Every call returns an shll_status; SHLL_OK = 0.  shll_last_error() gives a
 *     thread-local message for the last failure.  Nothing aborts.
 *   - Host arrays are owned by the caller and copied; device memory, stream and
 *     graphs are owned by the context.
 *   - Primitive arrays are SoA doubles: rho, u, [v,] T  (T is temperature, p = rho*R*T,
 *     base_shll.c:141), index i*ny + j with y contiguous (base_shll_2d.c:146-221).
 *   - A context is used by one host thread at a time.  Calls that return data
 *     synchronise its stream.
 *   - No CPU fallback: without a usable GPU the calls return SHLL_EDEVICE.
 */
#ifndef SHLL_H
#define SHLL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SHLL_OK = 0,
    SHLL_EINVAL = 1,      /* invalid argument */
    SHLL_ENOMEM = 2,      /* host or device allocation failed */
    SHLL_EDEVICE = 3,     /* HIP runtime error / no device */
    SHLL_ENONFINITE = 4,  /* state holds a non-finite or non-positive density/energy */
    SHLL_ESTATE = 5       /* call not valid in the object's state */
} shll_status;

typedef enum {
    SHLL_F64 = 0,    /* FP64 state and arithmetic, reference association (FP64 oracle form) */
    SHLL_REF32 = 1,  /* the reference's own float arithmetic with its double promotions */
    SHLL_C32 = 2,    /* FP32 fluxes, compensated (hi, lo) FP32 state, conservative face update
                        added exactly (TwoSum): conserved sums exact to ~1e-14 */
    SHLL_C32K = 3,   /* as SHLL_C32 but the face terms enter each cell through one Kahan
                        update (3x fewer update instructions); same accuracy against FP64,
                        conserved sums to ~1e-10 */
    SHLL_F64F = 4    /* 1D only: FP64 state and fluxes in SHLL_C32K's face-term arrangement
                        (fused multiply-adds; 1/rho and 1/sqrt(gamma R T) from faithful
                        Newton sequences in place of divisions and a square root; fluxes
                        scaled by DT/DX); not the reference's association, so within FP64
                        round-off of SHLL_F64 rather than bitwise (L1(rho) ~1e-12 on
                        second-order tubes) */
} shll_precision;

typedef enum {
    SHLL_BC_REFLECT = 0,  /* base_shll.c:93-110, base_shll_2d.c:150-216 */
    SHLL_BC_OUTFLOW = 1   /* 2nd_order_base_shll.c:214-256,280-323 */
} shll_bc;

/* Slope limiter of the second-order scheme. */
typedef enum {
    SHLL_LIM_MINMOD = 0,  /* minmod of the one-sided differences (2nd_order_base_shll.c:268-276,336-343) */
    SHLL_LIM_MC = 1       /* minmod(0.5*(f[i+1]-f[i-1]), theta*minmod(one-sided differences)),
                             base-omp/2nd_order_base_shll.c:44,317-325,385-392 (theta = alpha = 1.25) */
} shll_limiter;

/* Problem description: the reference's compile-time constants made runtime
 * (base_shll.c:16-26, base_shll_2d.c:21-35, base-omp/2nd_order_base_shll.c:29-44).
 * gamma = 1.4, R = 1, CFL = 0.25 as in the reference; DT = (CFL/(R+1))*DX.  A
 * zero-initialised tail (limiter, reserved, theta) selects the minmod limiter. */
typedef struct {
    int32_t dim;        /* 1 or 2 */
    int32_t order;      /* 1 (base_shll*.c) or 2 (2nd_order_base_shll.c) */
    int32_t bc;         /* shll_bc */
    int32_t precision;  /* shll_precision */
    int64_t nx;         /* cells in x (>= 2) */
    int64_t ny;         /* cells in y (1 in 1D, >= 2 in 2D) */
    double lx;          /* domain length in x (reference L = 1) */
    double ly;          /* domain length in y (reference H = 1) */
    int32_t limiter;    /* shll_limiter (order 2 only; SHLL_LIM_MINMOD for order 1) */
    int32_t reserved;   /* 0 */
    double theta;       /* SHLL_LIM_MC weight in [1, 2]; the reference's is 1.25 */
} shll_problem;

typedef struct shll_ctx shll_ctx;
typedef struct shll_batch shll_batch;

/* Error norms of one state against another (FP64, from the primitives). */
typedef struct {
    double l1_rho, l1_u, l1_p;     /* sum |d| * cell volume over the domain */
    double linf_rho, linf_p;       /* max |d| */
    double max_rel_rho, max_rel_p; /* max |d| / |ref|, elementwise */
    double max_du;                 /* max |du| (and |dv| in 2D) */
    double max_abs_u_ref;          /* max |u| (and |v|) of the reference */
    double nonfinite;              /* count of non-finite entries */
} shll_errors;

/* ---- single instance ------------------------------------------------------ */

/* Replaces Allocate_and_Init_Memory + Compute_U_from_P + Compute_P_from_U
 * (base_shll.c:28-61,74-85,161-177; base_shll_2d.c:37-137; 2nd_order_base_shll.c:45-189).
 * prim_soa: the initial primitives (the caller's replacement of the reference's
 * hard-coded initial data, e.g. from shll_ic_*). */
shll_status shll_init(const shll_problem *pb, const double *prim_soa, shll_ctx **out);

/* Replaces nsteps iterations of the loop body {Compute_F_from_P; Update_U_from_F;
 * Compute_P_from_U} (base_shll.c:208-218).  Synchronous. */
shll_status shll_step(shll_ctx *ctx, int64_t nsteps);

/* The reference loop itself: float time = 0; while (time < t_end) {step; time += DT;}
 * (base_shll.c:200-218).  *steps_out (may be NULL) receives NO_STEPS; the final state
 * is checked (SHLL_ENONFINITE). */
shll_status shll_run_to_time(shll_ctx *ctx, float t_end, int64_t *steps_out);

/* Step count of the reference's float time loop for a problem (no GPU needed); -1 for a
 * NULL or invalid problem, or when the float sum time += DT stops growing before it reaches
 * t_end (the reference's loop would not end). */
int64_t shll_steps_to_time(const shll_problem *pb, float t_end);

/* Replaces the arrays Save_Results prints (p0..p3 after Compute_P_from_U,
 * base_shll.c:180-192).  out_soa: nvar*nx*ny doubles (rho, u, [v,] T). */
shll_status shll_get_primitives(shll_ctx *ctx, double *out_soa);
/* Conserved variables (u0..u3) as doubles; C32 gives (double)hi + (double)lo. */
shll_status shll_get_conserved(shll_ctx *ctx, double *out_soa);
/* The state in its storage format: F64 and F64F nvar*N doubles; REF32 nvar*N floats;
 * C32 nvar*N hi floats then nvar*N lo floats.  bytes must equal shll_state_bytes(). */
size_t shll_state_bytes(const shll_ctx *ctx);
shll_status shll_get_state_raw(shll_ctx *ctx, void *out, size_t bytes);
/* Restart: replace the state by a buffer of shll_get_state_raw's layout (a checkpoint of a
 * context with the same problem).  Continuing from it gives bit for bit the results of the
 * uninterrupted run.  No reference counterpart (the reference has no checkpoint). */
shll_status shll_set_state_raw(shll_ctx *ctx, const void *in, size_t bytes);
/* Replaces Save_Results (base_shll.c:180-192, base_shll_2d.c:322-340): the same text
 * format, "%e\t%e\t%e\t%e\n" (x, rho, u, T) in 1D, six columns (x, y, rho, u, v, T) in 2D,
 * written from float values as the reference does. */
shll_status shll_write_results(shll_ctx *ctx, const char *path);
/* Replaces the ten-column Save_Results of omp-vectors/2nd_order_arm_sve.c:921-940: in 2D
 * x, y, rho, u, v, T, u0, u1, u2, u3 (primitives then conserved variables); in 1D the same
 * layout with seven columns (x, rho, u, T, u0, u1, u2).  Float values, "%e" tab-separated. */
shll_status shll_write_results_full(shll_ctx *ctx, const char *path);

/* Stage-level evaluation for checks against the reference's stage functions: from the
 * context's current state, the sound speed a of Compute_P_from_U and the split fluxes of
 * Compute_F_from_P (base_shll.c:130-177, base_shll_2d.c:239-319), computed exactly as the
 * step kernels compute them (compensated forms: from the hi word), as doubles.  a: nx*ny;
 * fp, fm (x) and hp, hm (y, 2D only): nvar*nx*ny each, SoA.  Any output may be NULL. */
shll_status shll_stage_fluxes(shll_ctx *ctx, double *a, double *fp, double *fm, double *hp, double *hm);

/* FP64 sums of the conserved variables times the cell volume (mass, momentum(s),
 * energy); out has nvar entries.  No reference counterpart (diagnostic). */
shll_status shll_conserved_sums(shll_ctx *ctx, double *out);
/* shll_conserved_sums without allocating anything.  No counterpart in the reference: it is
 * for callers that cannot allocate device or host memory (serving loops, code that forbids
 * hipMalloc).  Same kernel, grid and summation order as shll_conserved_sums, so the same bits;
 * the block partials go to a static device scratch of the library (256 blocks x 6 doubles =
 * 12 KiB, held by every GPU the library runs on until the process ends) and to the stack.
 * Calls that overlap on one GPU share that scratch and race: serialise them. */
shll_status shll_conserved_sums_noalloc(shll_ctx *ctx, double *out);
/* Scan the state for non-finite or non-positive density/energy: SHLL_OK or
 * SHLL_ENONFINITE; *bad_cells (may be NULL) receives the count. */
shll_status shll_check_state(shll_ctx *ctx, int64_t *bad_cells);
/* Error norms of ctx against ref (same dim/nx/ny; any precisions), FP64 on the GPU. */
shll_status shll_compare(shll_ctx *ctx, shll_ctx *ref, shll_errors *out);
/* Error norms against host primitives (nvar*N doubles). */
shll_status shll_compare_host(shll_ctx *ctx, const double *ref_prim_soa, shll_errors *out);

/* Replaces Free_Memory (base_shll.c:63-72). */
void shll_free(shll_ctx *ctx);

/* Launch shape of the step kernels.  cells_per_lane: 1D in {1,2,4,8}; 2D the owned rows
 * of a work item, the cells of its column each lane owns (2-4096; 0 or 1: 16, 32 or 64
 * rows for 1, 2 or more steps per launch), steps_per_launch >= 1 (temporal blocking; 2D:
 * 1, 2, 3, 4 or 8 in first order; 1, or 2 for SHLL_C32/SHLL_C32K in second order),
 * waves_per_block in {1,4} (1D, 4 with 2+ cells per lane: the block's wavefronts hold
 * one segment and share their edge fluxes), graph_steps: steps captured per HIP graph
 * (0 = no graph; at most 4096 launches per graph; a single instance's graphs are captured
 * when the shape is set).  shll_get_launch returns the values as set (a 2D 0 reads back as 1).
 * Results do not depend on the launch shape (bitwise). */
shll_status shll_set_launch(shll_ctx *ctx, int32_t cells_per_lane, int32_t steps_per_launch,
                            int32_t waves_per_block, int32_t graph_steps);
shll_status shll_get_launch(const shll_ctx *ctx, int32_t out[4]);
/* Time candidate launch shapes on this context's problem (state is restored) and
 * keep the fastest; the choice is cached per (device, kernel, shape class) in
 * $SHLL_TUNE_CACHE (if set).  Candidates: 1D cells per lane 1-8, 1-64 steps per launch,
 * 1 or 4 waves per block (and 256-4096 steps per launch when the whole tube fits one
 * segment or block, which then has no halo); 2D the valid steps per launch, 1 or 4
 * waves per block, work items of 3/4, 5/4 and 3/2 of the default rows, and shorter ones
 * (down to 1/8) when the default ones give fewer than 4 wavefronts per SIMD; graphs of
 * 128 (2D: 64) steps, or two launches if longer, and for a single instance also of 2048
 * (2D: 1024) steps and direct launches (graph_steps 0). */
shll_status shll_autotune(shll_ctx *ctx, int32_t trial_steps);

/* Average duration (ms) of the step kernel launches of the last shll_step /
 * shll_run_to_time / shll_batch_* call, measured with HIP events on the context's
 * stream (0 if not timed); enable with shll_set_timing. */
shll_status shll_set_timing(shll_ctx *ctx, int32_t on);
shll_status shll_last_timing(const shll_ctx *ctx, double *avg_kernel_ms, int64_t *launches, double *total_ms);

/* ---- batches of independent instances (tubes or 2D domains) --------------- */

/* n instances that share dim, order and precision; sizes, walls and step counts
 * differ.  prims[k]: the instance's primitives; nsteps[k]: its step count
 * (e.g. shll_steps_to_time).  All advance in one launch per steps_per_launch
 * steps over an (instance, tile) work list ordered by remaining work. */
shll_status shll_batch_init(const shll_problem *items, const double *const *prims, const int64_t *nsteps,
                            int32_t n, shll_batch **out);
/* Advance every unfinished instance by exactly min(nsteps, its remaining steps): whole
 * chunks of steps_per_launch-step launches replay HIP graphs, the last launch of a call
 * takes the steps that remain.  *max_remaining (may be NULL) receives the steps the longest
 * instance still has to take. */
shll_status shll_batch_advance(shll_batch *b, int64_t nsteps, int64_t *max_remaining);
/* Run every instance to its own step count. */
shll_status shll_batch_run(shll_batch *b);
shll_status shll_batch_get_primitives(shll_batch *b, int32_t k, double *out_soa);
shll_status shll_batch_get_conserved(shll_batch *b, int32_t k, double *out_soa);
/* FP64 conserved sums per instance: out[k*4 + v] (v < nvar), cell volume included. */
shll_status shll_batch_conserved_sums(shll_batch *b, double *out);
/* Per-instance census of non-finite / non-positive density or energy: bad_cells[k]
 * (may be NULL) and the number of instances with any (may be NULL). */
shll_status shll_batch_check_state(shll_batch *b, int64_t *bad_cells, int32_t *bad_instances);
/* Error norms of instance k of b against instance kref of ref (same shape), FP64 on the GPU. */
shll_status shll_batch_compare(shll_batch *b, int32_t k, shll_batch *ref, int32_t kref, shll_errors *out);
shll_status shll_batch_set_launch(shll_batch *b, int32_t cells_per_lane, int32_t steps_per_launch,
                                  int32_t waves_per_block, int32_t graph_steps);
shll_status shll_batch_get_launch(const shll_batch *b, int32_t out[4]);
shll_status shll_batch_autotune(shll_batch *b, int32_t trial_steps);
shll_status shll_batch_set_timing(shll_batch *b, int32_t on);
shll_status shll_batch_last_timing(const shll_batch *b, double *avg_kernel_ms, int64_t *launches, double *total_ms);
/* Total cell-updates the batch performs from start to finish. */
int64_t shll_batch_total_cell_updates(const shll_batch *b);
shll_status shll_batch_reset(shll_batch *b); /* back to the initial state */
void shll_batch_free(shll_batch *b);

/* ---- initial data (host; the reference hard-codes these in
 *      Allocate_and_Init_Memory) ---------------------------------------------- */

/* Tube: cells with i < x0*n get the left state (base_shll.c:54-60 with x0 = 0.5). */
void shll_ic_tube(int64_t n, double x0, double rhoL, double uL, double TL, double rhoR, double uR, double TR,
                  double *prim_soa);
/* base_shll_2d.c:93-103 square implosion. */
void shll_ic_implosion(int64_t nx, int64_t ny, double *prim_soa);
/* 2nd_order_base_shll.c:133-148 four shocks. */
void shll_ic_four_shocks(int64_t nx, int64_t ny, double *prim_soa);
/* base-omp/2nd_order_base_shll.c:149-167 four contacts ("Configuration 6"). */
void shll_ic_four_contacts(int64_t nx, int64_t ny, double *prim_soa);
/* Circle of (rho_in, T_in) for cell centres strictly inside radius r about (cx, cy). */
void shll_ic_circle(int64_t nx, int64_t ny, double lx, double ly, double cx, double cy, double r, double rho_in,
                    double T_in, double rho_out, double T_out, double *prim_soa);

/* Seeded batch generators (SplitMix64, counter-based: instance k depends only on
 * (seed, k)).  Tubes: N = round(exp(U[ln nmin, ln nmax])), rhoL/rhoR ~ U[2,20] (rhoR = 1),
 * TL/TR ~ U[0.5,2] (TR = 1), x0 ~ U[0.3,0.7], u = 0. */
typedef struct {
    int64_t n;
    double x0, rhoL, TL, rhoR, TR;
} shll_tube_spec;
void shll_gen_tube(uint64_t seed, int64_t k, int64_t nmin, int64_t nmax, shll_tube_spec *out);
/* 2D domains: side lengths log-uniform in [smin, smax], square cells of size 1/max(nx,ny)
 * (lx = nx*h, ly = ny*h), 1-4 circles with rho ~ U[2,20], T ~ U[0.5,2],
 * radius ~ U[0.05,0.3] * min(lx, ly), centres uniform inside the domain. */
typedef struct {
    int64_t nx, ny;
    double lx, ly;
    int32_t ncircles, pad;
    double cx[4], cy[4], r[4], rho[4], T[4];
} shll_domain_spec;
void shll_gen_domain(uint64_t seed, int64_t k, int64_t smin, int64_t smax, shll_domain_spec *out);
void shll_ic_domain(const shll_domain_spec *d, double *prim_soa);

const char *shll_last_error(void);
const char *shll_version(void);

#ifdef __cplusplus
}
#endif
#endif
