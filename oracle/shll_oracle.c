/*
 * shll_oracle.c -- CPU restatement of the SHLL time-step loop (see shll_oracle.h).
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the product.
 *
 * Build: gcc -O3 -ffp-contract=off -fopenmp (no -march, no -ffast-math), so every
 * float/double operation is one IEEE-754 operation in program order, like the
 * reference's x86-64 gcc -O3 build (base-c/makefile:2-4).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   REF32 u_from_p / p_from_u / flux / gather / update / minmod, all four schemes:
 *       bitwise against golden vectors written by the compiled reference
 *       (oracle/_ref, tests/golden): base_shll.c full Sod run (N=256, 410 steps),
 *       base_shll_2d.c implosion (256^2, 1/10/205 steps), 2nd_order_base_shll.c four
 *       shocks (256^2, 1/10 steps), plus injected random / circle / y-uniform Sod
 *       initial data.  The derived 1D second-order scheme is pinned by the 2D
 *       second-order reference run on y-uniform Sod data (every row equal).
 *   REF32 MC limiter: bitwise against base-omp/2nd_order_base_shll.c at its 1024^2
 *       size (four contacts, random, y-uniform data; built with clang, see Makefile).
 *   MC slope, every mode: equals the textbook minmod(theta*dl, (dl+dr)/2, theta*dr)
 *       on dyadic data, minmod at theta = 1; MC beats minmod in L1 on Sod.
 *   F64: same source text as REF32 (double instantiation); additionally pinned by
 *       invariants: rest state is a fixed point, F+ + F- = F, M=0 closed form,
 *       mirror symmetry of a mirrored Sod run, mass/energy conservation with
 *       reflective walls, momentum change = wall impulse, 2D run on y-uniform data
 *       equals the 1D run, L1 convergence to the exact Riemann solution.
 *   C32: no reference exists (it is this project's compensated scheme); pinned by
 *       the same invariants and by its distance to F64 (L1(rho) <= 1e-6).
 *   F64F: no reference exists (the compensated modes' face-term arrangement in double,
 *       1/sqrt(gamma R T) in place of a square root and a division); pinned by the same
 *       invariants (rest state, F+ + F- = F, M = 0 closed form, mirror symmetry,
 *       conservation, wall impulse, exact Riemann convergence) and by its distance to
 *       the F64 mode, which is FP64 round-off only: elementwise relative 1e-13 after
 *       short runs of either order and L1(rho) <= 1e-10 at t = 0.2 on second-order
 *       tubes, where a wrong sign, index or dropped term shows as 1e-4 or more.  Its
 *       1/sqrt moved by 0-2 ulp (orc_f64f_rsq_nudge, the kernels' faithful sequence)
 *       moves long runs by 2e-14-3e-14, the size of the kernels' distance to it.
 *   Initial conditions: bitwise against the reference's own initial arrays
 *       (golden step-0 dumps) for Sod, implosion and four shocks; the circle by a
 *       brute-force count of cells inside the radius.
 *   orc_steps_to_time: the published step counts (README.md:89-99, :201-204, :247-248).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "shll_oracle.h"

static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- two instantiations of the reference arithmetic ---- */
#define RT float
#define SFX(x) x##_r32
#include "oracle_scheme.inc"
#undef RT
#undef SFX

#define RT double
#define SFX(x) x##_f64
#include "oracle_scheme.inc"
#undef RT
#undef SFX

int orc_nvar(const orc_problem *pb) { return pb->dim == 1 ? 3 : 4; }

static int orc_check(const orc_problem *pb)
{
    if (!pb) return -1;
    if (pb->dim != 1 && pb->dim != 2) return -1;
    if (pb->order != 1 && pb->order != 2) return -1;
    if (pb->bc != ORC_BC_REFLECT && pb->bc != ORC_BC_OUTFLOW) return -1;
    if (pb->limiter != ORC_LIM_MINMOD && pb->limiter != ORC_LIM_MC) return -1;
    if (pb->limiter == ORC_LIM_MC && (pb->order != 2 || !(pb->theta >= 1.0 && pb->theta <= 2.0))) return -1;
    if (pb->nx < 2) return -1;
    if (pb->dim == 1 && pb->ny != 1) return -1;
    if (pb->dim == 2 && pb->ny < 2) return -1;
    if (!(pb->lx > 0) || (pb->dim == 2 && !(pb->ly > 0))) return -1;
    if ((double)pb->nx * (double)pb->ny > 2.0e9) return -1;
    return 0;
}

/* base_shll.c:200-218: float time = 0; while (time < TOTAL_TIME) { ...; time += DT; NO_STEPS++; } */
int64_t orc_steps_to_time(const orc_problem *pb, float t_end)
{
    cst_r32 k;
    float time = 0.0f;
    int64_t n = 0;
    make_cst_r32(&k, pb->nx, pb->ny, pb->lx, pb->dim == 2 ? pb->ly : 1.0);
    while (time < t_end) {
        time += k.dt;
        n++;
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* FP32-compensated mode.                                                    */
/* Constants are the FP64 oracle's, rounded once to float.  Split fluxes are  */
/* float functions of the hi word (fused multiply-adds where a product feeds  */
/* a sum).  The update is regrouped into face terms t = lam*G (one float per  */
/* face, shared by both cells) that are added to the (hi, lo) pair with       */
/* TwoSum (Knuth, TAOCP vol. 2, 4.2.2) and renormalised with Fast2Sum, so the */
/* state keeps ~48 significant bits and sum(hi+lo) telescopes: mass and       */
/* energy are conserved to ~1e-14 relative with any wall treatment that       */
/* conserves them.                                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, n, nv, dim, order, bc, kahan, lim;
    float R, gR, rcv, lamx, lamy, hlamx, hlamy, theta;
    double cv64, gR64;
    float *hi[4], *lo[4], *fp[4], *fm[4], *hp[4], *hm[4];
    float *a;  /* sound speed of the hi state (stage output) */
    void *block;
} c32grid;

static void two_sum(float a, float b, float *s, float *e)
{
    float ss = a + b;
    float bb = ss - a;
    *e = (a - (ss - bb)) + (b - bb);
    *s = ss;
}

/* C32 minmod: the median of (l, r, 0), i.e. the smaller magnitude when l and r
 * share a sign and 0 otherwise (the GPU evaluates it as one median instruction).
 * It differs from the reference's product test only where l*r underflows and in
 * the sign of a zero result, which never changes a face flux sum. */
static float mm_c32(float l, float r)
{
    if (l > 0.0f && r > 0.0f) return l < r ? l : r;
    if (l < 0.0f && r < 0.0f) return l > r ? l : r;
    return 0.0f;
}

static int c32_alloc(c32grid *g, const orc_problem *pb)
{
    cst_f64 k;
    int v;
    size_t n = (size_t)pb->nx * (size_t)pb->ny;
    float *base;
    memset(g, 0, sizeof(*g));
    g->nx = pb->nx; g->ny = pb->ny; g->n = (int)n; g->nv = orc_nvar(pb);
    g->dim = pb->dim; g->order = pb->order; g->bc = pb->bc;
    g->lim = pb->limiter;
    g->theta = (float)pb->theta;
    make_cst_f64(&k, pb->nx, pb->ny, pb->lx, pb->dim == 2 ? pb->ly : 1.0);
    g->R = (float)k.R;
    g->gR = (float)k.gR;
    g->rcv = (float)(1.0 / k.cv);
    g->lamx = (float)k.dtdx;
    g->lamy = (float)k.dtdy;
    g->hlamx = 0.5f * g->lamx;
    g->hlamy = 0.5f * g->lamy;
    g->cv64 = k.cv;
    g->gR64 = k.gR;
    g->block = malloc(n * sizeof(float) * (size_t)(g->nv * (pb->dim == 1 ? 4 : 6) + 1) + 64);
    if (!g->block) return -1;
    base = (float *)g->block;
    for (v = 0; v < g->nv; v++) {
        g->hi[v] = base; base += n;
        g->lo[v] = base; base += n;
        g->fp[v] = base; base += n;
        g->fm[v] = base; base += n;
        if (pb->dim == 2) { g->hp[v] = base; base += n; g->hm[v] = base; base += n; }
    }
    g->a = base;
    return 0;
}

/* FP64 primitive -> conserved with the scheme's kinetic-energy association. */
static void p2u_f64(int dim, int ke, double cv, const double *p, double *u)
{
    if (dim == 1) {
        u[0] = p[0];
        u[1] = p[0] * p[1];
        u[2] = ke == 0 ? p[0] * (p[2] * cv + (0.5 * p[1]) * p[1])
                       : p[0] * (p[2] * cv + 0.5 * (p[1] * p[1]));
    } else {
        u[0] = p[0];
        u[1] = p[0] * p[1];
        u[2] = p[0] * p[2];
        u[3] = p[0] * (p[3] * cv + 0.5 * (p[1] * p[1] + p[2] * p[2]));
    }
}

static void u2p_f64(int dim, int ke, double cv, const double *u, double *p)
{
    double r = u[0];
    double vx = u[1] / r;
    if (dim == 1) {
        p[0] = r; p[1] = vx;
        p[2] = ke == 0 ? ((u[2] / r) - (0.5 * vx) * vx) / cv : ((u[2] / r) - 0.5 * (vx * vx)) / cv;
    } else {
        double vy = u[2] / r;
        p[0] = r; p[1] = vx; p[2] = vy;
        p[3] = ((u[3] / r) - 0.5 * (vx * vx + vy * vy)) / cv;
    }
}

static void c32_init(c32grid *g, const double *prim)
{
    int c, v, ke = (g->dim == 1 && g->order == 1) ? 0 : 1;
    for (c = 0; c < g->n; c++) {
        double p[4], u[4];
        for (v = 0; v < g->nv; v++) p[v] = prim[(size_t)v * g->n + c];
        p2u_f64(g->dim, ke, g->cv64, p, u);
        for (v = 0; v < g->nv; v++) {
            float hi = (float)u[v];
            g->hi[v][c] = hi;
            g->lo[v][c] = (float)(u[v] - (double)hi);
        }
    }
}

/* C32 split of one flux component: F+ = F*Z1 + U*Z2, F- = -F*Z3 - U*Z2 with U*Z2
 * rounded once and each sum fused (fmaf is correctly rounded in C99/glibc). */
static void c32_split(float f, float U, float z1, float z2, float z3, float *fp, float *fm)
{
    float uz = U * z2;
    *fp = fmaf(f, z1, uz);
    *fm = fmaf(-f, z3, -uz);
}

/* z1 = (M+1)/2, z2 = (a/2)(1-M^2), z3 = (M-1)/2 (M/2 is exact, so z1 and z3 round
 * exactly like 0.5*(M+-1)) */
static void c32_z(float M, float a, float *z1, float *z2, float *z3)
{
    float h = 0.5f * M;
    *z1 = h + 0.5f;
    *z2 = (0.5f * a) * fmaf(-M, M, 1.0f);
    *z3 = h - 0.5f;
}

/* Per-cell float flux of the compensated mode: the reference's formulas
 * (base_shll.c:135-157, base_shll_2d.c:240-298) in float, with fused multiply-adds
 * where a product feeds a sum. */
static void c32_cell_flux(c32grid *g, int c)
{
    int v;
    if (g->dim == 1) {
        float r = g->hi[0][c], m = g->hi[1][c], E = g->hi[2][c];
        float inv = 1.0f / r;
        float vx = m * inv, e = E * inv;
        float T = fmaf(-0.5f * vx, vx, e) * g->rcv;
        float a = sqrtf(g->gR * T);
        float z1, z2, z3;
        float P = (r * g->R) * T;
        g->a[c] = a;
        float f[3], U[3];
        c32_z(vx / a, a, &z1, &z2, &z3);
        f[0] = m; f[1] = fmaf(m, vx, P); f[2] = vx * (E + P);
        U[0] = r; U[1] = m; U[2] = E;
        for (v = 0; v < 3; v++) c32_split(f[v], U[v], z1, z2, z3, &g->fp[v][c], &g->fm[v][c]);
    } else {
        float r = g->hi[0][c], mx = g->hi[1][c], my = g->hi[2][c], E = g->hi[3][c];
        float inv = 1.0f / r;
        float vx = mx * inv, vy = my * inv, e = E * inv;
        float T = fmaf(-0.5f, fmaf(vx, vx, vy * vy), e) * g->rcv;
        float a = sqrtf(g->gR * T);
        float P = (r * g->R) * T;
        float U[4], f[4], z1, z2, z3;
        g->a[c] = a;
        U[0] = r; U[1] = mx; U[2] = my; U[3] = E;
        c32_z(vx / a, a, &z1, &z2, &z3);
        f[0] = mx; f[1] = fmaf(mx, vx, P); f[2] = mx * vy; f[3] = vx * (E + P);
        for (v = 0; v < 4; v++) c32_split(f[v], U[v], z1, z2, z3, &g->fp[v][c], &g->fm[v][c]);
        c32_z(vy / a, a, &z1, &z2, &z3);
        f[0] = my; f[1] = my * vx; f[2] = fmaf(my, vy, P); f[3] = vy * (E + P);
        for (v = 0; v < 4; v++) c32_split(f[v], U[v], z1, z2, z3, &g->hp[v][c], &g->hm[v][c]);
    }
}

static const float sx1_c32[3] = {-1.0f, 1.0f, -1.0f};
static const float sx2_c32[4] = {-1.0f, 1.0f, -1.0f, -1.0f};
static const float sy2_c32[4] = {-1.0f, -1.0f, 1.0f, -1.0f};

/* slope of the split flux q at line position s (0 on the two boundary cells):
 * minmod of the one-sided differences, or with the MC limiter
 * mm(0.5*(q_r - q_l), theta*mm(q_c - q_l, q_r - q_c)), all in float */
static float c32_slope(const c32grid *g, const float *q, int c, int s, int stride, int len)
{
    float m;
    if (s <= 0 || s >= len - 1) return 0.0f;
    m = mm_c32(q[c] - q[c - stride], q[c + stride] - q[c]);
    if (g->lim != ORC_LIM_MC) return m;
    return mm_c32(0.5f * (q[c + stride] - q[c - stride]), g->theta * m);
}

/*
 * Face term t_k = lam * G_k of face k on a line of len cells (face k lies between
 * cells k-1 and k; c0 is the cell of line position 0).  G is the reference's update
 * regrouped by faces, which telescopes:
 *   interior      G_k = (fp_{k-1} + fm_k) [+ 0.5*(dfp_{k-1} - dfm_k)  second order]
 *   wall, k = 0   G_0 = fm_0 + S*fm_0          (reflective: L = S*fm_0)
 *                 G_0 = fp_0 + fm_0            (outflow:    L = fp_0)
 *   wall, k = len G_len = fp_{len-1} + S*fp_{len-1}  /  fp_{len-1} + fm_{len-1}
 * so that u_s -= lam*(((fp-fm)+R)-L) [- 0.5*lam*(((dfp+dfm)-Rdf)-Ldf)] equals
 * u_s += t_s - t_{s+1} in exact arithmetic.  Second order evaluates t_k as
 * fmaf(lam/2, d_k, lam*G_k) with d_k = dfp_{k-1} - dfm_k (0 at the walls).
 */
static float c32_face(const c32grid *g, float **qp, float **qm, const float *sgn, int v, int c0, int k,
                      int stride, int len, float lam)
{
    float G, d = 0.0f;
    const float *p = qp[v], *m = qm[v];
    if (k == 0) {
        G = g->bc == ORC_BC_REFLECT ? m[c0] + sgn[v] * m[c0] : p[c0] + m[c0];
    } else if (k == len) {
        int c = c0 + (len - 1) * stride;
        G = g->bc == ORC_BC_REFLECT ? p[c] + sgn[v] * p[c] : p[c] + m[c];
    } else {
        int cl = c0 + (k - 1) * stride, cr = cl + stride;
        G = p[cl] + m[cr];
        if (g->order == 2) d = c32_slope(g, p, cl, k - 1, stride, len) - c32_slope(g, m, cr, k, stride, len);
    }
    /* second order: lam*(G + d/2) as one fused multiply-add (d = 0 at the walls) */
    return g->order == 2 ? fmaf(0.5f * lam, d, lam * G) : lam * G;
}

/* Fast2Sum (Dekker): renormalise (a, b) with |b| below an ulp of a */
static void fast_two_sum(float a, float b, float *s, float *e)
{
    float ss = a + b;
    *e = b - (ss - a);
    *s = ss;
}

static void c32_step(c32grid *g)
{
    int c, v;
    for (c = 0; c < g->n; c++) c32_cell_flux(g, c);
    for (c = 0; c < g->n; c++) {
        int i = g->dim == 1 ? c : c / g->ny;
        int j = g->dim == 1 ? 0 : c % g->ny;
        int sx = g->dim == 1 ? 1 : g->ny;
        int cx0 = c - i * sx;      /* line position 0 of this cell's x line */
        int cy0 = c - j;           /* line position 0 of this cell's y line */
        for (v = 0; v < g->nv; v++) {
            /* the fluxes of this step all come from the old hi, so (hi, lo) of this
             * cell can be overwritten in place */
            float a = g->hi[v][c], b = g->lo[v][c], e;
            float tl = c32_face(g, g->fp, g->fm, g->dim == 1 ? sx1_c32 : sx2_c32, v, cx0, i, sx, g->nx, g->lamx);
            float tr = c32_face(g, g->fp, g->fm, g->dim == 1 ? sx1_c32 : sx2_c32, v, cx0, i + 1, sx, g->nx, g->lamx);
            float tb = 0.0f, tt = 0.0f;
            if (g->dim == 2) {
                tb = c32_face(g, g->hp, g->hm, sy2_c32, v, cy0, j, 1, g->ny, g->lamy);
                tt = c32_face(g, g->hp, g->hm, sy2_c32, v, cy0, j + 1, 1, g->ny, g->lamy);
            }
            if (g->kahan) {
                /* Kahan (compensated summation): lo is the part of the value not yet in
                 * hi; the cell's net face term y joins it and hi takes what fits */
                float y = g->dim == 1 ? (tl - tr) + b : ((tl - tr) + (tb - tt)) + b;
                float s = a + y;
                g->lo[v][c] = y - (s - a);
                g->hi[v][c] = s;
                continue;
            }
            two_sum(a, tl, &a, &e); b = b + e;
            two_sum(a, -tr, &a, &e); b = b + e;
            if (g->dim == 2) {
                two_sum(a, tb, &a, &e); b = b + e;
                two_sum(a, -tt, &a, &e); b = b + e;
            }
            fast_two_sum(a, b, &g->hi[v][c], &g->lo[v][c]);
        }
    }
}

static void c32_output(const c32grid *g, double *prim_out, double *cons_out)
{
    int c, v, ke = (g->dim == 1 && g->order == 1) ? 0 : 1;
    for (c = 0; c < g->n; c++) {
        double u[4], p[4];
        for (v = 0; v < g->nv; v++) u[v] = (double)g->hi[v][c] + (double)g->lo[v][c];
        u2p_f64(g->dim, ke, g->cv64, u, p);
        for (v = 0; v < g->nv; v++) {
            if (cons_out) cons_out[(size_t)v * g->n + c] = u[v];
            if (prim_out) prim_out[(size_t)v * g->n + c] = p[v];
        }
    }
}

static int c32_run(const orc_problem *pb, int kahan, const double *prim_in, int64_t nsteps, double *prim_out,
                   double *cons_out, float *hi_out, float *lo_out, double *elapsed)
{
    c32grid g;
    int64_t s;
    double t0;
    if (c32_alloc(&g, pb)) return -2;
    g.kahan = kahan;
    c32_init(&g, prim_in);
    t0 = orc_now();
    for (s = 0; s < nsteps; s++) c32_step(&g);
    if (elapsed) *elapsed = orc_now() - t0;
    c32_output(&g, prim_out, cons_out);
    if (hi_out || lo_out) {
        int v;
        for (v = 0; v < g.nv; v++) {
            if (hi_out) memcpy(hi_out + (size_t)v * g.n, g.hi[v], (size_t)g.n * sizeof(float));
            if (lo_out) memcpy(lo_out + (size_t)v * g.n, g.lo[v], (size_t)g.n * sizeof(float));
        }
    }
    free(g.block);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* FP64 face form (F64F, 1D): the compensated modes' face-term arrangement in   */
/* double.  Split fluxes scaled by DT/DX (so a face term is G + d/2 and the      */
/* update u += t_left - t_right); p = (E - (rho u) u/2) R/cv with no E/rho;      */
/* 1/a = 1/sqrt(gamma R p/rho), M = u/a and a/2 from it (no square root and no   */
/* division besides 1/rho and 1/sqrt); fused multiply-adds where a product       */
/* feeds a sum (fma is correctly rounded in C99/glibc); minmod as the median of  */
/* (l, r, 0).  Here 1/rho and 1/sqrt are 1.0/r and 1.0/sqrt(x); the kernels use  */
/* a hardware seed and two Newton steps for each (the same 1/rho, and 1/sqrt     */
/* within 2 ulp, measured), so kernel and oracle agree to round-off, not bit for */
/* bit.  The initial state is the F64 mode's.                                    */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, n, order, bc, lim, ke;
    double gR, rcv, cv, lamx, theta;
    double *u[3], *fp[3], *fm[3], *a;
    void *block;
} f64fgrid;

static double mm_f64f(double l, double r)
{
    if (l > 0.0 && r > 0.0) return l < r ? l : r;
    if (l < 0.0 && r < 0.0) return l > r ? l : r;
    return 0.0;
}

static int f64f_alloc(f64fgrid *g, const orc_problem *pb, const double *prim)
{
    cst_f64 k;
    int v, c;
    size_t n = (size_t)pb->nx;
    double *base;
    memset(g, 0, sizeof(*g));
    g->nx = pb->nx; g->n = (int)n;
    g->order = pb->order; g->bc = pb->bc; g->lim = pb->limiter;
    g->ke = pb->order == 1 ? 0 : 1;
    g->theta = pb->theta;
    make_cst_f64(&k, pb->nx, 1, pb->lx, 1.0);
    g->gR = k.gR;
    g->cv = k.cv;
    g->rcv = 1.0 / k.cv;
    g->lamx = k.dtdx;
    g->block = malloc(n * sizeof(double) * 10 + 64);
    if (!g->block) return -1;
    base = (double *)g->block;
    for (v = 0; v < 3; v++) {
        g->u[v] = base; base += n;
        g->fp[v] = base; base += n;
        g->fm[v] = base; base += n;
    }
    g->a = base;
    for (c = 0; c < g->n; c++) {
        double p[3], u[3];
        for (v = 0; v < 3; v++) p[v] = prim[(size_t)v * g->n + c];
        p2u_f64(1, g->ke, g->cv, p, u);
        for (v = 0; v < 3; v++) g->u[v][c] = u[v];
    }
    return 0;
}

static int f64f_rsq_ulps = 0; /* orc_f64f_rsq_nudge */

void orc_f64f_rsq_nudge(int ulps) { f64f_rsq_ulps = ulps < 0 ? -ulps : ulps; }

/* 1.0/sqrt(x), moved by the test switch's pseudo-random offset when it is set */
static double f64f_rsq(double x)
{
    double y = 1.0 / sqrt(x);
    uint64_t b;
    int d, i;
    if (!f64f_rsq_ulps) return y;
    memcpy(&b, &x, sizeof(b));
    b ^= b >> 29;
    b *= 0xbf58476d1ce4e5b9ULL;
    b ^= b >> 32;
    d = (int)(b % (uint64_t)(2 * f64f_rsq_ulps + 1)) - f64f_rsq_ulps;
    for (i = 0; i < (d < 0 ? -d : d); i++) y = nextafter(y, d > 0 ? INFINITY : 0.0);
    return y;
}

/* split fluxes of one cell scaled by sc (base_shll.c:135-157 in the F64F
 * arrangement; the step passes DT/DX, the stage entry 1) and the sound speed */
static void f64f_cell_flux(f64fgrid *g, int c, double sc)
{
    int v;
    double r = g->u[0][c], m = g->u[1][c], E = g->u[2][c];
    double inv = 1.0 / r;
    double vx = m * inv;
    double P = fma(-0.5 * m, vx, E) * g->rcv; /* rho R T, R = 1 */
    double x = g->gR * (P * inv);            /* a^2 */
    double y = f64f_rsq(x);                  /* 1/a */
    double M = vx * y, hs = 0.5 * sc;
    double z1 = fma(hs, M, hs), z3 = fma(hs, M, -hs), z2 = ((hs * x) * y) * fma(-M, M, 1.0);
    double f[3], U[3];
    g->a[c] = x * y;
    f[0] = m; f[1] = fma(m, vx, P); f[2] = vx * (E + P);
    U[0] = r; U[1] = m; U[2] = E;
    for (v = 0; v < 3; v++) {
        double uz = U[v] * z2;
        g->fp[v][c] = fma(f[v], z1, uz);
        g->fm[v][c] = fma(-f[v], z3, -uz);
    }
}

static double f64f_slope(const f64fgrid *g, const double *q, int s)
{
    double m;
    if (s <= 0 || s >= g->nx - 1) return 0.0;
    m = mm_f64f(q[s] - q[s - 1], q[s + 1] - q[s]);
    if (g->lim != ORC_LIM_MC) return m;
    return mm_f64f(0.5 * (q[s + 1] - q[s - 1]), g->theta * m);
}

/* face term of face k (between cells k-1 and k) from the scaled split fluxes */
static double f64f_face(const f64fgrid *g, int v, int k)
{
    static const double sgn[3] = {-1.0, 1.0, -1.0};
    const double *p = g->fp[v], *m = g->fm[v];
    double G, d = 0.0;
    int len = g->nx;
    if (k == 0) {
        G = g->bc == ORC_BC_REFLECT ? m[0] + sgn[v] * m[0] : p[0] + m[0];
    } else if (k == len) {
        G = g->bc == ORC_BC_REFLECT ? p[len - 1] + sgn[v] * p[len - 1] : p[len - 1] + m[len - 1];
    } else {
        G = p[k - 1] + m[k];
        if (g->order == 2) d = f64f_slope(g, p, k - 1) - f64f_slope(g, m, k);
    }
    return g->order == 2 ? fma(0.5, d, G) : G;
}

static void f64f_step(f64fgrid *g)
{
    int c, v;
    for (c = 0; c < g->n; c++) f64f_cell_flux(g, c, g->lamx);
    for (v = 0; v < 3; v++) {
        double tl = f64f_face(g, v, 0);
        for (c = 0; c < g->n; c++) {
            double tr = f64f_face(g, v, c + 1);
            g->u[v][c] = g->u[v][c] + (tl - tr);
            tl = tr;
        }
    }
}

static void f64f_output(const f64fgrid *g, double *prim_out, double *cons_out)
{
    int c, v;
    for (c = 0; c < g->n; c++) {
        double u[3], p[3];
        for (v = 0; v < 3; v++) u[v] = g->u[v][c];
        u2p_f64(1, g->ke, g->cv, u, p);
        for (v = 0; v < 3; v++) {
            if (cons_out) cons_out[(size_t)v * g->n + c] = u[v];
            if (prim_out) prim_out[(size_t)v * g->n + c] = p[v];
        }
    }
}

static int f64f_run(const orc_problem *pb, const double *prim_in, int64_t nsteps, double *prim_out,
                    double *cons_out, double *elapsed)
{
    f64fgrid g;
    int64_t s;
    double t0;
    if (pb->dim != 1) return -1;
    if (f64f_alloc(&g, pb, prim_in)) return -2;
    t0 = orc_now();
    for (s = 0; s < nsteps; s++) f64f_step(&g);
    if (elapsed) *elapsed = orc_now() - t0;
    f64f_output(&g, prim_out, cons_out);
    free(g.block);
    return 0;
}

/* ------------------------------------------------------------------------- */
int orc_run(const orc_problem *pb, int mode, const double *prim_in, int64_t nsteps,
            double *prim_out, double *cons_out, int nthreads, double *elapsed_s)
{
    if (orc_check(pb) || !prim_in || nsteps < 0) return -1;
    switch (mode) {
    case ORC_REF32: return run_r32(pb, prim_in, nsteps, prim_out, cons_out, nthreads, elapsed_s);
    case ORC_F64: return run_f64(pb, prim_in, nsteps, prim_out, cons_out, nthreads, elapsed_s);
    case ORC_C32: return c32_run(pb, 0, prim_in, nsteps, prim_out, cons_out, NULL, NULL, elapsed_s);
    case ORC_C32K: return c32_run(pb, 1, prim_in, nsteps, prim_out, cons_out, NULL, NULL, elapsed_s);
    case ORC_F64F: return f64f_run(pb, prim_in, nsteps, prim_out, cons_out, elapsed_s);
    default: return -1;
    }
}

int orc_run_c32_raw(const orc_problem *pb, int mode, const double *prim_in, int64_t nsteps,
                    float *hi_out, float *lo_out)
{
    if (orc_check(pb) || !prim_in || nsteps < 0 || (mode != ORC_C32 && mode != ORC_C32K)) return -1;
    return c32_run(pb, mode == ORC_C32K, prim_in, nsteps, NULL, NULL, hi_out, lo_out, NULL);
}

void orc_constants(const orc_problem *pb, int mode, double out[6])
{
    if (mode == ORC_REF32) {
        cst_r32 k;
        make_cst_r32(&k, pb->nx, pb->ny, pb->lx, pb->dim == 2 ? pb->ly : 1.0);
        out[0] = k.gamma; out[1] = k.R; out[2] = k.cv; out[3] = k.dt; out[4] = k.dtdx; out[5] = k.dtdy;
    } else {
        cst_f64 k;
        make_cst_f64(&k, pb->nx, pb->ny, pb->lx, pb->dim == 2 ? pb->ly : 1.0);
        out[0] = k.gamma; out[1] = k.R; out[2] = k.cv; out[3] = k.dt; out[4] = k.dtdx; out[5] = k.dtdy;
    }
}

/* Stage evaluation on whole arrays (Compute_U_from_P, Compute_P_from_U, Compute_F_from_P once). */
#define STAGE_BODY(SUF)                                                                        \
    do {                                                                                       \
        grid##SUF g;                                                                           \
        int v, c;                                                                              \
        if (grid_alloc##SUF(&g, pb)) return -2;                                                \
        for (v = 0; v < g.nv; v++)                                                             \
            for (c = 0; c < g.n; c++) g.p[v][c] = prim_in[(size_t)v * g.n + c];                \
        u_from_p##SUF(&g);                                                                     \
        p_from_u##SUF(&g);                                                                     \
        flux##SUF(&g);                                                                         \
        for (v = 0; v < g.nv; v++)                                                             \
            for (c = 0; c < g.n; c++) {                                                        \
                size_t o = (size_t)v * g.n + c;                                                \
                if (cons) cons[o] = g.u[v][c];                                                 \
                if (prim) prim[o] = g.p[v][c];                                                 \
                if (fp) fp[o] = g.fp[v][c];                                                    \
                if (fm) fm[o] = g.fm[v][c];                                                    \
                if (hp && g.dim == 2) hp[o] = g.hp[v][c];                                      \
                if (hm && g.dim == 2) hm[o] = g.hm[v][c];                                      \
            }                                                                                  \
        if (a)                                                                                 \
            for (c = 0; c < g.n; c++) a[c] = g.a[c];                                           \
        free(g.block);                                                                         \
    } while (0)

int orc_stage(const orc_problem *pb, int mode, const double *prim_in, double *cons, double *prim,
              double *a, double *fp, double *fm, double *hp, double *hm)
{
    if (orc_check(pb) || !prim_in) return -1;
    if (mode == ORC_REF32) {
        STAGE_BODY(_r32);
    } else if (mode == ORC_F64) {
        STAGE_BODY(_f64);
    } else if (mode == ORC_C32 || mode == ORC_C32K) {
        c32grid g;
        int v, c;
        if (c32_alloc(&g, pb)) return -2;
        c32_init(&g, prim_in);
        for (c = 0; c < g.n; c++) c32_cell_flux(&g, c);
        for (v = 0; v < g.nv; v++)
            for (c = 0; c < g.n; c++) {
                size_t o = (size_t)v * g.n + c;
                if (cons) cons[o] = (double)g.hi[v][c] + (double)g.lo[v][c];
                if (fp) fp[o] = g.fp[v][c];
                if (fm) fm[o] = g.fm[v][c];
                if (hp && g.dim == 2) hp[o] = g.hp[v][c];
                if (hm && g.dim == 2) hm[o] = g.hm[v][c];
            }
        if (a)
            for (c = 0; c < g.n; c++) a[c] = g.a[c];
        if (prim) c32_output(&g, prim, NULL);
        free(g.block);
    } else if (mode == ORC_F64F && pb->dim == 1) {
        f64fgrid g;
        int v, c;
        if (f64f_alloc(&g, pb, prim_in)) return -2;
        for (c = 0; c < g.n; c++) f64f_cell_flux(&g, c, 1.0);
        for (v = 0; v < 3; v++)
            for (c = 0; c < g.n; c++) {
                size_t o = (size_t)v * g.n + c;
                if (cons) cons[o] = g.u[v][c];
                if (fp) fp[o] = g.fp[v][c];
                if (fm) fm[o] = g.fm[v][c];
            }
        if (a)
            for (c = 0; c < g.n; c++) a[c] = g.a[c];
        if (prim) f64f_output(&g, prim, NULL);
        free(g.block);
    } else {
        return -1;
    }
    return 0;
}
#undef STAGE_BODY

double orc_minmod(int mode, double l, double r)
{
    if (mode == ORC_REF32) return (double)mm_r32((float)l, (float)r);
    if (mode == ORC_C32 || mode == ORC_C32K) return (double)mm_c32((float)l, (float)r);
    if (mode == ORC_F64F) return mm_f64f(l, r);
    return mm_f64(l, r);
}

double orc_slope(int mode, int limiter, double theta, double ql, double qc, double qr)
{
    if (mode == ORC_REF32) {
        grid_r32 g;
        g.lim = limiter; g.theta = (float)theta;
        return (double)slope_r32(&g, (float)ql, (float)qc, (float)qr);
    }
    if (mode == ORC_C32 || mode == ORC_C32K) {
        c32grid g;
        float q[3] = {(float)ql, (float)qc, (float)qr};
        g.lim = limiter; g.theta = (float)theta;
        return (double)c32_slope(&g, q, 1, 1, 1, 3);
    }
    if (mode == ORC_F64F) {
        f64fgrid g;
        double q[3] = {ql, qc, qr};
        g.nx = 3; g.lim = limiter; g.theta = theta;
        return f64f_slope(&g, q, 1);
    }
    {
        grid_f64 g;
        g.lim = limiter; g.theta = theta;
        return slope_f64(&g, ql, qc, qr);
    }
}

/* ------------------------------------------------------------------------- */
/* Initial conditions.  Values pass through float exactly as the reference's  */
/* float arrays hold them.                                                    */
/* ------------------------------------------------------------------------- */

/* base_shll.c:54-60:  if (i < 0.5*N) left else right  (comparison in double) */
void orc_ic_tube(int n, double x0, double rhoL, double uL, double TL,
                 double rhoR, double uR, double TR, double *prim)
{
    int i;
    for (i = 0; i < n; i++) {
        int left = (double)i < x0 * (double)n;
        prim[i] = (double)(float)(left ? rhoL : rhoR);
        prim[n + i] = (double)(float)(left ? uL : uR);
        prim[2 * n + i] = (double)(float)(left ? TL : TR);
    }
}

/* base_shll_2d.c:93-103: rho = 1 where (i > 0.2*NX) && (i < 0.8*NX) && (j > 0.2*NY) && (j < 0.8*NY),
 * else 10; u = v = 0; T = 1. */
void orc_ic_implosion(int nx, int ny, double *prim)
{
    int i, j;
    size_t n = (size_t)nx * ny;
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            size_t c = (size_t)i * ny + j;
            int in = (i > 0.2 * nx) && (i < 0.8 * nx) && (j > 0.2 * ny) && (j < 0.8 * ny);
            prim[c] = in ? 1.0 : 10.0;
            prim[n + c] = 0.0;
            prim[2 * n + c] = 0.0;
            prim[3 * n + c] = 1.0;
        }
}

/* 2nd_order_base_shll.c:133-148 (four shocks).  T = p/(rho*R) is evaluated in double
 * from the float rho (the literals are double), then stored as float.  Cells on the
 * lines i == 0.75*NX or j == 0.75*NY fall through to the last branch. */
void orc_ic_four_shocks(int nx, int ny, double *prim)
{
    int i, j;
    size_t n = (size_t)nx * ny;
    const float R = 1.0f;
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            size_t c = (size_t)i * ny + j;
            float r, vx, vy, T;
            if ((i < 0.75 * nx) && (j < 0.75 * ny)) {
                r = 0.138f; vx = 1.206f; vy = 1.206f; T = (float)(0.029 / (double)(r * R));
            } else if ((i > 0.75 * nx) && (j < 0.75 * ny)) {
                r = 0.5323f; vx = 0.0f; vy = 1.206f; T = (float)(0.3 / (double)(r * R));
            } else if ((i < 0.75 * nx) && (j > 0.75 * ny)) {
                r = 0.5323f; vx = 1.206f; vy = 0.0f; T = (float)(0.3 / (double)(r * R));
            } else {
                r = 1.5f; vx = 0.0f; vy = 0.0f; T = (float)(1.5 / (double)(r * R));
            }
            prim[c] = r; prim[n + c] = vx; prim[2 * n + c] = vy; prim[3 * n + c] = T;
        }
}

/* base-omp/2nd_order_base_shll.c:149-167 ("Configuration 6", four contacts), with the
 * cell counter starting at 0 (that file's reset of it sits in a commented-out block).
 * Regions by the double comparisons of the reference; cells on the lines i == 0.5*NX or
 * j == 0.5*NY fall through to the last branch.  T = 1.0/(rho*R) in double from the
 * float product, stored as float. */
void orc_ic_four_contacts(int nx, int ny, double *prim)
{
    int i, j;
    size_t n = (size_t)nx * ny;
    const float R = 1.0f;
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            size_t c = (size_t)i * ny + j;
            float r, vx, vy;
            if ((i < 0.5 * nx) && (j < 0.5 * ny)) {
                r = 1.0f; vx = -0.75f; vy = 0.5f;
            } else if ((i > 0.5 * nx) && (j < 0.5 * ny)) {
                r = 3.0f; vx = -0.75f; vy = -0.5f;
            } else if ((i < 0.5 * nx) && (j > 0.5 * ny)) {
                r = 2.0f; vx = 0.75f; vy = 0.5f;
            } else {
                r = 1.0f; vx = 0.75f; vy = -0.5f;
            }
            prim[c] = r; prim[n + c] = vx; prim[2 * n + c] = vy;
            prim[3 * n + c] = (float)(1.0 / (double)(r * R));
        }
}

void orc_ic_circle(int nx, int ny, double lx, double ly, double cx, double cy, double r,
                   double rho_in, double T_in, double rho_out, double T_out, double *prim)
{
    int i, j;
    size_t n = (size_t)nx * ny;
    double dx = lx / nx, dy = ly / ny;
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            size_t c = (size_t)i * ny + j;
            double x = (i + 0.5) * dx - cx, y = (j + 0.5) * dy - cy;
            int in = x * x + y * y < r * r;
            prim[c] = (double)(float)(in ? rho_in : rho_out);
            prim[n + c] = 0.0;
            prim[2 * n + c] = 0.0;
            prim[3 * n + c] = (double)(float)(in ? T_in : T_out);
        }
}
