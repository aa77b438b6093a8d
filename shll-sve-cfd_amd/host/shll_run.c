/*
 * shll_run.c -- C host driver: the reference's main() (base_shll.c:198-225,
 * base_shll_2d.c:343-370, 2nd_order_base_shll.c:565-592, base-omp/2nd_order_base_shll.c
 * :615-658) over the GPU library.
 *
 *   Allocate_and_Init_Memory; Compute_U_from_P; Compute_P_from_U   -> shll_init
 *   while (time < TOTAL_TIME) { F; Update; P; time += DT; }         -> shll_run_to_time
 *   printf("Completed in %d steps\n", NO_STEPS)                    -> same line
 *   Save_Results                                                    -> shll_write_results
 *                                        (ten columns, --full)      -> shll_write_results_full
 *   Free_Memory                                                     -> shll_free
 *
 * The reference's compile-time constants become flags.  Defaults reproduce
 * base_shll.c (1D, N=256, t=0.2, Sod, reflective walls, reference float arithmetic).
 *
 *   shll_run [--dim 1|2] [--order 1|2] [--nx N] [--ny N] [--precision f64|ref32|c32|c32k|f64f]
 *            [--bc reflect|outflow] [--ic sod|implosion|four-shocks|four-contacts|circle]
 *            [--limiter minmod|mc] [--theta A] [--t-end T] [--steps K] [--out results.dat]
 *            [--full] [--sums] [--dump PATH]
 * --dump writes the end state in binary: the primitives (nvar*nx*ny doubles, SoA) followed by
 * the state in its storage format (shll_get_state_raw).
 * --limiter mc (theta 1.25) with --dim 2 --order 2 --nx 1024 is base-omp/2nd_order_base_shll.c:
 * its default initial data are the four contacts and its TOTAL_TIME 0.3 (:43).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shll.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: shll_run [--dim 1|2] [--order 1|2] [--nx N] [--ny N] [--precision f64|ref32|c32|c32k|f64f]\n"
            "                [--bc reflect|outflow] [--ic sod|implosion|four-shocks|four-contacts|circle]\n"
            "                [--limiter minmod|mc] [--theta A] [--t-end T] [--steps K] [--out PATH]\n"
            "                [--full] [--sums] [--dump PATH]\n");
}

int main(int argc, char **argv)
{
    shll_problem pb;
    const char *ic = NULL, *out = "results.dat", *dump = NULL;
    float t_end = -1.0f;
    long long steps = -1;
    int sums = 0, full = 0, i, order_set = 0, bc_set = 0, nv;
    double *prim;
    shll_ctx *ctx = NULL;
    shll_status st;
    int64_t done = 0, n;

    memset(&pb, 0, sizeof(pb));
    pb.dim = 1; pb.order = 1; pb.bc = SHLL_BC_REFLECT; pb.precision = SHLL_REF32;
    pb.nx = 256; pb.ny = 1; pb.lx = 1.0; pb.ly = 1.0;
    pb.limiter = SHLL_LIM_MINMOD; pb.theta = 1.25;
    for (i = 1; i < argc; i++) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--sums")) { sums = 1; continue; }
        if (!strcmp(a, "--full")) { full = 1; continue; }
        if (!v) { usage(); return 2; }
        if (!strcmp(a, "--dim")) pb.dim = atoi(v);
        else if (!strcmp(a, "--order")) { pb.order = atoi(v); order_set = 1; }
        else if (!strcmp(a, "--nx")) pb.nx = atoll(v);
        else if (!strcmp(a, "--ny")) pb.ny = atoll(v);
        else if (!strcmp(a, "--precision"))
            pb.precision = !strcmp(v, "f64")    ? SHLL_F64
                           : !strcmp(v, "c32")  ? SHLL_C32
                           : !strcmp(v, "c32k") ? SHLL_C32K
                           : !strcmp(v, "f64f") ? SHLL_F64F
                                                : SHLL_REF32;
        else if (!strcmp(a, "--bc")) { pb.bc = !strcmp(v, "outflow") ? SHLL_BC_OUTFLOW : SHLL_BC_REFLECT; bc_set = 1; }
        else if (!strcmp(a, "--ic")) ic = v;
        else if (!strcmp(a, "--limiter")) pb.limiter = !strcmp(v, "mc") ? SHLL_LIM_MC : SHLL_LIM_MINMOD;
        else if (!strcmp(a, "--theta")) pb.theta = atof(v);
        else if (!strcmp(a, "--t-end")) t_end = (float)atof(v);
        else if (!strcmp(a, "--steps")) steps = atoll(v);
        else if (!strcmp(a, "--out")) out = v;
        else if (!strcmp(a, "--dump")) dump = v;
        else { usage(); return 2; }
        i++;
    }
    (void)order_set;
    if (pb.dim == 2 && pb.ny == 1) pb.ny = pb.nx;
    if (!bc_set) pb.bc = pb.order == 1 ? SHLL_BC_REFLECT : SHLL_BC_OUTFLOW;
    if (!ic) ic = pb.dim == 1 ? "sod" : (pb.order == 1 ? "implosion" : (pb.limiter == SHLL_LIM_MC ? "four-contacts" : "four-shocks"));
    /* TOTAL_TIME of each reference file (base_shll.c:26, base_shll_2d.c:35, 2nd_order_base_shll.c:43,
     * base-omp/2nd_order_base_shll.c:43) */
    if (t_end < 0) t_end = pb.dim == 1 ? 0.2f : (pb.order == 1 ? 0.1f : (pb.limiter == SHLL_LIM_MC ? 0.3f : 0.8f));

    nv = pb.dim == 1 ? 3 : 4;
    prim = (double *)malloc(sizeof(double) * (size_t)nv * (size_t)(pb.nx * pb.ny));
    if (!prim) { fprintf(stderr, "out of memory\n"); return 1; }
    if (!strcmp(ic, "sod")) shll_ic_tube(pb.nx, 0.5, 10.0, 0.0, 1.0, 1.0, 0.0, 1.0, prim);
    else if (!strcmp(ic, "implosion")) shll_ic_implosion(pb.nx, pb.ny, prim);
    else if (!strcmp(ic, "four-shocks")) shll_ic_four_shocks(pb.nx, pb.ny, prim);
    else if (!strcmp(ic, "four-contacts")) shll_ic_four_contacts(pb.nx, pb.ny, prim);
    else if (!strcmp(ic, "circle")) shll_ic_circle(pb.nx, pb.ny, 1.0, 1.0, 0.5, 0.5, 0.25, 10.0, 1.0, 1.0, 1.0, prim);
    else { fprintf(stderr, "unknown --ic %s\n", ic); free(prim); return 2; }

    st = shll_init(&pb, prim, &ctx);
    free(prim);
    if (st != SHLL_OK) { fprintf(stderr, "shll_init: %s\n", shll_last_error()); return 1; }
    if (steps >= 0) {
        st = shll_step(ctx, steps);
        done = steps;
        if (st == SHLL_OK) st = shll_check_state(ctx, NULL);
    } else {
        st = shll_run_to_time(ctx, t_end, &done);
    }
    if (st != SHLL_OK) { fprintf(stderr, "step loop: %s\n", shll_last_error()); shll_free(ctx); return 1; }
    printf("Completed in %lld steps\n", (long long)done);
    if (sums) {
        double s[4];
        if (shll_conserved_sums(ctx, s) == SHLL_OK) {
            printf("sums:");
            for (n = 0; n < nv; n++) printf(" %.17g", s[n]);
            printf("\n");
        }
    }
    if (dump) {
        size_t ncell = (size_t)(pb.nx * pb.ny), raw = shll_state_bytes(ctx);
        double *p = (double *)malloc(sizeof(double) * (size_t)nv * ncell);
        void *r = malloc(raw);
        FILE *f = fopen(dump, "wb");
        st = (p && r && f) ? shll_get_primitives(ctx, p) : SHLL_ENOMEM;
        if (st == SHLL_OK) st = shll_get_state_raw(ctx, r, raw);
        if (st == SHLL_OK && (fwrite(p, sizeof(double), (size_t)nv * ncell, f) != (size_t)nv * ncell ||
                              fwrite(r, 1, raw, f) != raw))
            st = SHLL_EINVAL;
        if (f && fclose(f) != 0) st = SHLL_EINVAL;
        free(p);
        free(r);
        if (st != SHLL_OK) {
            fprintf(stderr, "dump to %s failed: %s\n", dump, shll_last_error());
            shll_free(ctx);
            return 1;
        }
    }
    if (out && strcmp(out, "-") != 0) {
        if ((full ? shll_write_results_full(ctx, out) : shll_write_results(ctx, out)) != SHLL_OK) {
            fprintf(stderr, "write: %s\n", shll_last_error());
            shll_free(ctx);
            return 1;
        }
    }
    shll_free(ctx);
    return 0;
}
