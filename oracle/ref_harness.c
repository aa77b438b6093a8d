/*
 * ref_harness.c -- drives a compiled reference solver to write golden vectors.
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference's own
 * object file (compiled from /root/reference/base-c/<file>.c with main renamed to
 * reference_main); the binaries land in oracle/_ref/ and are never committed.
 *
 * The reference keeps its state in file-scope float arrays with external linkage
 * (base_shll.c:9-15, base_shll_2d.c:9-20, 2nd_order_base_shll.c:9-28, and the same in
 * base-omp/2nd_order_base_shll.c:9-29) and exposes
 * its stages as external void functions, so this file can
 *   - run the reference's own main() and dump the arrays when Free_Memory() makes
 *     its first free() call (link with -Wl,--wrap=free), or
 *   - repeat main()'s sequence itself (Allocate_and_Init_Memory, optional
 *     replacement of the primitive initial data, Compute_U_from_P,
 *     Compute_P_from_U, then {Compute_F_from_P; Update_U_from_F;
 *     Compute_P_from_U} per step) and dump after chosen step counts.
 *
 * Dump format (little endian): int32 magic 0x4C4C4853 ("SHLL"), int32 nvar,
 * int32 ncells, int32 steps, then float p[nvar][ncells], u[nvar][ncells], a[ncells].
 *
 * Usage:
 *   <bin> main  <out.bin>
 *   <bin> steps <out_prefix> <ic.bin|-> <k1,k2,...>   -> <out_prefix>_<k>.bin
 *   ic.bin: float prim[nvar][ncells] replacing the reference initial data.
 *   <bin> time  <nsteps> <nthreads>
 *       CPU baseline: the reference's own initial data and stage sequence for nsteps
 *       steps, timed with CLOCK_MONOTONIC around the step loop only (set-up, initial
 *       U <- P and the output are outside).  In the OpenMP build (REF_TIMED_OMP, the
 *       base-omp file compiled with OpenMP on) the loop runs as base-omp's main() runs
 *       it (base-omp/2nd_order_base_shll.c:620-653): every thread of one parallel
 *       region calls each stage, whose orphaned `omp for` loops share the cells, with a
 *       barrier per step.  Other builds are serial (nthreads must be 1).
 *       Prints: cells <N> steps <nsteps> threads <nthreads> seconds <t> checksum <sum of rho>
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#if defined(REF_O1_1D)
#define NVAR 3
extern float *p0, *p1, *p2, *u0, *u1, *u2, *a;
#else
#define NVAR 4
extern float *p0, *p1, *p2, *p3, *u0, *u1, *u2, *u3, *a;
#endif
extern const int N;
#if defined(REF_OMP_O2_2D)
/* base-omp/2nd_order_base_shll.c counts steps in a local of main (:626); the "main"
 * mode's dump then reports 0 steps */
int NO_STEPS = 0;
#else
extern int NO_STEPS;
#endif

void Allocate_and_Init_Memory(void);
void Compute_U_from_P(void);
void Compute_P_from_U(void);
void Compute_F_from_P(void);
void Update_U_from_F(void);
void Free_Memory(void);
int reference_main(void);

void __real_free(void *ptr);

static const char *g_dump_on_free = NULL;

static void arrays(float **P, float **U)
{
#if defined(REF_O1_1D)
    P[0] = p0; P[1] = p1; P[2] = p2;
    U[0] = u0; U[1] = u1; U[2] = u2;
#else
    P[0] = p0; P[1] = p1; P[2] = p2; P[3] = p3;
    U[0] = u0; U[1] = u1; U[2] = u2; U[3] = u3;
#endif
}

static void dump(const char *path, int steps)
{
    float *P[4], *U[4];
    int32_t hdr[4];
    int v;
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    arrays(P, U);
    hdr[0] = 0x4C4C4853; hdr[1] = NVAR; hdr[2] = N; hdr[3] = steps;
    fwrite(hdr, sizeof(int32_t), 4, f);
    for (v = 0; v < NVAR; v++) fwrite(P[v], sizeof(float), (size_t)N, f);
    for (v = 0; v < NVAR; v++) fwrite(U[v], sizeof(float), (size_t)N, f);
    fwrite(a, sizeof(float), (size_t)N, f);
    fclose(f);
}

void __wrap_free(void *ptr)
{
    if (g_dump_on_free) {
        const char *path = g_dump_on_free;
        g_dump_on_free = NULL;
        dump(path, NO_STEPS);
    }
    __real_free(ptr);
}

static void load_ic(const char *path)
{
    float *P[4], *U[4];
    int v;
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); exit(2); }
    arrays(P, U);
    for (v = 0; v < NVAR; v++)
        if (fread(P[v], sizeof(float), (size_t)N, f) != (size_t)N) {
            fprintf(stderr, "short initial-data file %s\n", path);
            exit(2);
        }
    fclose(f);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* the timed step loop of the "time" mode */
static int time_steps(long nsteps, int nthreads)
{
    double t0 = 0.0, t1 = 0.0, sum = 0.0;
    long i;
    float *P[4], *U[4];
    Allocate_and_Init_Memory();
#if defined(REF_TIMED_OMP)
#pragma omp parallel num_threads(nthreads)
    {
        long s;
        Compute_U_from_P();
        Compute_P_from_U();
#pragma omp barrier
#pragma omp master
        t0 = now_s();
        for (s = 0; s < nsteps; s++) {
            Compute_F_from_P();
            Update_U_from_F();
            Compute_P_from_U();
#pragma omp barrier
        }
#pragma omp master
        t1 = now_s();
    }
#else
    long s;
    if (nthreads != 1) {
        fprintf(stderr, "this reference build is serial: use 1 thread\n");
        return 1;
    }
    Compute_U_from_P();
    Compute_P_from_U();
    t0 = now_s();
    for (s = 0; s < nsteps; s++) {
        Compute_F_from_P();
        Update_U_from_F();
        Compute_P_from_U();
    }
    t1 = now_s();
#endif
    arrays(P, U);
    for (i = 0; i < N; i++) sum += P[0][i];
    printf("cells %d steps %ld threads %d seconds %.6f checksum %.9e\n", N, nsteps, nthreads, t1 - t0, sum);
    Free_Memory();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 4 && strcmp(argv[1], "time") == 0) return time_steps(atol(argv[2]), atoi(argv[3]));
    if (argc == 3 && strcmp(argv[1], "main") == 0) {
        g_dump_on_free = argv[2];
        reference_main();
        return 0;
    }
    if (argc == 5 && strcmp(argv[1], "steps") == 0) {
        int ks[64], nk = 0, kmax = 0, s, i;
        char buf[4096], *tok;
        strncpy(buf, argv[4], sizeof(buf) - 1);
        buf[sizeof(buf) - 1] = 0;
        for (tok = strtok(buf, ","); tok && nk < 64; tok = strtok(NULL, ",")) {
            ks[nk] = atoi(tok);
            if (ks[nk] > kmax) kmax = ks[nk];
            nk++;
        }
        Allocate_and_Init_Memory();
        if (strcmp(argv[3], "-") != 0) load_ic(argv[3]);
        Compute_U_from_P();
        Compute_P_from_U();
        for (s = 0;; s++) {
            for (i = 0; i < nk; i++)
                if (ks[i] == s) {
                    char path[4096];
                    snprintf(path, sizeof(path), "%s_%d.bin", argv[2], s);
                    dump(path, s);
                }
            if (s == kmax) break;
            Compute_F_from_P();
            Update_U_from_F();
            Compute_P_from_U();
        }
        Free_Memory();
        return 0;
    }
    fprintf(stderr, "usage: %s main <out.bin> | steps <prefix> <ic.bin|-> <k1,k2,...>\n", argv[0]);
    return 1;
}
