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
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

int main(int argc, char **argv)
{
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
