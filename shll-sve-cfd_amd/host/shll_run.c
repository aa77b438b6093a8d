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
 * The reference's compile-time constants (base_shll.c:16-26 and the like) become options,
 * given as flags or in a JSON file.  Defaults reproduce base_shll.c (1D, N=256, t=0.2, Sod,
 * reflective walls, reference float arithmetic).
 *
 *   shll_run [--config FILE.json] [--dim 1|2] [--order 1|2] [--nx N] [--ny N]
 *            [--precision f64|ref32|c32|c32k|f64f] [--bc reflect|outflow]
 *            [--ic sod|implosion|four-shocks|four-contacts|circle] [--limiter minmod|mc]
 *            [--theta A] [--t-end T] [--steps K] [--out PATH] [--full] [--sums] [--dump PATH]
 *            [--print-config]
 *
 * --config reads a flat JSON object whose keys are the flag names without the dashes
 * ("t_end" or "t-end"); values are strings, numbers or, for full and sums, true/false:
 *     {"dim": 2, "order": 2, "nx": 1024, "limiter": "mc", "theta": 1.25, "out": "results.dat"}
 * Options apply in command-line order, so flags after --config override the file.
 * --print-config prints the resolved options as such a JSON object and exits without
 * touching the GPU.
 * --dump writes the end state in binary: the primitives (nvar*nx*ny doubles, SoA) followed by
 * the state in its storage format (shll_get_state_raw).
 * --limiter mc (theta 1.25) with --dim 2 --order 2 --nx 1024 is base-omp/2nd_order_base_shll.c:
 * its default initial data are the four contacts and its TOTAL_TIME 0.3 (:43).
 * Exit status: 0 done, 1 a library or I/O error, 2 a usage or configuration error.
 */
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shll.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: shll_run [--config FILE.json] [--dim 1|2] [--order 1|2] [--nx N] [--ny N]\n"
            "                [--precision f64|ref32|c32|c32k|f64f] [--bc reflect|outflow]\n"
            "                [--ic sod|implosion|four-shocks|four-contacts|circle] [--limiter minmod|mc]\n"
            "                [--theta A] [--t-end T] [--steps K] [--out PATH] [--full] [--sums]\n"
            "                [--dump PATH] [--print-config]\n");
}

#define STR_MAX 1024

typedef struct {
    shll_problem pb;
    int bc_set, ny_set;
    char ic[64], out[STR_MAX], dump[STR_MAX];
    double t_end; /* < 0: the reference file's TOTAL_TIME */
    long long steps; /* < 0: run to t_end */
    int sums, full;
} options;

static int parse_ll(const char *v, long long lo, long long hi, long long *out)
{
    char *e;
    long long x;
    errno = 0;
    x = strtoll(v, &e, 10);
    if (errno || e == v || *e || x < lo || x > hi) return -1;
    *out = x;
    return 0;
}

static int parse_d(const char *v, double *out)
{
    char *e;
    double x;
    errno = 0;
    x = strtod(v, &e);
    if (errno || e == v || *e) return -1;
    *out = x;
    return 0;
}

static int parse_bool(const char *v, int *out)
{
    if (!strcmp(v, "true") || !strcmp(v, "1")) { *out = 1; return 0; }
    if (!strcmp(v, "false") || !strcmp(v, "0")) { *out = 0; return 0; }
    return -1;
}

static int copy_str(char *dst, size_t cap, const char *v)
{
    if (strlen(v) >= cap) return -1;
    strcpy(dst, v);
    return 0;
}

/* One option by name (a flag without its dashes, or a JSON key).  Returns 0, or -1 with a
 * message on stderr. */
static int set_option(options *o, const char *key, const char *v, const char *where)
{
    long long x;
    int bad = 0;
    if (!strcmp(key, "dim")) {
        bad = parse_ll(v, 1, 2, &x);
        if (!bad) o->pb.dim = (int32_t)x;
    } else if (!strcmp(key, "order")) {
        bad = parse_ll(v, 1, 2, &x);
        if (!bad) o->pb.order = (int32_t)x;
    } else if (!strcmp(key, "nx")) {
        bad = parse_ll(v, 1, 1LL << 40, &x);
        if (!bad) o->pb.nx = x;
    } else if (!strcmp(key, "ny")) {
        bad = parse_ll(v, 1, 1LL << 40, &x);
        if (!bad) { o->pb.ny = x; o->ny_set = 1; }
    } else if (!strcmp(key, "precision")) {
        if (!strcmp(v, "f64")) o->pb.precision = SHLL_F64;
        else if (!strcmp(v, "ref32")) o->pb.precision = SHLL_REF32;
        else if (!strcmp(v, "c32")) o->pb.precision = SHLL_C32;
        else if (!strcmp(v, "c32k")) o->pb.precision = SHLL_C32K;
        else if (!strcmp(v, "f64f")) o->pb.precision = SHLL_F64F;
        else bad = 1;
    } else if (!strcmp(key, "bc")) {
        if (!strcmp(v, "reflect")) o->pb.bc = SHLL_BC_REFLECT;
        else if (!strcmp(v, "outflow")) o->pb.bc = SHLL_BC_OUTFLOW;
        else bad = 1;
        if (!bad) o->bc_set = 1;
    } else if (!strcmp(key, "ic")) {
        if (strcmp(v, "sod") && strcmp(v, "implosion") && strcmp(v, "four-shocks") && strcmp(v, "four-contacts") &&
            strcmp(v, "circle"))
            bad = 1;
        else
            bad = copy_str(o->ic, sizeof(o->ic), v);
    } else if (!strcmp(key, "limiter")) {
        if (!strcmp(v, "minmod")) o->pb.limiter = SHLL_LIM_MINMOD;
        else if (!strcmp(v, "mc")) o->pb.limiter = SHLL_LIM_MC;
        else bad = 1;
    } else if (!strcmp(key, "theta")) {
        bad = parse_d(v, &o->pb.theta);
    } else if (!strcmp(key, "t-end") || !strcmp(key, "t_end")) {
        bad = parse_d(v, &o->t_end) || !(o->t_end >= 0);
    } else if (!strcmp(key, "steps")) {
        bad = parse_ll(v, -1, 1LL << 62, &o->steps);  /* -1: run to t_end */
    } else if (!strcmp(key, "out")) {
        bad = copy_str(o->out, sizeof(o->out), v);
    } else if (!strcmp(key, "dump")) {
        bad = copy_str(o->dump, sizeof(o->dump), v);
    } else if (!strcmp(key, "full")) {
        bad = parse_bool(v, &o->full);
    } else if (!strcmp(key, "sums")) {
        bad = parse_bool(v, &o->sums);
    } else {
        fprintf(stderr, "%s: unknown option '%s'\n", where, key);
        return -1;
    }
    if (bad) {
        fprintf(stderr, "%s: bad value '%s' for '%s'\n", where, v, key);
        return -1;
    }
    return 0;
}

/* A flat JSON object: {"key": value, ...} with string, number, true/false values (no
 * nesting, no null; strings without escapes other than \" \\ \/).  Each pair goes to
 * set_option.  Returns 0, or -1 with a message. */
static int read_config(options *o, const char *path)
{
    FILE *f = fopen(path, "rb");
    char *s, key[64], val[STR_MAX];
    long len;
    size_t p = 0, n;
    int first = 1, rc = -1;
    if (!f) { fprintf(stderr, "%s: cannot open\n", path); return -1; }
    if (fseek(f, 0, SEEK_END) != 0 || (len = ftell(f)) < 0 || fseek(f, 0, SEEK_SET) != 0 || len > (1 << 20)) {
        fprintf(stderr, "%s: cannot read (or larger than 1 MiB)\n", path);
        fclose(f);
        return -1;
    }
    s = (char *)malloc((size_t)len + 1);
    if (!s) { fclose(f); return -1; }
    n = fread(s, 1, (size_t)len, f);
    fclose(f);
    s[n] = 0;
#define SKIP_WS() while (p < n && isspace((unsigned char)s[p])) p++
    SKIP_WS();
    if (p >= n || s[p] != '{') { fprintf(stderr, "%s: expected a JSON object\n", path); goto done; }
    p++;
    for (;;) {
        size_t k;
        SKIP_WS();
        if (p < n && s[p] == '}' && first) { p++; break; }
        /* key */
        if (p >= n || s[p] != '"') { fprintf(stderr, "%s: expected a key at byte %zu\n", path, p); goto done; }
        p++;
        for (k = 0; p < n && s[p] != '"'; p++) {
            if (k + 1 >= sizeof(key)) { fprintf(stderr, "%s: key too long at byte %zu\n", path, p); goto done; }
            key[k++] = s[p];
        }
        key[k] = 0;
        if (p >= n) { fprintf(stderr, "%s: unterminated key\n", path); goto done; }
        p++;
        SKIP_WS();
        if (p >= n || s[p] != ':') { fprintf(stderr, "%s: expected ':' after \"%s\"\n", path, key); goto done; }
        p++;
        SKIP_WS();
        /* value */
        k = 0;
        if (p < n && s[p] == '"') {
            for (p++; p < n && s[p] != '"'; p++) {
                char c = s[p];
                if (c == '\\') {
                    if (p + 1 >= n || (s[p + 1] != '"' && s[p + 1] != '\\' && s[p + 1] != '/')) {
                        fprintf(stderr, "%s: unsupported escape in the value of \"%s\"\n", path, key);
                        goto done;
                    }
                    c = s[++p];
                }
                if (k + 1 >= sizeof(val)) { fprintf(stderr, "%s: value of \"%s\" too long\n", path, key); goto done; }
                val[k++] = c;
            }
            if (p >= n) { fprintf(stderr, "%s: unterminated string\n", path); goto done; }
            p++;
        } else {
            while (p < n && (isalnum((unsigned char)s[p]) || s[p] == '.' || s[p] == '-' || s[p] == '+')) {
                if (k + 1 >= sizeof(val)) { fprintf(stderr, "%s: value of \"%s\" too long\n", path, key); goto done; }
                val[k++] = s[p++];
            }
            if (k == 0) { fprintf(stderr, "%s: expected a value for \"%s\"\n", path, key); goto done; }
        }
        val[k] = 0;
        if (set_option(o, key, val, path) != 0) goto done;
        first = 0;
        SKIP_WS();
        if (p < n && s[p] == ',') { p++; continue; }
        if (p < n && s[p] == '}') { p++; break; }
        fprintf(stderr, "%s: expected ',' or '}' at byte %zu\n", path, p);
        goto done;
    }
    SKIP_WS();
    if (p != n) { fprintf(stderr, "%s: trailing characters after the object\n", path); goto done; }
    rc = 0;
#undef SKIP_WS
done:
    free(s);
    return rc;
}

static void print_json_str(const char *s)
{
    putchar('"');
    for (; *s; s++) {
        if (*s == '"' || *s == '\\') putchar('\\');
        putchar(*s);
    }
    putchar('"');
}

int main(int argc, char **argv)
{
    options o;
    int i, nv, print_only = 0;
    double *prim;
    shll_ctx *ctx = NULL;
    shll_status st;
    int64_t done = 0, n;
    static const char *prec_names[] = {"f64", "ref32", "c32", "c32k", "f64f"};

    memset(&o, 0, sizeof(o));
    o.pb.dim = 1; o.pb.order = 1; o.pb.bc = SHLL_BC_REFLECT; o.pb.precision = SHLL_REF32;
    o.pb.nx = 256; o.pb.ny = 1; o.pb.lx = 1.0; o.pb.ly = 1.0;
    o.pb.limiter = SHLL_LIM_MINMOD; o.pb.theta = 1.25;
    o.t_end = -1.0;
    o.steps = -1;
    strcpy(o.out, "results.dat");
    for (i = 1; i < argc; i++) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--sums")) { o.sums = 1; continue; }
        if (!strcmp(a, "--full")) { o.full = 1; continue; }
        if (!strcmp(a, "--print-config")) { print_only = 1; continue; }
        if (strncmp(a, "--", 2) != 0 || !v) { usage(); return 2; }
        if (!strcmp(a, "--config")) {
            if (read_config(&o, v) != 0) return 2;
        } else if (set_option(&o, a + 2, v, "command line") != 0) {
            usage();
            return 2;
        }
        i++;
    }
    if (o.pb.dim == 1) o.pb.ny = 1;
    else if (!o.ny_set) o.pb.ny = o.pb.nx;
    if (!o.bc_set) o.pb.bc = o.pb.order == 1 ? SHLL_BC_REFLECT : SHLL_BC_OUTFLOW;
    if (!o.ic[0])
        strcpy(o.ic, o.pb.dim == 1 ? "sod"
                     : (o.pb.order == 1 ? "implosion"
                                        : (o.pb.limiter == SHLL_LIM_MC ? "four-contacts" : "four-shocks")));
    /* the library's limits (shll_init rejects the rest), checked before the host buffer of
     * nvar*nx*ny doubles is sized from them */
    if (o.pb.nx < 2 || o.pb.nx > (1LL << 30) ||
        (o.pb.dim == 2 && (o.pb.ny < 2 || o.pb.ny > (1LL << 30) || (double)o.pb.nx * (double)o.pb.ny > 4.0e9))) {
        fprintf(stderr, "grid of %lld x %lld cells out of range (2 to 2^30 a side, at most 4e9 cells)\n",
                (long long)o.pb.nx, (long long)o.pb.ny);
        return 2;
    }
    if (o.pb.dim == 1 && strcmp(o.ic, "sod") != 0) {
        fprintf(stderr, "initial data '%s' is 2D; 1D runs take 'sod'\n", o.ic);
        return 2;
    }
    if (o.pb.dim == 2 && !strcmp(o.ic, "sod")) {
        fprintf(stderr, "initial data 'sod' is 1D\n");
        return 2;
    }
    /* TOTAL_TIME of each reference file (base_shll.c:26, base_shll_2d.c:35, 2nd_order_base_shll.c:43,
     * base-omp/2nd_order_base_shll.c:43) */
    if (o.t_end < 0) o.t_end = o.pb.dim == 1 ? 0.2 : (o.pb.order == 1 ? 0.1 : (o.pb.limiter == SHLL_LIM_MC ? 0.3 : 0.8));

    if (print_only) {
        printf("{\"dim\": %d, \"order\": %d, \"nx\": %lld, \"ny\": %lld, \"precision\": ", (int)o.pb.dim,
               (int)o.pb.order, (long long)o.pb.nx, (long long)o.pb.ny);
        print_json_str(prec_names[o.pb.precision]);
        printf(", \"bc\": \"%s\", \"ic\": ", o.pb.bc == SHLL_BC_OUTFLOW ? "outflow" : "reflect");
        print_json_str(o.ic);
        printf(", \"limiter\": \"%s\", \"theta\": %.17g, \"t_end\": %.9g, \"steps\": %lld, \"out\": ",
               o.pb.limiter == SHLL_LIM_MC ? "mc" : "minmod", o.pb.theta, (double)(float)o.t_end, o.steps);
        print_json_str(o.out);
        printf(", \"dump\": ");
        print_json_str(o.dump);
        printf(", \"full\": %s, \"sums\": %s}\n", o.full ? "true" : "false", o.sums ? "true" : "false");
        return 0;
    }

    nv = o.pb.dim == 1 ? 3 : 4;
    prim = (double *)malloc(sizeof(double) * (size_t)nv * (size_t)(o.pb.nx * o.pb.ny));
    if (!prim) { fprintf(stderr, "out of memory\n"); return 1; }
    if (!strcmp(o.ic, "sod")) shll_ic_tube(o.pb.nx, 0.5, 10.0, 0.0, 1.0, 1.0, 0.0, 1.0, prim);
    else if (!strcmp(o.ic, "implosion")) shll_ic_implosion(o.pb.nx, o.pb.ny, prim);
    else if (!strcmp(o.ic, "four-shocks")) shll_ic_four_shocks(o.pb.nx, o.pb.ny, prim);
    else if (!strcmp(o.ic, "four-contacts")) shll_ic_four_contacts(o.pb.nx, o.pb.ny, prim);
    else shll_ic_circle(o.pb.nx, o.pb.ny, 1.0, 1.0, 0.5, 0.5, 0.25, 10.0, 1.0, 1.0, 1.0, prim);

    st = shll_init(&o.pb, prim, &ctx);
    free(prim);
    if (st != SHLL_OK) { fprintf(stderr, "shll_init: %s\n", shll_last_error()); return 1; }
    if (o.steps >= 0) {
        st = shll_step(ctx, o.steps);
        done = o.steps;
        if (st == SHLL_OK) st = shll_check_state(ctx, NULL);
    } else {
        st = shll_run_to_time(ctx, (float)o.t_end, &done);
    }
    if (st != SHLL_OK) { fprintf(stderr, "step loop: %s\n", shll_last_error()); shll_free(ctx); return 1; }
    printf("Completed in %lld steps\n", (long long)done);
    if (o.sums) {
        double s[4];
        if (shll_conserved_sums(ctx, s) == SHLL_OK) {
            printf("sums:");
            for (n = 0; n < nv; n++) printf(" %.17g", s[n]);
            printf("\n");
        }
    }
    if (o.dump[0]) {
        size_t ncell = (size_t)(o.pb.nx * o.pb.ny), raw = shll_state_bytes(ctx);
        double *p = (double *)malloc(sizeof(double) * (size_t)nv * ncell);
        void *r = malloc(raw);
        FILE *f = fopen(o.dump, "wb");
        st = (p && r && f) ? shll_get_primitives(ctx, p) : SHLL_ENOMEM;
        if (st == SHLL_OK) st = shll_get_state_raw(ctx, r, raw);
        if (st == SHLL_OK && (fwrite(p, sizeof(double), (size_t)nv * ncell, f) != (size_t)nv * ncell ||
                              fwrite(r, 1, raw, f) != raw))
            st = SHLL_EINVAL;
        if (f && fclose(f) != 0) st = SHLL_EINVAL;
        free(p);
        free(r);
        if (st != SHLL_OK) {
            fprintf(stderr, "dump to %s failed: %s\n", o.dump, shll_last_error());
            shll_free(ctx);
            return 1;
        }
    }
    if (strcmp(o.out, "-") != 0) {
        if ((o.full ? shll_write_results_full(ctx, o.out) : shll_write_results(ctx, o.out)) != SHLL_OK) {
            fprintf(stderr, "write: %s\n", shll_last_error());
            shll_free(ctx);
            return 1;
        }
    }
    shll_free(ctx);
    return 0;
}
