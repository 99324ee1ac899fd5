/*
 * infoli_host.c -- host side of the C ABI: configuration, seeded synthetic
 * inputs and the reference's file formats.  Plain C, no device code.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "infoli_host.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */

static __thread char g_errmsg[1024];

int infoli_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof g_errmsg, fmt, ap);
    va_end(ap);
    return code;
}

void infoli_clear_error(void) { g_errmsg[0] = '\0'; }

const char *infoli_last_error(void) { return g_errmsg; }

const char *infoli_strerror(int code)
{
    switch (code) {
    case INFOLI_OK: return "ok";
    case INFOLI_ERR_ARG: return "invalid argument";
    case INFOLI_ERR_HIP: return "HIP runtime error";
    case INFOLI_ERR_NOMEM: return "out of memory";
    case INFOLI_ERR_STATE: return "call out of order";
    case INFOLI_ERR_NONFINITE: return "non-finite voltage";
    case INFOLI_ERR_IO: return "file input/output error";
    case INFOLI_ERR_CONFIG: return "malformed configuration";
    case INFOLI_ERR_BARRIER: return "grid barrier timed out";
    case INFOLI_ERR_NODEVICE: return "no usable device";
    default: return "unknown error";
    }
}

/* ------------------------------------------------------------------------ */
/* configuration: bin/default.conf                                           */
/* ------------------------------------------------------------------------ */

/* The reference's default model (bin/default.conf:1-42), as text so that the
 * float and double readings are the ones a parser of that file produces. */
static const char *const k_default_text[INFOLI_NPARAMS] = {
    "0.05", "0.04", "1", "150", "9", "5", "0.016", "35", "4.5", "0.016",
    "0.125", "240", "0", "20", "0.016", "0.25", "0.15", "0.13", "55", "-75",
    "120", "-43", "10",
    "-60", "0.0112788", "0.0049291", "0.0337836", "3.7152", "0.5", "0.68",
    "-60", "1.0127807", "0.3596066", "0.2369847", "0.2369847", "0.1",
    "0.7423159", "0.0321349", "-60", "0.003596066", "0.9", "0.2369847"};

int infoli_params_default(infoli_params *out)
{
    if (!out) return infoli_fail(INFOLI_ERR_ARG, "infoli_params_default: null output");
    for (int i = 0; i < INFOLI_NPARAMS; i++) {
        out->v[i] = strtod(k_default_text[i], NULL);
        out->f[i] = strtof(k_default_text[i], NULL);
    }
    return INFOLI_OK;
}

/* read_parameters (src/utils/utils.c:137-160): 42 entries "NAME = value";
 * whitespace (including blank lines) between entries is skipped, the names
 * are not checked.  The float reading uses strtof, as fscanf("%f") does. */
int infoli_read_config(const char *path, infoli_params *out)
{
    if (!path || !out) return infoli_fail(INFOLI_ERR_ARG, "infoli_read_config: null argument");
    FILE *fd = fopen(path, "r");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "Error: Couldn't open %s", path);
    char name[256], value[256];
    for (int i = 0; i < INFOLI_NPARAMS; i++) {
        int got = fscanf(fd, " %255s = %255s", name, value);
        char *end = NULL;
        double dv = 0;
        if (got == 2) {
            errno = 0;
            dv = strtod(value, &end);
        }
        if (got != 2 || end == value || *end != '\0') {
            fclose(fd);
            return infoli_fail(INFOLI_ERR_CONFIG,
                               "Configuration Error: Improper Parameter %s in file %s, line %d.",
                               got >= 1 ? name : "", path, i + 1);
        }
        out->v[i] = dv;
        out->f[i] = strtof(value, NULL);
    }
    fclose(fd);
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* counter-based random numbers                                              */
/* ------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t x)
{
    /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t infoli_rng_bits(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b)
{
    uint64_t h = mix64(seed + stream * 0xD1B54A32D192ED03ull);
    h = mix64(h + a * 0x9E6C63D0676A9A99ull);
    return mix64(h + b);
}

double infoli_rng_u01(uint64_t bits)
{
    return (double)((bits >> 11) + 1) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------ */
/* graphs                                                                    */
/* ------------------------------------------------------------------------ */

/* Geometric skipping over a run of `len` Bernoulli(p) candidates: calls
 * emit(pos) for each success, pos ascending.  Draw k of run `row` uses the
 * bits (seed, stream, row, k). */
typedef void (*emit_fn)(void *user, int64_t row, int64_t pos);

static void skip_run(uint64_t seed, uint64_t stream, int64_t row, int64_t len,
                     double p, emit_fn emit, void *user)
{
    if (len <= 0 || !(p > 0)) return;
    if (p >= 1) {
        for (int64_t x = 0; x < len; x++) emit(user, row, x);
        return;
    }
    const double lq = log1p(-p);
    int64_t x = -1;
    for (uint64_t k = 0;; k++) {
        double u = infoli_rng_u01(infoli_rng_bits(seed, stream, (uint64_t)row, k));
        double skip = floor(log(u) / lq);
        if (!(skip < (double)(len - 1 - x))) break;
        x += (int64_t)skip + 1;
        emit(user, row, x);
    }
}

typedef struct {
    int64_t n;
    int64_t *row_ptr;   /* counting pass: degrees in row_ptr[i+1] */
    int64_t *cursor;    /* filling pass */
    int32_t *col;
    double *w;
    uint64_t seed;
    double w_lo, w_hi;
} gen_ctx;

static void sym_count(void *user, int64_t i, int64_t x)
{
    gen_ctx *g = (gen_ctx *)user;
    int64_t j = i + 1 + x;
    g->row_ptr[i + 1]++;
    g->row_ptr[j + 1]++;
}

static void sym_fill(void *user, int64_t i, int64_t x)
{
    gen_ctx *g = (gen_ctx *)user;
    int64_t j = i + 1 + x;
    double wv = g->w_lo;
    if (g->w_hi != g->w_lo)
        wv = g->w_lo + (g->w_hi - g->w_lo) *
                           infoli_rng_u01(infoli_rng_bits(g->seed, INFOLI_STREAM_W,
                                                          (uint64_t)i, (uint64_t)j));
    int64_t a = g->cursor[i]++, b = g->cursor[j]++;
    g->col[a] = (int32_t)j; g->w[a] = wv;
    g->col[b] = (int32_t)i; g->w[b] = wv;
}

static int check_gen_args(int64_t n, double p, int64_t *row_ptr, int64_t *nnz,
                          int32_t *col, double *w, const char *who)
{
    if (n <= 0 || n > INT32_MAX)
        return infoli_fail(INFOLI_ERR_ARG, "%s: n = %lld out of range", who, (long long)n);
    if (!(p >= 0 && p <= 1))
        return infoli_fail(INFOLI_ERR_ARG, "%s: p = %g not in [0, 1]", who, p);
    if (!row_ptr || !nnz)
        return infoli_fail(INFOLI_ERR_ARG, "%s: null row_ptr or nnz", who);
    if ((col == NULL) != (w == NULL))
        return infoli_fail(INFOLI_ERR_ARG, "%s: col and w must both be given or both be NULL", who);
    return INFOLI_OK;
}

int infoli_gen_symmetric(uint64_t seed, int64_t n, double p, double w_lo,
                         double w_hi, int64_t *row_ptr, int32_t *col,
                         double *w, int64_t *nnz)
{
    int rc = check_gen_args(n, p, row_ptr, nnz, col, w, "infoli_gen_symmetric");
    if (rc) return rc;
    gen_ctx g = {n, row_ptr, NULL, col, w, seed, w_lo, w_hi};
    memset(row_ptr, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n - 1; i++)
        skip_run(seed, INFOLI_STREAM_GRAPH, i, n - 1 - i, p, sym_count, &g);
    for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
    *nnz = row_ptr[n];
    if (!col) return INFOLI_OK;
    g.cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!g.cursor) return infoli_fail(INFOLI_ERR_NOMEM, "infoli_gen_symmetric: no memory");
    memcpy(g.cursor, row_ptr, sizeof(int64_t) * (size_t)n);
    /* row i ascending: lower neighbours of a row arrive (ascending) before
     * the row emits its own upper neighbours (ascending) */
    for (int64_t i = 0; i < n - 1; i++)
        skip_run(seed, INFOLI_STREAM_GRAPH, i, n - 1 - i, p, sym_fill, &g);
    free(g.cursor);
    return INFOLI_OK;
}

static void dir_count(void *user, int64_t i, int64_t x)
{
    (void)x;
    ((gen_ctx *)user)->row_ptr[i + 1]++;
}

static void dir_fill(void *user, int64_t i, int64_t x)
{
    gen_ctx *g = (gen_ctx *)user;
    int64_t j = x < i ? x : x + 1;   /* skip the receiver itself (OMP:343-344) */
    int64_t a = g->cursor[i]++;
    g->col[a] = (int32_t)j;
    g->w[a] = g->w_lo;
}

int infoli_gen_directed(uint64_t seed, int64_t n, double p, double w0,
                        int64_t *row_ptr, int32_t *col, double *w, int64_t *nnz)
{
    int rc = check_gen_args(n, p, row_ptr, nnz, col, w, "infoli_gen_directed");
    if (rc) return rc;
    gen_ctx g = {n, row_ptr, NULL, col, w, seed, w0, w0};
    memset(row_ptr, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; i++)
        skip_run(seed, INFOLI_STREAM_DIRECTED, i, n - 1, p, dir_count, &g);
    for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
    *nnz = row_ptr[n];
    if (!col) return INFOLI_OK;
    g.cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!g.cursor) return infoli_fail(INFOLI_ERR_NOMEM, "infoli_gen_directed: no memory");
    memcpy(g.cursor, row_ptr, sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++)
        skip_run(seed, INFOLI_STREAM_DIRECTED, i, n - 1, p, dir_fill, &g);
    free(g.cursor);
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* per-cell synthetic values                                                 */
/* ------------------------------------------------------------------------ */

int infoli_gen_uniform(uint64_t seed, int stream, int64_t n, double lo,
                       double hi, double *out)
{
    if (n < 0 || (n > 0 && !out)) return infoli_fail(INFOLI_ERR_ARG, "infoli_gen_uniform: bad n/out");
    for (int64_t i = 0; i < n; i++)
        out[i] = (hi == lo) ? lo
                 : lo + (hi - lo) * infoli_rng_u01(infoli_rng_bits(seed, (uint64_t)stream, (uint64_t)i, 0));
    return INFOLI_OK;
}

int infoli_gen_normal(uint64_t seed, int stream, int64_t n, double mean,
                      double sd, double *out)
{
    if (n < 0 || (n > 0 && !out)) return infoli_fail(INFOLI_ERR_ARG, "infoli_gen_normal: bad n/out");
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i++) {
        double u1 = infoli_rng_u01(infoli_rng_bits(seed, (uint64_t)stream, (uint64_t)i, 0));
        double u2 = infoli_rng_u01(infoli_rng_bits(seed, (uint64_t)stream, (uint64_t)i, 1));
        out[i] = mean + sd * (sqrt(-2.0 * log(u1)) * cos(two_pi * u2));
    }
    return INFOLI_OK;
}

int infoli_gen_mask(uint64_t seed, int stream, int64_t n, double frac, uint8_t *mask)
{
    if (n < 0 || (n > 0 && !mask)) return infoli_fail(INFOLI_ERR_ARG, "infoli_gen_mask: bad n/mask");
    for (int64_t i = 0; i < n; i++)
        mask[i] = infoli_rng_u01(infoli_rng_bits(seed, (uint64_t)stream, (uint64_t)i, 0)) <= frac;
    return INFOLI_OK;
}

/* RAND_INIT's per-cell randomness (src/single_threaded/infoli.c:212-222)
 * from the counter generator, stream 7: g_CaL = 0.6 + 0.2*(k % 100)/100
 * and warm-up steps (u % ceil(100/DELTA)) | 1, k and u the two 64-bit
 * draws of cell i. */
int infoli_gen_rand_init(uint64_t seed, int64_t n, double delta, double *g_cal, int64_t *steps)
{
    if (n < 0 || (n > 0 && (!g_cal || !steps)) || !(delta > 0))
        return infoli_fail(INFOLI_ERR_ARG, "infoli_gen_rand_init: bad n/delta/outputs");
    const uint64_t span = (uint64_t)ceil(100 / delta);
    for (int64_t i = 0; i < n; i++) {
        const uint64_t k = infoli_rng_bits(seed, 7, (uint64_t)i, 0);
        const uint64_t u = infoli_rng_bits(seed, 7, (uint64_t)i, 1);
        g_cal[i] = 0.6 + (0.2 * (double)(k % 100)) / 100;
        steps[i] = (int64_t)((u % span) | 1u);
    }
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* files                                                                     *//* ------------------------------------------------------------------------ */
/* files                                                                     */
/* ------------------------------------------------------------------------ */

static int is_junction(double v, double threshold)
{
    /* threshold 0: the ST rule "cond_value == 0 is no junction" (ST:171);
     * threshold > 0: stable_version's "value < threshold is none" (:204) */
    return threshold > 0 ? (v >= threshold) : (v != 0);
}

/* dense [sender][receiver] matrix: ST:158-191 reads it row by row, and a
 * non-zero entry (row s, column r) makes s a neighbour of r, so each
 * receiver's senders are ascending. */
int infoli_read_conn_dense(const char *path, int64_t n, double threshold,
                           int64_t *row_ptr, int32_t *col, double *w, int64_t *nnz)
{
    if (!path || n <= 0 || n > INT32_MAX || !row_ptr || !nnz || ((col == NULL) != (w == NULL)))
        return infoli_fail(INFOLI_ERR_ARG, "infoli_read_conn_dense: bad argument");
    FILE *fd = fopen(path, "r");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "Error: Couldn't open %s", path);
    int64_t *cursor = NULL;
    memset(row_ptr, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int pass = 0; pass < (col ? 2 : 1); pass++) {
        if (pass == 1) {
            for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
            cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
            if (!cursor) { fclose(fd); return infoli_fail(INFOLI_ERR_NOMEM, "no memory"); }
            memcpy(cursor, row_ptr, sizeof(int64_t) * (size_t)n);
            rewind(fd);
        }
        for (int64_t s = 0; s < n; s++) {
            for (int64_t r = 0; r < n; r++) {
                double v;
                if (fscanf(fd, "%lf", &v) != 1) {
                    free(cursor);
                    fclose(fd);
                    return infoli_fail(INFOLI_ERR_IO,
                                       "%s: expected %lld x %lld values, short at row %lld column %lld",
                                       path, (long long)n, (long long)n, (long long)s, (long long)r);
                }
                if (!is_junction(v, threshold)) continue;
                if (pass == 0) row_ptr[r + 1]++;
                else { int64_t a = cursor[r]++; col[a] = (int32_t)s; w[a] = v; }
            }
        }
    }
    if (!col) for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
    *nnz = row_ptr[n];
    free(cursor);
    fclose(fd);
    return INFOLI_OK;
}

/* "send rec cond" lines (stable_version.c:226-252, truncated/openmp-oldAPI
 * :321-345): each accepted line appends send to rec's list, file order. */
int infoli_read_conn_triplets(const char *path, int64_t n, double threshold,
                              int64_t *row_ptr, int32_t *col, double *w, int64_t *nnz)
{
    if (!path || n <= 0 || n > INT32_MAX || !row_ptr || !nnz || ((col == NULL) != (w == NULL)))
        return infoli_fail(INFOLI_ERR_ARG, "infoli_read_conn_triplets: bad argument");
    FILE *fd = fopen(path, "r");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "Error: Couldn't open %s", path);
    int64_t *cursor = NULL;
    memset(row_ptr, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int pass = 0; pass < (col ? 2 : 1); pass++) {
        if (pass == 1) {
            for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
            cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
            if (!cursor) { fclose(fd); return infoli_fail(INFOLI_ERR_NOMEM, "no memory"); }
            memcpy(cursor, row_ptr, sizeof(int64_t) * (size_t)n);
            rewind(fd);
        }
        long long snd, rcv;
        double v;
        int64_t line = 0;
        int got;
        while ((got = fscanf(fd, "%lld %lld %lf", &snd, &rcv, &v)) == 3) {
            line++;
            if (snd < 0 || snd >= n || rcv < 0 || rcv >= n) {
                free(cursor);
                fclose(fd);
                return infoli_fail(INFOLI_ERR_IO, "%s: line %lld: cell id out of range [0, %lld)",
                                   path, (long long)line, (long long)n);
            }
            if (!is_junction(v, threshold)) continue;
            if (pass == 0) row_ptr[rcv + 1]++;
            else { int64_t a = cursor[rcv]++; col[a] = (int32_t)snd; w[a] = v; }
        }
        if (got != EOF) {
            free(cursor);
            fclose(fd);
            return infoli_fail(INFOLI_ERR_IO, "%s: malformed triplet after line %lld", path,
                               (long long)line);
        }
    }
    if (!col) for (int64_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
    *nnz = row_ptr[n];
    free(cursor);
    fclose(fd);
    return INFOLI_OK;
}

int infoli_read_values(const char *path, int64_t n, double *out)
{
    if (!path || n <= 0 || !out) return infoli_fail(INFOLI_ERR_ARG, "infoli_read_values: bad argument");
    FILE *fd = fopen(path, "r");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "Error: Couldn't open %s", path);
    for (int64_t i = 0; i < n; i++) {
        if (fscanf(fd, "%lf", &out[i]) != 1) {
            fclose(fd);
            return infoli_fail(INFOLI_ERR_IO, "%s: expected %lld values, found %lld", path,
                               (long long)n, (long long)i);
        }
    }
    fclose(fd);
    return INFOLI_OK;
}

/* ReadFileLine (src/utils/utils.c:70-122): one step per line, the first n
 * space-separated values are the cells' Iapp; fewer than n is an error
 * ("Input line doesn't have enough elements"). */
int infoli_read_iapp_file(const char *path, int64_t n, int64_t max_steps,
                          double *table, int64_t *n_steps)
{
    if (!path || n <= 0 || max_steps < 0 || !n_steps)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_read_iapp_file: bad argument");
    FILE *fd = fopen(path, "r");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "Error: Couldn't open %s", path);
    char *line = NULL;
    size_t cap = 0;
    int64_t s = 0;
    while (s < max_steps && getline(&line, &cap, fd) != -1) {
        char *save = NULL, *tok = strtok_r(line, " \t\r\n", &save);
        int64_t i = 0;
        for (; tok && i < n; i++, tok = strtok_r(NULL, " \t\r\n", &save))
            if (table) table[s * n + i] = atof(tok);
        if (i < n) {
            free(line);
            fclose(fd);
            return infoli_fail(INFOLI_ERR_IO,
                               "Error: Input line doesn't have enough elements, only %lld (line %lld of %s)",
                               (long long)i, (long long)(s + 1), path);
        }
        s++;
    }
    free(line);
    fclose(fd);
    *n_steps = s;
    return INFOLI_OK;
}
