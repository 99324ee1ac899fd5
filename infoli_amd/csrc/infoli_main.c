/*
 * infoli_main.c -- command-line driver of the GPU InfOli simulator.
 *
 * Keeps the reference simulator's interface (src/openmp/infoli.c:43-735):
 *   infoli.x -n <cells> -p <density> -t <ms> [-i <iapp file>] [-c <config>]
 *            [-o <output>]
 * with the same defaults (1000 cells, p = 0.5, 5000 ms), the same config
 * lookup (default.conf next to the executable, OMP:113-121), the same
 * connectivity semantics (each receiver takes each other cell as a sender
 * with probability p, conductance CONDUCTANCE, OMP:337-375 -- seeded here)
 * and the same trace file format (OMP:383-396, 486-491, 682-696, 723-732).
 * The per-timestep update runs on the GPU through include/infoli_gpu.h.
 *
 * Additional options (no reference counterpart unless noted):
 *   -g <file>  per-cell g_CaL from a file and the regression trace format
 *              (the reference's -DG_CAL_FROM_FILE build, OMP:195-197,294-296)
 *   -m <file>  dense [sender][receiver] conductance matrix (ST:158-191)
 *   -r <file>  "send rec cond" triplets (stable_version.c:226-252)
 *   -x <prec>  fp32 (default, the reference's float), fp32k, fp64
 *   -s <seed>  connectivity seed (default: time, as srand(time(NULL)))
 *   -k <steps> trace every k steps (default 100, OMP:194; 1 under -g)
 *   -D <dev>   HIP device
 *   -a         autotune the launch strategy before the run
 */
#define _GNU_SOURCE
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include "infoli_gpu.h"

static void print_usage(void)
{
    /* src/utils/utils.c:13-26 */
    printf("Usage for infoli simulator:\n");
    printf("/path/to/bin/infoli.x -n <integer> -p <[0,1]> -t <float>\n");
    printf("-i <filename> -c <filename> -o <filename>\n");
    printf("Argument summary:\n");
    printf("-n denotes number of neurons to be simulated. Default value: 1000.\n");
    printf("-p denotes network density in [0,1] format. Default value: 0.5 (=50%%).\n");
    printf("-t denotes simulation time in milliseconds. Default value: 5000.\n");
    printf("-i denotes a filename for input stimulus. By default, stimulus is hardcoded.\n");
    printf("-c denotes a filename for configuration file. By default, configuration is read from default.conf.\n");
    printf("-o denotes a filename for simulation output. Default value: InferiorOlive_Output.txt.\n");
    printf("GPU build options: -g <gcal file> -m <conn matrix> -r <conn triplets> -x fp32|fp32k|fp64\n");
    printf("                   -s <seed> -k <trace stride> -D <device> -a (autotune)\n");
    printf("                   -b <degree> (fp32k: FP64 sums for cells of at least this degree; 0 = none)\n");
    printf("                   -R <seed> (RAND_INIT: random g_CaL and isolated warm-up) -S <state dump file>\n");
    printf("                   -L <checkpoint to resume from> -C <checkpoint to write at the end>\n");
}

static void die(int rc, const char *what)
{
    printf("Error: %s: %s\n", what, infoli_last_error()[0] ? infoli_last_error() : infoli_strerror(rc));
    exit(EXIT_FAILURE);
}

/* One chunk of trace records to format (OMP:682-696): steps [t, t + k),
 * records at multiples of stride, nrec of them in v. */
struct fmt_job {
    FILE *out;
    const double *v;
    int64_t t, k, stride, n, nrec;
    int gcal_mode, file_mode;
    double pulse_amp;
    int64_t pulse_t0, pulse_t1;
    int bad;
};

static void *format_records(void *arg)
{
    struct fmt_job *j = (struct fmt_job *)arg;
    const int64_t n = j->n;
    int64_t r = 0;
    for (int64_t s = j->t; s < j->t + j->k; s++) {
        if (s % j->stride) continue;
        const double *v = j->v + r * n;
        if (j->gcal_mode) {
            for (int64_t i = 0; i < n; i++) fprintf(j->out, "%.8f ", v[i]);
        } else if (j->file_mode) {
            fprintf(j->out, "%lld %.2f ", (long long)(s + 1), s * 0.05);
            for (int64_t i = 0; i < n; i++) fprintf(j->out, "%.16f ", v[i]);
        } else {
            const double iapp = (s >= j->pulse_t0 && s < j->pulse_t1) ? j->pulse_amp : 0.0;
            fprintf(j->out, " %lld %.2f ", (long long)(s + 1), iapp);
            for (int64_t i = 0; i < n; i++) fprintf(j->out, "%lld : %.8f ", (long long)(i + 1), v[i]);
        }
        fputs("\n", j->out);
        r++;
    }
    j->bad = (r != j->nrec);
    return NULL;
}

static double now_ms(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

int main(int argc, char **argv)
{
    int64_t n = 1000;
    double p = 0.5, sim_time = 5000;
    const char *in_file = NULL, *gcal_file = NULL, *conn_dense = NULL, *conn_trip = NULL;
    const char *out_file = "InferiorOlive_Output.txt";
    char params_file[PATH_MAX + 32];
    int have_params_file = 0, precision = INFOLI_FP32, device = 0, autotune = 0;
    uint64_t seed = (uint64_t)time(NULL);
    int64_t stride = 100;
    int stride_given = 0;
    int64_t fp64_degree = -1;   /* -1: the library's default */
    int rand_init = 0;
    uint64_t rand_seed = 0;
    const char *state_file = NULL, *ck_in = NULL, *ck_out = NULL;

    /* config next to the executable (OMP:113-121) */
    ssize_t len = readlink("/proc/self/exe", params_file, PATH_MAX);
    if (len < 0) {
        printf("Error: Couldn't read simlink to running executable.\n");
        exit(EXIT_FAILURE);
    }
    params_file[len] = '\0';
    char *slash = strrchr(params_file, '/');
    if (slash) slash[1] = '\0';
    strcat(params_file, "default.conf");

    int opt;
    while ((opt = getopt(argc, argv, "hn:p:t:i:c:o:g:m:r:x:s:k:D:ab:R:S:L:C:")) != -1) {
        switch (opt) {
        case 'h': print_usage(); return 0;
        case 'n':
            n = atoll(optarg);
            if (n <= 0) {
                printf("Incorrect network size argument. Using default value of 1000.\n");
                n = 1000;
            }
            break;
        case 'p':
            p = atof(optarg);
            if (p < 0) {
                printf("Incorrect network density argument. Using default value of 0.5.\n");
                p = 0.5;
            }
            if (p == 0) printf("Warning: Simulating non-connected network (network density = 0).\n");
            if (p > 1) {
                printf("Warning: Setting network density to 1 (=100%%).\n");
                p = 1;
            }
            break;
        case 't':
            sim_time = atof(optarg);
            if (sim_time <= 0) {
                printf("Incorrect simulation time argument. Using default value of 5000ms.\n");
                sim_time = 5000;
            }
            break;
        case 'i': in_file = optarg; break;
        case 'c':
            snprintf(params_file, sizeof params_file, "%s", optarg);
            have_params_file = 1;
            break;
        case 'o': out_file = optarg; break;
        case 'g': gcal_file = optarg; break;
        case 'm': conn_dense = optarg; break;
        case 'r': conn_trip = optarg; break;
        case 'x':
            if (!strcmp(optarg, "fp64")) precision = INFOLI_FP64;
            else if (!strcmp(optarg, "fp32")) precision = INFOLI_FP32;
            else if (!strcmp(optarg, "fp32k")) precision = INFOLI_FP32_KAHAN;
            else { print_usage(); exit(EXIT_FAILURE); }
            break;
        case 's': seed = strtoull(optarg, NULL, 10); break;
        case 'k':
            stride = atoll(optarg);
            stride_given = 1;
            if (stride <= 0) { print_usage(); exit(EXIT_FAILURE); }
            break;
        case 'D': device = atoi(optarg); break;
        case 'a': autotune = 1; break;
        case 'b':
            fp64_degree = strtoll(optarg, NULL, 10);
            if (fp64_degree < 0) { print_usage(); exit(EXIT_FAILURE); }
            break;
        case 'R': rand_init = 1; rand_seed = strtoull(optarg, NULL, 10); break;
        case 'S': state_file = optarg; break;
        case 'L': ck_in = optarg; break;
        case 'C': ck_out = optarg; break;
        default: print_usage(); exit(EXIT_FAILURE);
        }
    }
    if (gcal_file && !stride_given) stride = 1;   /* print_granularity = 1 (OMP:195-197) */

    infoli_params params;
    struct stat sb;
    int rc;
    if (have_params_file || stat(params_file, &sb) == 0) {
        rc = infoli_read_config(params_file, &params);
        if (rc) {
            printf("%s\n", infoli_last_error());
            if (rc == INFOLI_ERR_CONFIG)
                printf("For a proper Configuration file format, please see default.conf in bin folder.\n");
            exit(EXIT_FAILURE);
        }
    } else {
        infoli_params_default(&params);
    }

    /* connectivity */
    int64_t *row_ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!row_ptr) { printf("Error: Couldn't malloc for cellParamsPtr\n"); exit(EXIT_FAILURE); }
    int64_t nnz = 0;
    int32_t *col = NULL;
    double *w = NULL;
    for (int pass = 0; pass < 2; pass++) {
        if (conn_dense) rc = infoli_read_conn_dense(conn_dense, n, 0.0, row_ptr, col, w, &nnz);
        else if (conn_trip) rc = infoli_read_conn_triplets(conn_trip, n, 0.0, row_ptr, col, w, &nnz);
        else rc = infoli_gen_directed(seed, n, p, params.v[INFOLI_P_CONDUCTANCE], row_ptr, col, w, &nnz);
        if (rc) die(rc, "connectivity");
        if (pass == 0) {
            col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
            w = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
            if (!col || !w) { printf("Error: Couldn't malloc for cellParamsPtr\n"); exit(EXIT_FAILURE); }
        }
    }

    infoli_t *net = NULL;
    rc = infoli_create(&net, &params, n, precision, device);
    if (rc) die(rc, "create");
    if (fp64_degree >= 0) {
        rc = infoli_set_fp64_degree(net, fp64_degree);
        if (rc) die(rc, "FP64 degree");
    }
    rc = infoli_set_connectivity_csr(net, row_ptr, col, w);
    if (rc) die(rc, "connectivity");
    free(row_ptr); free(col); free(w);

    if (gcal_file) {
        double *g = (double *)malloc(sizeof(double) * (size_t)n);
        if (!g) { printf("Error: Couldn't allocate memory for channels\n"); exit(EXIT_FAILURE); }
        rc = infoli_read_values(gcal_file, n, g);
        if (rc) die(rc, "g_CaL file");
        rc = infoli_set_gcal(net, g);
        if (rc) die(rc, "g_CaL");
        free(g);
    }

    if (rand_init) {
        /* RAND_INIT (src/single_threaded/infoli.c:212-236): random g_CaL, then
         * every cell alone for a random odd number of steps */
        double *g = (double *)malloc(sizeof(double) * (size_t)n);
        int64_t *ws = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        if (!g || !ws) { printf("Error: Couldn't allocate memory for channels\n"); exit(EXIT_FAILURE); }
        rc = infoli_gen_rand_init(rand_seed, n, params.v[INFOLI_P_DELTA], g, ws);
        if (!rc) rc = infoli_set_gcal(net, g);
        if (!rc) rc = infoli_isolated_steps(net, ws);
        if (rc) die(rc, "random initialisation");
        free(g);
        free(ws);
    }
    if (ck_in) {
        rc = infoli_load_checkpoint(net, ck_in);
        if (rc) die(rc, "checkpoint");
    }

    int64_t total_steps = (int64_t)(sim_time / params.v[INFOLI_P_DELTA]);  /* OMP:478 */
    double pulse_amp = 6;
    int64_t pulse_t0 = 20000, pulse_t1 = 20500 - 1;                        /* OMP:482-485 */
    if (in_file) {
        int64_t lines = 0;
        rc = infoli_read_iapp_file(in_file, n, INT64_MAX, NULL, &lines);
        if (rc) die(rc, "stimulus file");
        double *tab = (double *)malloc(sizeof(double) * (size_t)(lines * n > 0 ? lines * n : 1));
        if (!tab) { printf("Error: Couldn't allocate memory for the stimulus\n"); exit(EXIT_FAILURE); }
        rc = infoli_read_iapp_file(in_file, n, lines, tab, &lines);
        if (rc) die(rc, "stimulus file");
        rc = infoli_set_iapp_table(net, 0, lines, tab);
        if (rc) die(rc, "stimulus");
        free(tab);
        total_steps = lines;          /* one step per line (ST:248) */
        pulse_amp = 0;
    }
    rc = infoli_set_iapp_pulse(net, NULL, pulse_amp, pulse_t0, pulse_t1);
    if (rc) die(rc, "stimulus");
    if (autotune) {
        infoli_exec e;
        rc = infoli_autotune(net, 200, &e);
        if (rc) die(rc, "autotune");
    }

    FILE *out = fopen(out_file, "w+");
    if (!out) { printf("Error: Couldn't create %s\n", out_file); exit(EXIT_FAILURE); }
    if (!gcal_file) {
        fputs("Execution Time for Simulation in ms:              \n", out);
        fputs("#simSteps Input(Iapp) Output(V_axon)", out);
        for (int64_t i = 0; i < n; i++) fprintf(out, "%lld ", (long long)i);
        fputs("\n", out);
    }

    /* run in chunks into two trace buffers; a formatter thread writes chunk
     * c's records (in step order, cells ascending) while the GPU runs chunk
     * c + 1, and is joined before the next chunk's formatting starts */
    const int64_t max_rec = (int64_t)((64u << 20) / (sizeof(double) * (size_t)n)) + 1;
    const int64_t chunk = max_rec * stride;
    double *trace[2];
    trace[0] = (double *)malloc(sizeof(double) * (size_t)max_rec * (size_t)n);
    trace[1] = (double *)malloc(sizeof(double) * (size_t)max_rec * (size_t)n);
    if (!trace[0] || !trace[1]) { printf("Error: Couldn't allocate the trace buffer\n"); exit(EXIT_FAILURE); }
    struct fmt_job jobs[2];
    pthread_t fmt_thread;
    int fmt_running = 0, buf = 0;
    int nonfinite = 0;
    double tic = now_ms();
    const int64_t start = infoli_step_count(net);   /* > 0 when resumed from a checkpoint */
    for (int64_t t = start; t < start + total_steps; t += chunk, buf ^= 1) {
        const int64_t k = start + total_steps - t < chunk ? start + total_steps - t : chunk;
        const int64_t nrec = infoli_trace_records(net, k, stride);
        rc = infoli_advance(net, k, stride, trace[buf]);
        if (rc == INFOLI_ERR_NONFINITE) nonfinite = 1;
        else if (rc) die(rc, "advance");
        if (fmt_running) {
            pthread_join(fmt_thread, NULL);
            fmt_running = 0;
            if (jobs[buf ^ 1].bad) { printf("Error: trace record count mismatch\n"); exit(EXIT_FAILURE); }
        }
        jobs[buf] = (struct fmt_job){out, trace[buf], t, k, stride, n, nrec, gcal_file != NULL, in_file != NULL,
                                     pulse_amp, pulse_t0, pulse_t1, 0};
        if (pthread_create(&fmt_thread, NULL, format_records, &jobs[buf]) == 0) {
            fmt_running = 1;
        } else {
            format_records(&jobs[buf]);   /* no thread: format here */
            if (jobs[buf].bad) { printf("Error: trace record count mismatch\n"); exit(EXIT_FAILURE); }
        }
    }
    if (fmt_running) {
        pthread_join(fmt_thread, NULL);
        if (jobs[buf ^ 1].bad) { printf("Error: trace record count mismatch\n"); exit(EXIT_FAILURE); }
    }
    double toc = now_ms();
    fclose(out);
    chmod(out_file, 0x01B6);
    printf("Execution Time for Simulation: %0.2f ms.\n", toc - tic);
    if (nonfinite) printf("Warning: a voltage became NaN or Inf during the simulation.\n");
    if (!gcal_file) {
        out = fopen(out_file, "rb+");
        if (out) {
            fprintf(out, "Execution Time for Simulation in ms: %0.2f\n", toc - tic);
            fclose(out);
        }
    }
    if (state_file) {
        rc = infoli_write_state(net, state_file);
        if (rc) die(rc, "state dump");
    }
    if (ck_out) {
        rc = infoli_save_checkpoint(net, ck_out);
        if (rc) die(rc, "checkpoint");
    }
    free(trace[0]);
    free(trace[1]);
    infoli_destroy(net);
    return 0;
}
