/*
 * infoli_gpu.h -- C ABI of the MI355X-native InfOli network update.
 *
 * One context holds one network (or a batch of independent networks laid out
 * block-diagonally) resident in HBM: structure-of-arrays cell state, the
 * receiver-major neighbour lists in the layout the chosen gather reads
 * (sliced ELL, per-block staged tables or receiver-block x sender-chunk
 * tiles), per-cell g_CaL and applied-current schedule.  infoli_advance()
 * runs the per-timestep update (gap-junction currents, then the 3-compartment
 * Hodgkin-Huxley update with forward-Euler voltages) on the GPU for any
 * number of steps and returns the axon-voltage trace.
 *
 * The reference has no library boundary: its step is inline in main()'s loop
 * (src/openmp/infoli.c:480-698) over globals (src/openmp/infoli.h:43-72) and
 * SoA arrays (src/openmp/infoli.c:204-225).  Each entry point below names the
 * reference code it replaces.
 *
 * Conventions (all entry points):
 *   - return 0 (INFOLI_OK) or a negative INFOLI_ERR_* code, never exit();
 *     infoli_last_error() gives a message for the calling thread;
 *   - the library copies every input array; the caller keeps ownership;
 *   - one host thread per context; contexts are not reentrant;
 *   - all device work of a context runs on one HIP stream it owns.
 */
#ifndef INFOLI_GPU_H_
#define INFOLI_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define INFOLI_OK               0
#define INFOLI_ERR_ARG         -1  /* bad argument (null, size, range)            */
#define INFOLI_ERR_HIP         -2  /* HIP runtime error                          */
#define INFOLI_ERR_NOMEM       -3  /* host or device allocation failed           */
#define INFOLI_ERR_STATE       -4  /* call out of order (e.g. no connectivity)   */
#define INFOLI_ERR_NONFINITE   -5  /* a voltage became NaN/Inf (watchdog)        */
#define INFOLI_ERR_IO          -6  /* file open/read/write failed                */
#define INFOLI_ERR_CONFIG      -7  /* malformed configuration file               */
#define INFOLI_ERR_BARRIER     -8  /* persistent-kernel grid barrier timed out   */
#define INFOLI_ERR_NODEVICE    -9  /* no usable gfx950 device                    */

/* ---- arithmetic of the step ------------------------------------------ */
/* FP64: the double-precision arithmetic of src/single_threaded/infoli.c
 *       (mod_prec = double, src/single_threaded/infoli.h:68).
 * FP32: float arithmetic (mod_prec = float, src/openmp/infoli.h:39-40),
 *       plain neighbour sums.  The compartments follow
 *       src/openmp/infoli.c:548-679 (libm-accurate expf, IEEE division, the
 *       reference's literals and association).  The gap-junction term of
 *       :512-518 is the same formula evaluated with the hardware exp2,
 *       exp(-V^2/100) = exp2(V^2 * (-log2(e)/100)), and one fused
 *       multiply-add for 0.8 e + 0.2, so traces differ from an expf
 *       evaluation by up to 2e-4 mV over 2,000 steps without spikes and
 *       5e-4 mV with them.
 * FP32_KAHAN: FP32 with each cell's neighbour sum compensated (TwoSum).
 * In every arithmetic the removable singularities of beta_r (V_dend = -8.5
 * mV, src/openmp/infoli.c:568) and alpha_x (V = -25 mV, :627, :656) take
 * their limits 0.1 and 1.3 where the reference computes 0/0 = NaN. */
#define INFOLI_FP64        0
#define INFOLI_FP32        1
#define INFOLI_FP32_KAHAN  2

/* ---- model parameters: the 42 values of bin/default.conf, in file order
 * (read by read_parameters, src/utils/utils.c:137-160, assigned at
 * src/openmp/infoli.c:240-292).  Indices 0-22 are constants, 23-41 the
 * initial cell state. */
#define INFOLI_NPARAMS 42
typedef struct infoli_params {
    double v[INFOLI_NPARAMS];   /* value as written in the file (strtod)     */
    float  f[INFOLI_NPARAMS];   /* value as the reference reads it (%f)      */
} infoli_params;

enum {
    INFOLI_P_DELTA = 0, INFOLI_P_CONDUCTANCE, INFOLI_P_C_M, INFOLI_P_G_NA_S,
    INFOLI_P_G_KDR_S, INFOLI_P_G_K_S, INFOLI_P_G_LS, INFOLI_P_G_K_CA,
    INFOLI_P_G_CAH, INFOLI_P_G_LD, INFOLI_P_G_H, INFOLI_P_G_NA_A,
    INFOLI_P_G_NA_R, INFOLI_P_G_K_A, INFOLI_P_G_LA, INFOLI_P_P1, INFOLI_P_P2,
    INFOLI_P_G_INT, INFOLI_P_V_NA, INFOLI_P_V_K, INFOLI_P_V_CA, INFOLI_P_V_H,
    INFOLI_P_V_L,
    INFOLI_P_STATE0 = 23        /* first initial-state value (V_dend)        */
};

/* ---- per-cell variables: order of bin/default.conf lines 24-42 -------- */
enum {
    INFOLI_V_DEND = 0, INFOLI_CALCIUM_R, INFOLI_POTASSIUM_S, INFOLI_HCURRENT_Q,
    INFOLI_CA2PLUS, INFOLI_I_CAH, INFOLI_G_CAL, INFOLI_V_SOMA, INFOLI_SODIUM_M,
    INFOLI_SODIUM_H, INFOLI_POTASSIUM_N, INFOLI_POTASSIUM_P,
    INFOLI_POTASSIUM_X_S, INFOLI_CALCIUM_K, INFOLI_CALCIUM_L, INFOLI_V_AXON,
    INFOLI_SODIUM_M_A, INFOLI_SODIUM_H_A, INFOLI_POTASSIUM_X_A,
    INFOLI_NVARS                /* 19 */
};

typedef struct infoli_ctx infoli_t;

/* ---- configuration ---------------------------------------------------- */

/* Built-in bin/default.conf values.  Replaces the file read when no -c is
 * given and the file next to the executable is absent. */
int infoli_params_default(infoli_params *out);

/* Parse a 42-line "NAME = value" file.  Names are ignored, order matters.
 * Replaces read_parameters (src/utils/utils.c:137-160); a line that does not
 * scan is INFOLI_ERR_CONFIG (the reference prints and exits, utils.c:149-153). */
int infoli_read_config(const char *path, infoli_params *out);

/* ---- context lifecycle ------------------------------------------------ */

/* Allocate a network of n_cells on HIP device `device`, every cell at the
 * initial state params->v[23..41] (src/openmp/infoli.c:267-292,
 * initState src/single_threaded/infoli.c:862-902), no neighbours, no applied
 * current.  Replaces the SoA allocation and initialisation at
 * src/openmp/infoli.c:204-292. */
int infoli_create(infoli_t **ctx, const infoli_params *params,
                  int64_t n_cells, int precision, int device);

void infoli_destroy(infoli_t *ctx);

/* Receiver-major neighbour lists: cell i receives from col[row_ptr[i] ..
 * row_ptr[i+1]) with conductances w[...] (summed in the given order).
 * Replaces cellCompParams (src/openmp/infoli.h:79-83) filled at
 * src/openmp/infoli.c:320-375, and the gather of
 * perform_communication_step (src/single_threaded/infoli.c:1012-1037). */
int infoli_set_connectivity_csr(infoli_t *ctx, const int64_t *row_ptr,
                                const int32_t *col, const double *w);

/* Declare that the network is n_nets independent networks: cells
 * [offsets[k], offsets[k+1]) form network k and no neighbour list crosses a
 * boundary (checked).  The networks then run as one block-diagonal
 * network through the same kernels, many networks per launch.  No
 * reference counterpart: the reference runs one network per process. */
int infoli_set_partition(infoli_t *ctx, int n_nets, const int64_t *offsets);

/* INFOLI_FP32_KAHAN only: cells with at least min_degree neighbours compute
 * their gap-junction terms in FP64 (from the FP32 voltages and
 * conductances, the term of src/single_threaded/infoli.c:548-553) and sum
 * them in FP64; the others keep FP32 terms with compensated FP32 sums.
 * 0 = no cell.  Default INFOLI_FP64_DEGREE_DEFAULT (0 for the other
 * precisions, which reject a non-zero value).  No reference counterpart:
 * the reference sums every cell in its one precision
 * (src/openmp/infoli.c:496-520).
 *
 * The default keeps compensated FP32 inside the 1e-3 mV budget (max |dV|
 * against FP64 after 100 ms) with a margin: measured 1.7e-4 mV at degree
 * 100 (100,000 cells) and 7.2e-4 mV at degree 1000 (10,000 cells), fitted
 * as proportional to degree^0.63, which reaches 1e-3 mV near degree 1700. */
#define INFOLI_FP64_DEGREE_DEFAULT 1280
int infoli_set_fp64_degree(infoli_t *ctx, int64_t min_degree);
int64_t infoli_get_fp64_degree(const infoli_t *ctx);

/* Per-cell g_CaL.  Replaces read_g_CaL_from_file (src/utils/utils.c:124-135)
 * under -DG_CAL_FROM_FILE (src/openmp/infoli.c:294-296). */
int infoli_set_gcal(infoli_t *ctx, const double *g_cal);

/* Applied dendritic current Iapp_i(t) = base_i + (in_pulse(i,t) ? amp : 0),
 * with in_pulse(i,t) = (mask == NULL || mask[i]) && t0 <= t < t1, t the
 * 0-based step index.  The reference hard-codes base 0 and a pulse of 6 on
 * steps [20000, 20499) to every cell (src/openmp/infoli.c:482-485, 501). */
int infoli_set_iapp(infoli_t *ctx, const double *base /* NULL = zeros */);
int infoli_set_iapp_pulse(infoli_t *ctx, const uint8_t *mask, double amp,
                          int64_t t0, int64_t t1);

/* Per-step, per-cell applied current for steps [t_first, t_first+n_steps):
 * table[s*n_cells + i] replaces base+pulse on those steps.  Replaces
 * ReadFileLine-driven input (src/utils/utils.c:70-122,
 * src/single_threaded/infoli.c:242-286).  n_steps = 0 clears the table. */
int infoli_set_iapp_table(infoli_t *ctx, int64_t t_first, int64_t n_steps,
                          const double *table);

/* Read or overwrite one per-cell variable (INFOLI_V_DEND..INFOLI_POTASSIUM_X_A)
 * for all cells.  get_state replaces printState's per-cell dump
 * (src/single_threaded/infoli.c:971-1005). */
int infoli_set_state(infoli_t *ctx, int var, const double *values);
int infoli_get_state(infoli_t *ctx, int var, double *values);

/* Write every cell's state in printState's layout
 * (src/single_threaded/infoli.c:971-1005): one line per cell, 16 values
 * "%.8lf " in the order V_soma, Sodium_m, Potassium_n, Potassium_x_s,
 * Calcium_k, Calcium_l, V_dend, Calcium_r, Potassium_s, Hcurrent_q,
 * Ca2Plus, I_CaH, V_axon, Sodium_m_a, Sodium_h_a, Potassium_x_a.  Replaces
 * the PRINTSTATE dump (src/single_threaded/infoli.c:343-347). */
int infoli_write_state(infoli_t *ctx, const char *path);

/* Checkpoint and resume: all 19 per-cell variables in FP64 and the step
 * count (so a resumed run continues the pulse window and the Iapp table
 * where it stopped), in a binary file.  Loading needs the same cell count;
 * it converts between precisions.  No reference counterpart: printState
 * drops Sodium_h, Potassium_p and g_CaL and rounds to 8 decimals. */
int infoli_save_checkpoint(infoli_t *ctx, const char *path);
int infoli_load_checkpoint(infoli_t *ctx, const char *path);

/* Advance every cell alone, with no gap-junction current and no applied
 * current, for its own number of steps (steps[i] >= 0): the closed-circuit
 * warm-up of RAND_INIT (src/single_threaded/infoli.c:219-236, CompDend with
 * randomness 1).  The step count does not change. */
int infoli_isolated_steps(infoli_t *ctx, const int64_t *steps);

/* RAND_INIT's randomness from the counter generator (stream 7):
 * g_cal[i] = 0.6 + 0.2*(k % 100)/100 (src/single_threaded/infoli.c:212-214)
 * and steps[i] = (u % ceil(100/delta)) | 1 (src/single_threaded/infoli.c:221-222),
 * k, u the cell's two 64-bit draws.  Host only. */
int infoli_gen_rand_init(uint64_t seed, int64_t n, double delta, double *g_cal, int64_t *steps);

/* ---- the step ----------------------------------------------------------- */

/* Advance n_steps timesteps.  After step t (0-based, counted from context
 * creation) with t % trace_stride == 0, V_axon of every cell is appended to
 * `trace` (row-major [record][cell]; NULL = no trace), exactly the records the
 * reference prints at src/openmp/infoli.c:682-696 with
 * print_granularity = trace_stride.  Replaces the step loop body
 * src/openmp/infoli.c:480-698: gap-junction loop :496-520 and compartment
 * loop :548-679 (FP64: CompDend/CompSoma/CompAxon,
 * src/single_threaded/infoli.c:371-858).  Returns INFOLI_ERR_NONFINITE if a
 * voltage is NaN/Inf at the end (state is still advanced). */
int infoli_advance(infoli_t *ctx, int64_t n_steps, int64_t trace_stride,
                   double *trace);

/* Build now everything a following infoli_advance(ctx, n_steps,
 * trace_stride, ...) needs -- the neighbour layout of the execution
 * strategy, the device trace ring and pinned staging (trace_stride 0: no
 * trace), the hipGraph it replays -- without running a step, so that the
 * advance's time is the steps' alone.  No reference counterpart: the
 * reference's setup before its timer (src/openmp/infoli.c:204-428). */
int infoli_prepare(infoli_t *ctx, int64_t n_steps, int64_t trace_stride);

/* Number of records infoli_advance(ctx, n_steps, stride, ...) will write. */
int64_t infoli_trace_records(const infoli_t *ctx, int64_t n_steps,
                             int64_t trace_stride);

/* Steps completed since creation. */
int64_t infoli_step_count(const infoli_t *ctx);

/* I_c of every cell from the current V_dend, without stepping (the value
 * the gap-junction loop src/openmp/infoli.c:496-520, IcNeighbors
 * src/single_threaded/infoli.c:544-556, would compute now). */
int infoli_gap_currents(infoli_t *ctx, double *i_c);

/* infoli_gap_currents without allocating anything: the currents pass
 * through a static device scratch array of INFOLI_NOALLOC_SCRATCH_BYTES per
 * GPU instead of a buffer hipMalloc'd per call, and are copied straight
 * into i_c (no pinned staging), so the call makes no hipMalloc, hipFree or
 * other allocating call.  Same values as infoli_gap_currents, bit for bit.
 * No reference counterpart: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Limit: n_cells * (4 or 8 bytes) <= 16 MiB, i.e. 4,194,304
 * cells in FP32, 2,097,152 in FP64 (else INFOLI_ERR_ARG).  The neighbour
 * lists of the current strategy must be built (every infoli_advance and
 * infoli_gap_currents builds them), else INFOLI_ERR_STATE.  All contexts on
 * one GPU share the scratch array: calls that overlap on one GPU race. */
#define INFOLI_NOALLOC_SCRATCH_BYTES (16 << 20)
int infoli_gap_currents_noalloc(infoli_t *ctx, double *i_c);

/* ---- execution strategy and timing -------------------------------------- */

#define INFOLI_EXEC_AUTO        0  /* autotuned / heuristic choice           */
#define INFOLI_EXEC_STEP        1  /* one fused kernel per step              */
#define INFOLI_EXEC_GRAPH       2  /* per-step kernels replayed as hipGraph  */
#define INFOLI_EXEC_PERSISTENT  3  /* steps_per_launch steps in one kernel:
                                      whole networks per workgroup with no
                                      barrier between workgroups (GATHER_NETS),
                                      or the fused gathers with a grid barrier */

/* How a step reads the neighbours' V_dend. */
#define INFOLI_GATHER_DIRECT    0  /* random loads from the V_dend array        */
#define INFOLI_GATHER_LDS       1  /* whole V_dend array staged in LDS (small n) */
#define INFOLI_GATHER_STAGED    2  /* per cell block: senders sorted, copied to
                                      LDS, then read in each row's order        */
#define INFOLI_GATHER_TILED     3  /* receiver block x sender chunk tiles: the
                                      chunk's V_dend in LDS, per-tile partial
                                      sums merged by the compartment kernel     */
#define INFOLI_GATHER_SWEEP     4  /* the same tiles, one 1024-thread workgroup
                                      per receiver block walking its chunks in
                                      order, the partial sums merged in LDS
                                      (same results as TILED at equal shapes)  */
#define INFOLI_GATHER_NETS      5  /* whole networks per workgroup (the pieces
                                      of a block-diagonal network that no
                                      junction joins, up to 2048 cells each in
                                      FP32 and 1024 in FP64): V_dend in LDS,
                                      the state in registers, steps_per_launch
                                      steps per launch with no barrier between
                                      workgroups; persistent mode only, one
                                      thread per cell (the results of the other
                                      gathers at one lane per cell)            */

typedef struct infoli_exec {
    int     mode;            /* INFOLI_EXEC_*                                 */
    int     lanes_per_cell;  /* 1,2,4,8,16: lanes sharing one neighbour list  */
    int     block;           /* threads per workgroup (256; tiled 256, 512 or
                                1024; sweep 1024)                            */
    int     steps_per_launch;/* graph length / persistent steps per launch    */
    int     gather;          /* INFOLI_GATHER_*                               */
    int     lds_kb;          /* staged / tiled / sweep: LDS budget of the
                                gather list or chunk, KiB (the sweep's R sums
                                take 4 or 8 bytes each beside it, 160 KiB in
                                all)                                         */
    int     tile_cells;      /* tiled / sweep: receivers per tile (multiple
                                of 64)                                       */
    int     ode_split;       /* tiled compartment update: 0 one thread per
                                cell; 1 one wavefront per compartment; 2 somata,
                                axons and dendrite gates beside the tiles, the
                                dendrite voltages after (same results).  With
                                the sweep: 0 each block's cells in the
                                sweeping kernel itself, 1 a kernel after it,
                                2 a kernel running beside it (a second graph
                                branch) and the dendrite voltages after both */
    int     fast_math;       /* tiled / sweep, FP32: hardware exp2 and reciprocal in
                                the compartment update (0 = libm expf and IEEE
                                division, the reference's arithmetic)        */
} infoli_exec;

int infoli_get_exec(const infoli_t *ctx, infoli_exec *out);
int infoli_set_exec(infoli_t *ctx, const infoli_exec *exec);

/* Time candidate strategies on `trial_steps` steps each (state restored
 * afterwards) and keep the fastest; writes the choice to *chosen. */
int infoli_autotune(infoli_t *ctx, int64_t trial_steps, infoli_exec *chosen);

/* Device time in ms of the last infoli_advance (HIP events on the context's
 * stream), and of its dominant kernel launches: total ms and launch count. */
int infoli_last_timing(const infoli_t *ctx, double *advance_ms,
                       double *kernel_ms, int64_t *kernel_launches);

/* How the last infoli_advance launched its steps: hipGraph replays of
 * steps_per_launch steps each, steps launched one by one (profiling, step
 * mode, or a remainder shorter than a graph), and graphs captured during
 * the call (0 when infoli_prepare or an earlier call captured them). */
int infoli_last_launches(const infoli_t *ctx, int64_t *graph_replays,
                         int64_t *direct_steps, int64_t *captures);

/* Bracket every step-kernel launch with HIP events (per-step launches
 * instead of graph replays while on) so that infoli_last_timing reports
 * kernel time apart from launch gaps. */
int infoli_set_profiling(infoli_t *ctx, int on);

/* Block until all work of the context is done. */
int infoli_synchronize(infoli_t *ctx);

/* ---- synthetic inputs (seeded, counter-based; same bits on every host) -- */

/* Symmetric G(n, p): each unordered pair {i,j}, i != j, is a junction with
 * probability p, drawn by geometric skipping (O(edges)).  Rows ascending.
 * w_ij = w_ji = w_lo + (w_hi - w_lo) * U(0,1] per junction.  Two calls:
 * first with col == NULL to get *nnz (row_ptr filled), then with arrays of
 * *nnz entries.  No reference counterpart (the reference draws rand() per
 * ordered pair, src/openmp/infoli.c:337-352; see infoli_gen_directed). */
int infoli_gen_symmetric(uint64_t seed, int64_t n, double p, double w_lo,
                         double w_hi, int64_t *row_ptr, int32_t *col,
                         double *w, int64_t *nnz);

/* The reference generator's semantics (src/openmp/infoli.c:337-375):
 * receiver i gets sender j != i independently with probability p, senders
 * ascending, conductance w0; seeded instead of srand(time(NULL)).  Same
 * two-call protocol. */
int infoli_gen_directed(uint64_t seed, int64_t n, double p, double w0,
                        int64_t *row_ptr, int32_t *col, double *w,
                        int64_t *nnz);

/* out[i] = lo + (hi-lo)*U(0,1], stream-keyed. */
int infoli_gen_uniform(uint64_t seed, int stream, int64_t n, double lo,
                       double hi, double *out);
/* out[i] = mean + sd*N(0,1) (Box-Muller), stream-keyed. */
int infoli_gen_normal(uint64_t seed, int stream, int64_t n, double mean,
                      double sd, double *out);
/* mask[i] = 1 with probability frac, stream-keyed. */
int infoli_gen_mask(uint64_t seed, int stream, int64_t n, double frac,
                    uint8_t *mask);

/* ---- file formats (src/regression_test, tools/connectivity, tools/iapp) -- */

/* Dense n x n conductance matrix, row = sender, column = receiver
 * (src/single_threaded/infoli.c:158-191); entries with |value| < threshold
 * are no junction (stable_version.c:204 uses 1e-7; ST uses == 0).
 * Two-call protocol as infoli_gen_symmetric. */
int infoli_read_conn_dense(const char *path, int64_t n, double threshold,
                           int64_t *row_ptr, int32_t *col, double *w,
                           int64_t *nnz);

/* "send rec cond" triplets (src/regression_test/stable_version.c:226-252),
 * one junction per line, kept in file order per receiver. */
int infoli_read_conn_triplets(const char *path, int64_t n, double threshold,
                              int64_t *row_ptr, int32_t *col, double *w,
                              int64_t *nnz);

/* Whitespace-separated n values (gcal_file.txt, src/utils/utils.c:124-135). */
int infoli_read_values(const char *path, int64_t n, double *out);

/* Per-step Iapp file: one line per step with >= n values
 * (src/utils/utils.c:70-122).  Reads up to max_steps lines into
 * table[s*n + i]; *n_steps gets the count.  A short line is INFOLI_ERR_IO. */
int infoli_read_iapp_file(const char *path, int64_t n, int64_t max_steps,
                          double *table, int64_t *n_steps);

/* ---- diagnostics ------------------------------------------------------- */

/* Host-only check of the receiver-block x sender-chunk layout that
 * INFOLI_GATHER_TILED / INFOLI_GATHER_SWEEP upload for this network and shape
 * (needs no device).  Builds the tables with the library's own builder and
 * checks them against the network: every sender id inside its tile's staged
 * chunk (padding slots too, since they are read), each row's junctions in
 * row order and each junction in exactly one tile, sender ranges ascending
 * per block, slice widths, receiver lists that are permutations ordered by
 * in-chunk degree, descending.  Then one
 * step's scratch traffic as the kernels index it: every partial-sum and
 * dendrite-sum slot the compartment kernel reads is written by the launch
 * before it, and with ode_split 2 the `nsa` soma/axon workgroups update every
 * soma and axon exactly once.  INFOLI_OK, or INFOLI_ERR_STATE with the first
 * violation in infoli_last_error().  No reference counterpart: it checks the
 * layout that replaces cellCompParams (src/openmp/infoli.h:79-83). */
int infoli_tiled_layout_check(int64_t n_cells, const int64_t *row_ptr, const int32_t *col,
                              const double *w, int precision, int gather,
                              int lanes_per_cell, int tile_cells, int lds_kb,
                              int ode_split, int block, int64_t nsa);

const char *infoli_last_error(void);
const char *infoli_strerror(int code);
/* Library version string, e.g. "infoli-mi355x 0.1 (gfx950)". */
const char *infoli_version(void);

#ifdef __cplusplus
}
#endif
#endif /* INFOLI_GPU_H_ */
