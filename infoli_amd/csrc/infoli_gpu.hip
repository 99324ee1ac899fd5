/*
 * infoli_gpu.hip -- C ABI runtime of the MI355X InfOli network step.
 *
 * Owns the device copies of one network: state planes, the neighbour lists
 * in the layout the chosen gather reads (sliced ELL, staged per-block
 * tables or 2D tiles; see infoli_kernels.h), stimulus and trace ring.
 * Chooses a launch strategy (heuristic or measured by infoli_autotune),
 * launches the step kernels one per step, as replayed hipGraphs or as a
 * persistent kernel with a grid barrier, and copies the trace back in
 * pieces that overlap the following steps.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "infoli_gpu.h"
#include "infoli_host.h"
#include "infoli_kernels.h"

using namespace infoli;

/* Device scratch of infoli_gap_currents_noalloc: one call's gap currents
 * (n values in the context's precision) on their way to the caller.  Each
 * GPU the library runs on holds its own copy until the process ends; all
 * contexts on a GPU share it, so calls that overlap on one GPU race.  Host
 * code never uses this name as an address: infoli_create resolves the
 * device address once per device (scratch_setup) and the entry point passes
 * that pointer on. */
static __device__ __align__(256) unsigned char scratch[INFOLI_NOALLOC_SCRATCH_BYTES];

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return infoli_fail(INFOLI_ERR_HIP, "%s failed: %s (%s:%d)", #x,                \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define RCHK(x)                          \
    do {                                 \
        int r_ = (x);                    \
        if (r_ != INFOLI_OK) return r_;  \
    } while (0)

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};   // the scratch array's address on each device

/* The scratch array's device address on `device`, which must be current;
 * resolved once per device. */
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return infoli_fail(INFOLI_ERR_NODEVICE, "device %d: at most %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return INFOLI_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        return infoli_fail(INFOLI_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                           hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return INFOLI_OK;
}

}  // namespace

namespace {

constexpr size_t kTraceRingBytes = size_t(1) << 30;   // device trace ring budget
constexpr int64_t kTracePieceSteps = 1000;             // steps per trace piece (copy overlaps the next)
constexpr size_t kVldsMaxBytes = 96 * 1024;           // V_dend staged in LDS up to this size
constexpr size_t kLdsBytes = 160 * 1024;              // LDS of a CU (gfx950)
constexpr int kDefaultLdsKb = 76;                     // staged gather: two workgroups per CU

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct GraphKey {
    int mode = -1, L = 0, gather = 0, G = 0, grid = 0;
    size_t args_len = 0;
    unsigned char args[sizeof(StepArgs<double>)];
};

}  // namespace

struct infoli_ctx {
    int device = 0;
    int prec = INFOLI_FP64;
    int64_t n = 0, npad = 0;
    size_t tsize = 8;
    infoli_params params;
    hipStream_t stream = nullptr;

    void *d_state = nullptr;
    void *d_vd[2] = {nullptr, nullptr};
    void *d_vs[2] = {nullptr, nullptr};
    void *d_va[2] = {nullptr, nullptr};
    int32_t *d_deg = nullptr;
    int64_t *d_slice_ptr = nullptr;
    int32_t *d_col = nullptr;
    void *d_w = nullptr;
    int sell_L = 0;                 // lanes per cell of the layout on the device
    int64_t sell_entries = 0;

    // staged-gather layout (built for one lanes-per-cell value and LDS budget)
    bool staged_valid = false;
    int staged_kb = 0;
    int64_t st_nblocks = 0, st_elds = 0;
    int st_sb = 0;
    int32_t *d_blk_cell0 = nullptr, *d_blk_ncell = nullptr, *d_lrs = nullptr;
    int64_t *d_blk_gbase = nullptr;
    uint32_t *d_gl32 = nullptr;
    uint64_t *d_gl64 = nullptr;

    // 2D-tiled layout (built for one lanes-per-cell, chunk size and block size)
    bool tiled_valid = false;
    int tiled_L = 0, tiled_kb = 0, tiled_R = 0;
    int64_t tl_ntiles = 0, tl_nrb = 0, tl_C = 0;
    int64_t tl_nsa = 1;     // soma/axon workgroups launched beside the tiles (ode_split 2)
    bool tl_allwc = false;  // every tile has a constant conductance
    TileDesc *d_tiles = nullptr;
    int32_t *d_rb_tile0 = nullptr;
    int32_t *d_rb_one = nullptr;   // 0, 1, ..., nrb: one partial per block when swept
    int64_t *d_tsp = nullptr;
    uint16_t *d_rlist = nullptr, *d_rdeg = nullptr, *d_tcol = nullptr;
    void *d_tw = nullptr, *d_part = nullptr;
    void *d_dpre = nullptr;   // ode_split 2: [2][npad] dendrite current sum before I_c, and I_h

    // whole networks per workgroup (GATHER_NETS): cell ranges [rng[g], rng[g+1])
    bool nets_valid = false;
    NetsRange *d_rng = nullptr;
    unsigned char *d_nets_lists = nullptr;   // the ranges' neighbour lists kept in LDS
    int64_t nets_nr = 0;      // ranges = workgroups
    int nets_block = 0;       // threads per workgroup
    int nets_cpt = 1;         // cells per thread
    bool nets_split = false;  // one thread per compartment (k_persist_nets3)
    size_t nets_lds = 0;      // dynamic LDS per workgroup
    int64_t nets_in_lds = 0;  // ranges whose lists are in LDS

    void *d_iapp_base = nullptr;
    uint8_t *d_mask = nullptr;
    double amp = 0;
    int64_t t0 = 0, t1 = 0;
    void *d_tab = nullptr;
    int64_t tab_t0 = 0, tab_n = 0;

    void *d_trace = nullptr;
    int64_t trace_cap = 0;
    void *h_stage = nullptr;        // pinned staging for trace / state copies
    size_t h_stage_bytes = 0;

    int64_t *d_step = nullptr;
    AdvanceStatus *d_status = nullptr;   // barrier + watchdog flag of an advance
    GridBarrier *d_bar = nullptr;        // &d_status->bar
    int *d_flag = nullptr;               // &d_status->nonfinite
    AdvanceStatus *h_status = nullptr;   // pinned copy read after each advance
    int64_t step = 0;
    // the device step counter holds `step` and the status is clear, as the
    // last advance left them: the next advance needs no k_begin
    bool begin_done = false;

    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<double> w;
    int max_deg = 0;
    int64_t fp64_deg = 0;   // FP32 compensated: FP64 sums for cells of at least this degree
    std::vector<int64_t> part;

    infoli_exec exec{};
    int user_mode = -1;   // launch mode the caller set with infoli_set_exec (-1: none)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // swept tiles with ode_split 2: the compartment kernel's branch of a step
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_copy[2] = {nullptr, nullptr};   // trace pieces copied into staging halves
    std::vector<hipEvent_t> prof_ev;
    bool profiling = false;
    double last_ms = 0, kern_ms = 0;
    int64_t kern_launches = 0;
    // launches of the last infoli_advance: graph replays, steps launched
    // one by one, graph captures
    int64_t n_replays = 0, n_direct = 0, n_captures = 0;

    int num_cus = 0;
};

/* ------------------------------------------------------------------------ */
/* kernel dispatch                                                           */
/* ------------------------------------------------------------------------ */

namespace {

template <typename T>
Model<T> make_model(const infoli_params &p);

template <>
Model<float> make_model<float>(const infoli_params &p)
{
    const float *f = p.f;
    Model<float> m;
    m.DELTA = f[0]; m.C_M = f[2]; m.G_NA_S = f[3]; m.G_KDR_S = f[4]; m.G_K_S = f[5];
    m.G_LS = f[6]; m.G_K_CA = f[7]; m.G_CAH = f[8]; m.G_LD = f[9]; m.G_H = f[10];
    m.G_NA_A = f[11]; m.G_K_A = f[13]; m.G_LA = f[14];
    m.V_NA = f[18]; m.V_K = f[19]; m.V_CA = f[20]; m.V_H = f[21]; m.V_L = f[22];
    volatile float G_INT = f[17], P1 = f[15], P2 = f[16];   // float arithmetic, as OMP:587-666
    m.G_sd = G_INT / (1 - P1);
    m.G_ds = G_INT / P1;
    m.G_as = G_INT / (1 - P2);
    m.G_sa = G_INT / P2;
    return m;
}

template <>
Model<double> make_model<double>(const infoli_params &p)
{
    const double *v = p.v;
    Model<double> m;
    m.DELTA = v[0]; m.C_M = v[2]; m.G_NA_S = v[3]; m.G_KDR_S = v[4]; m.G_K_S = v[5];
    m.G_LS = v[6]; m.G_K_CA = v[7]; m.G_CAH = v[8]; m.G_LD = v[9]; m.G_H = v[10];
    m.G_NA_A = v[11]; m.G_K_A = v[13]; m.G_LA = v[14];
    m.V_NA = v[18]; m.V_K = v[19]; m.V_CA = v[20]; m.V_H = v[21]; m.V_L = v[22];
    m.G_sd = v[17] / (1 - v[15]);
    m.G_ds = v[17] / v[15];
    m.G_as = v[17] / (1 - v[16]);
    m.G_sa = v[17] / v[16];
    return m;
}

/* The sliced-ELL neighbour lists (sender id + conductance per entry) are
 * larger than half of the 256 MiB Infinity Cache: they are streamed from HBM
 * every step anyway, so their loads are marked nontemporal and leave the
 * caches to the gathered voltages. */
bool stream_exceeds_cache(const infoli_ctx *c)
{
    return (double)c->col.size() * (double)(4 + c->tsize) > 128.0 * 1024 * 1024;
}

/* the gathers built on the 2D-tiled layout */
bool tiled_layout(const infoli_exec &x) { return x.gather == GATHER_TILED || x.gather == GATHER_SWEEP; }

/* some cell sums FP64 gap-junction terms (FP32 compensated, degree budget) */
bool hyb_active(const infoli_ctx *c)
{
    return c->prec == INFOLI_FP32_KAHAN && c->fp64_deg > 0 && c->max_deg >= c->fp64_deg;
}

template <typename T>
StepArgs<T> make_args(const infoli_ctx *c, int64_t stride)
{
    StepArgs<T> a;
    memset(&a, 0, sizeof a);
    a.n = c->n;
    a.npad = c->npad;
    a.st = (T *)c->d_state;
    a.vd[0] = (T *)c->d_vd[0];
    a.vd[1] = (T *)c->d_vd[1];
    a.vs[0] = (T *)c->d_vs[0];
    a.vs[1] = (T *)c->d_vs[1];
    a.va[0] = (T *)c->d_va[0];
    a.va[1] = (T *)c->d_va[1];
    a.deg = c->d_deg;
    a.stream_nt = stream_exceeds_cache(c) ? 1 : 0;
    a.slice_ptr = c->d_slice_ptr;
    a.col = c->d_col;
    a.w = (const T *)c->d_w;
    a.iapp_base = (const T *)c->d_iapp_base;
    a.mask = c->d_mask;
    a.amp = (T)c->amp;
    a.t0 = c->t0;
    a.t1 = c->t1;
    a.tab = (const T *)c->d_tab;
    a.tab_t0 = c->tab_t0;
    a.tab_n = c->tab_n;
    a.trace = stride > 0 ? (T *)c->d_trace : nullptr;
    a.stride = stride > 0 ? stride : 1;
    a.trace_cap = c->trace_cap > 0 ? c->trace_cap : 1;
    a.d_step = nullptr;
    if (c->exec.gather == GATHER_STAGED) {
        a.nblocks = c->st_nblocks;
        a.blk_cell0 = c->d_blk_cell0;
        a.blk_ncell = c->d_blk_ncell;
        a.blk_gbase = c->d_blk_gbase;
        a.lrs = c->d_lrs;
        a.gl32 = c->d_gl32;
        a.gl64 = c->d_gl64;
        a.sb = c->st_sb;
    } else {
        a.nblocks = c->npad / kBlock;
    }
    if (tiled_layout(c->exec)) {
        a.tiles = c->d_tiles;
        a.rb_tile0 = c->d_rb_tile0;
        a.rb_part0 = c->exec.gather == GATHER_SWEEP ? c->d_rb_one : c->d_rb_tile0;
        a.tsp = c->d_tsp;
        a.rlist = c->d_rlist;
        a.rdeg = c->d_rdeg;
        a.tcol = c->d_tcol;
        a.tw = (const T *)c->d_tw;
        a.part = c->d_part;
        a.dpre = (T *)c->d_dpre;
        a.ntiles = (int32_t)c->tl_ntiles;
        a.nsa = (int32_t)c->tl_nsa;
        a.fast_math = c->exec.fast_math ? 1 : 0;
        a.sweep_ode = c->exec.gather == GATHER_SWEEP && c->exec.ode_split == 0 ? 1 : 0;
        a.fp64_deg = hyb_active(c) ? (int32_t)std::min<int64_t>(c->fp64_deg, INT32_MAX) : 0;
        a.rcells = c->tiled_R;
    }
    a.m = make_model<T>(c->params);
    return a;
}

typedef void (*step_fn_f)(StepArgs<float>, int64_t);
typedef void (*step_fn_d)(StepArgs<double>, int64_t);
typedef void (*pers_fn_f)(StepArgs<float>, GridBarrier *, int);
typedef void (*pers_fn_d)(StepArgs<double>, GridBarrier *, int);
typedef void (*gap_fn_f)(StepArgs<float>, int64_t, float *);
typedef void (*gap_fn_d)(StepArgs<double>, int64_t, double *);

int lidx(int L)
{
    switch (L) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    default: return -1;
    }
}

#define INFOLI_L_TABLE(KERN, T, ...) \
    {(const void *)KERN<T, 1, __VA_ARGS__>, (const void *)KERN<T, 2, __VA_ARGS__>, \
     (const void *)KERN<T, 4, __VA_ARGS__>, (const void *)KERN<T, 8, __VA_ARGS__>, \
     (const void *)KERN<T, 16, __VA_ARGS__>}

/* [precision][gather][L] */
#define INFOLI_G_TABLE(KERN, T, K) \
    {INFOLI_L_TABLE(KERN, T, K, GATHER_DIRECT), INFOLI_L_TABLE(KERN, T, K, GATHER_VLDS), \
     INFOLI_L_TABLE(KERN, T, K, GATHER_STAGED)}
const void *const k_step_tab[3][3][5] = {
    INFOLI_G_TABLE(k_step, double, false),
    INFOLI_G_TABLE(k_step, float, false),
    INFOLI_G_TABLE(k_step, float, true),
};
const void *const k_persist_tab[3][3][5] = {
    INFOLI_G_TABLE(k_persist, double, false),
    INFOLI_G_TABLE(k_persist, float, false),
    INFOLI_G_TABLE(k_persist, float, true),
};
/* [precision]: one thread per compartment */
const void *const k_persist_nets3_tab[3] = {(const void *)k_persist_nets3<double, false>,
                                            (const void *)k_persist_nets3<float, false>,
                                            (const void *)k_persist_nets3<float, true>};
/* [precision][cells per thread - 1] (FP64: one cell per thread) */
const void *const k_persist_nets_tab[3][2] = {
    {(const void *)k_persist_nets<double, false, 1>, (const void *)k_persist_nets<double, false, 1>},
    {(const void *)k_persist_nets<float, false, 1>, (const void *)k_persist_nets<float, false, 2>},
    {(const void *)k_persist_nets<float, true, 1>, (const void *)k_persist_nets<float, true, 2>},
};
/* [precision][threads: 256, 512, 1024][L] */
#define INFOLI_NT_TABLE(T, K) \
    {INFOLI_L_TABLE(k_tile_gj, T, K, 256), INFOLI_L_TABLE(k_tile_gj, T, K, 512), \
     INFOLI_L_TABLE(k_tile_gj, T, K, 1024)}
const void *const k_tile_gj_tab[3][3][5] = {
    INFOLI_NT_TABLE(double, false),
    INFOLI_NT_TABLE(float, false),
    INFOLI_NT_TABLE(float, true),
};
int ntidx(int nt) { return nt == 1024 ? 2 : nt == 512 ? 1 : 0; }
/* [fast math][precision] */
const void *const k_tile_ode_tab[2][3] = {
    {(const void *)k_tile_ode<double, false, false>, (const void *)k_tile_ode<float, false, false>,
     (const void *)k_tile_ode<float, true, false>},
    {(const void *)k_tile_ode<double, false, false>, (const void *)k_tile_ode<float, false, true>,
     (const void *)k_tile_ode<float, true, true>},
};
const void *const k_tile_ode3_tab[2][3] = {
    {(const void *)k_tile_ode3<double, false, false>, (const void *)k_tile_ode3<float, false, false>,
     (const void *)k_tile_ode3<float, true, false>},
    {(const void *)k_tile_ode3<double, false, false>, (const void *)k_tile_ode3<float, false, true>,
     (const void *)k_tile_ode3<float, true, true>},
};
/* FP32 compensated with FP64 sums for high-degree cells: [threads][L] */
/* every tile of constant conductance, FP32, 1024 threads: [compensated][L] */
const void *const k_tile_gj_awc_tab[2][5] = {
    INFOLI_L_TABLE(k_tile_gj, float, false, 1024, false, true),
    INFOLI_L_TABLE(k_tile_gj, float, true, 1024, false, true),
};
const void *const k_tile_gj_hyb_tab[3][5] = {
    INFOLI_L_TABLE(k_tile_gj, float, true, 256, true),
    INFOLI_L_TABLE(k_tile_gj, float, true, 512, true),
    INFOLI_L_TABLE(k_tile_gj, float, true, 1024, true),
};
const void *const k_gap_hyb_tab[5] = INFOLI_L_TABLE(k_gap_currents, float, true, true);
/* one workgroup of 1024 threads per receiver block: [precision][L], and
 * FP32 compensated with FP64 sums for high-degree cells: [L] */
const void *const k_tile_sweep_tab[3][5] = {
    INFOLI_L_TABLE(k_tile_sweep, double, false, 1024),
    INFOLI_L_TABLE(k_tile_sweep, float, false, 1024),
    INFOLI_L_TABLE(k_tile_sweep, float, true, 1024),
};
const void *const k_tile_sweep_hyb_tab[5] = INFOLI_L_TABLE(k_tile_sweep, float, true, 1024, true);
/* blocks of fewer slices than wavefronts: the compartments beside the rows */
const void *const k_tile_sweep_small_tab[3][5] = {
    INFOLI_L_TABLE(k_tile_sweep, double, false, 1024, false, true),
    INFOLI_L_TABLE(k_tile_sweep, float, false, 1024, false, true),
    INFOLI_L_TABLE(k_tile_sweep, float, true, 1024, false, true),
};
const void *const k_tile_sweep_small_hyb_tab[5] = INFOLI_L_TABLE(k_tile_sweep, float, true, 1024, true, true);
/* [fast math][precision] */
const void *const k_tile_dend_tab[2][3] = {
    {(const void *)k_tile_dend<double, false, false>, (const void *)k_tile_dend<float, false, false>,
     (const void *)k_tile_dend<float, true, false>},
    {(const void *)k_tile_dend<double, false, false>, (const void *)k_tile_dend<float, false, true>,
     (const void *)k_tile_dend<float, true, true>},
};
/* beside the swept tiles: [fast math][precision] */
const void *const k_tile_beside_tab[2][3] = {
    {(const void *)k_tile_beside<double, false>, (const void *)k_tile_beside<float, false>,
     (const void *)k_tile_beside<float, false>},
    {(const void *)k_tile_beside<double, false>, (const void *)k_tile_beside<float, true>,
     (const void *)k_tile_beside<float, true>},
};
const void *const k_gap_tab[3][5] = {
    INFOLI_L_TABLE(k_gap_currents, double, false),
    INFOLI_L_TABLE(k_gap_currents, float, false),
    INFOLI_L_TABLE(k_gap_currents, float, true),
};

bool is_f32(const infoli_ctx *c) { return c->prec != INFOLI_FP64; }

/* bytes of one partial (sum, compensation) pair */
size_t part_bytes(const infoli_ctx *c) { return c->prec == INFOLI_FP32_KAHAN ? 8 : c->tsize; }

/* the sweep's LDS: one accumulator per receiver of the block, then the chunk */
size_t sweep_lds_bytes_of(int R, int64_t C, size_t tsize, size_t psz) { return (size_t)R * psz + (size_t)C * tsize; }

size_t sweep_lds_bytes(const infoli_ctx *c, int R, int64_t C)
{
    return sweep_lds_bytes_of(R, C, c->tsize, part_bytes(c));
}

/* the largest chunk budget (KiB) that fits beside R receivers' sums */
int sweep_max_kb(const infoli_ctx *c, int R)
{
    const int64_t room = (int64_t)kLdsBytes - (int64_t)R * (int64_t)part_bytes(c);
    return (int)std::min<int64_t>(156, room / 1024);
}

size_t dyn_lds_bytes(const infoli_ctx *c)
{
    switch (c->exec.gather) {
    case GATHER_VLDS: return (size_t)c->n * c->tsize;
    case GATHER_STAGED:   // + slack: padded slots of a row's last slice read past its end
        return (size_t)(std::max<int64_t>(c->st_elds, 1) + c->max_deg + 64) * c->tsize;
    default: return 0;
    }
}

/* the tile kernel of the context's strategy and layout: the one whose rows
 * prefetch their next ids when every tile has a constant conductance and
 * the tiles run alone (measured: dense_10k 13.5 -> 12.4 us per step with
 * the compartments after the tiles; batched_512 83.4 -> 90.8 us with the
 * somata and axons beside them, so not then) */
const void *tile_kernel(const infoli_ctx *c)
{
    const int li = lidx(c->exec.lanes_per_cell);
    if (hyb_active(c)) return k_tile_gj_hyb_tab[ntidx(c->exec.block)][li];
    if (c->tl_allwc && c->exec.ode_split != 2 && is_f32(c) && c->exec.block == 1024)
        return k_tile_gj_awc_tab[c->prec == INFOLI_FP32_KAHAN ? 1 : 0][li];
    return k_tile_gj_tab[c->prec][ntidx(c->exec.block)][li];
}

/* the sweeping kernel of the context's precision and lanes */
const void *sweep_kernel(const infoli_ctx *c)
{
    const int li = lidx(c->exec.lanes_per_cell);
    // a block of fewer slices than the workgroup's 16 wavefronts, with the
    // compartments in the sweeping kernel: the idle wavefronts run the
    // somata, axons and dendrite gates beside the rows
    if (c->exec.ode_split == 0 && c->exec.tile_cells / (64 / c->exec.lanes_per_cell) < 16)
        return hyb_active(c) ? k_tile_sweep_small_hyb_tab[li] : k_tile_sweep_small_tab[c->prec][li];
    return hyb_active(c) ? k_tile_sweep_hyb_tab[li] : k_tile_sweep_tab[c->prec][li];
}

/* raise the dynamic-LDS limit of a kernel above 64 KiB when needed */
void allow_dyn_lds(const void *fn, size_t bytes)
{
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int64_t n_blocks(const infoli_ctx *c)
{
    return c->exec.gather == GATHER_STAGED ? c->st_nblocks : c->npad / kBlock;
}

int launch_step(infoli_ctx *c, const void *args, int64_t k)
{
    if (tiled_layout(c->exec)) {
        const int fm = c->exec.fast_math ? 1 : 0;
        const void *fo = k_tile_ode_tab[fm][c->prec];
        int64_t kk = k;
        void *kargs[] = {const_cast<void *>(args), &kk};
        if (c->exec.gather == GATHER_SWEEP) {
            // one workgroup per receiver block; the compartments in the same
            // kernel (ode_split 0), after it (1), or (2) the somata, axons
            // and dendrite gates in a kernel beside it and the dendrite
            // voltages after both
            const void *fs = sweep_kernel(c);
            const size_t lds = sweep_lds_bytes(c, c->tiled_R, c->tl_C);
            allow_dyn_lds(fs, lds);
            if (c->exec.ode_split == 2) {
                // fork: the sweep is launched first, so that its workgroups
                // take the CUs before the smaller ones fill them (the other
                // order measured 282 against 264 us per step at 1M cells);
                // join before the dendrite voltages
                HIPCHK(hipEventRecord(c->ev_fork, c->stream));
                HIPCHK(hipLaunchKernel(fs, dim3((unsigned)c->tl_nrb), dim3(1024), kargs, lds, c->stream));
                HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
                HIPCHK(hipLaunchKernel(k_tile_beside_tab[fm][c->prec], dim3((unsigned)(c->npad / 64)), dim3(192),
                                       kargs, 0, c->side));
                HIPCHK(hipEventRecord(c->ev_join, c->side));
                HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
                HIPCHK(hipLaunchKernel(k_tile_dend_tab[fm][c->prec], dim3((unsigned)(c->npad / kBlock)), dim3(kBlock),
                                       kargs, 0, c->stream));
                return INFOLI_OK;
            }
            HIPCHK(hipLaunchKernel(fs, dim3((unsigned)c->tl_nrb), dim3(1024), kargs, lds, c->stream));
            if (c->exec.ode_split == 0) return INFOLI_OK;   // the compartments ran in the sweep
        } else {
            const void *fg = tile_kernel(c);
            const size_t lds = (size_t)c->tl_C * c->tsize;
            allow_dyn_lds(fg, lds);
            if (c->exec.ode_split == 2) {
                // tiles, then the soma/axon workgroups in the same launch;
                // the dendrites (which need the tiles' partials) after
                HIPCHK(hipLaunchKernel(fg, dim3((unsigned)(c->tl_ntiles + c->tl_nsa)), dim3(c->exec.block), kargs,
                                       lds, c->stream));
                HIPCHK(hipLaunchKernel(k_tile_dend_tab[fm][c->prec], dim3((unsigned)(c->npad / kBlock)), dim3(kBlock),
                                       kargs, 0, c->stream));
                return INFOLI_OK;
            }
            HIPCHK(hipLaunchKernel(fg, dim3((unsigned)c->tl_ntiles), dim3(c->exec.block), kargs, lds, c->stream));
        }
        if (c->exec.ode_split)
            HIPCHK(hipLaunchKernel(k_tile_ode3_tab[fm][c->prec], dim3((unsigned)(c->npad / 64)), dim3(192), kargs, 0,
                                   c->stream));
        else
            HIPCHK(hipLaunchKernel(fo, dim3((unsigned)(c->npad / kBlock)), dim3(kBlock), kargs, 0, c->stream));
        return INFOLI_OK;
    }
    const void *fn = k_step_tab[c->prec][c->exec.gather][lidx(c->exec.lanes_per_cell)];
    const int64_t nblk = n_blocks(c);
    int64_t kk = k;
    void *kargs[] = {const_cast<void *>(args), &kk};
    allow_dyn_lds(fn, dyn_lds_bytes(c));
    HIPCHK(hipLaunchKernel(fn, dim3((unsigned)nblk), dim3(kBlock), kargs, dyn_lds_bytes(c), c->stream));
    return INFOLI_OK;
}

size_t nets_lds_bytes(const infoli_ctx *c) { return c->nets_lds; }

int persist_grid(infoli_ctx *c)
{
    if (c->exec.gather == GATHER_NETS) return (int)c->nets_nr;   // one workgroup per cell range
    const void *fn = k_persist_tab[c->prec][c->exec.gather][lidx(c->exec.lanes_per_cell)];
    int per_cu = 0;
    allow_dyn_lds(fn, dyn_lds_bytes(c));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, dyn_lds_bytes(c)));
    if (per_cu < 1) return infoli_fail(INFOLI_ERR_HIP, "persistent kernel does not fit on a CU");
    const int64_t nblk = n_blocks(c);
    int64_t g = (int64_t)per_cu * c->num_cus;
    return (int)std::min<int64_t>(g, nblk);
}

int launch_persist(infoli_ctx *c, const void *args, int nsteps, int grid)
{
    if (c->exec.gather == GATHER_NETS) {
        // whole networks per workgroup: no grid barrier, nothing to reset
        const void *fn = c->nets_split ? k_persist_nets3_tab[c->prec] : k_persist_nets_tab[c->prec][c->nets_cpt - 1];
        const NetsRange *rng = c->d_rng;
        const unsigned char *lists = c->d_nets_lists;
        int ns = nsteps;
        void *kargs[] = {const_cast<void *>(args), &rng, &lists, &ns};
        allow_dyn_lds(fn, nets_lds_bytes(c));
        HIPCHK(hipLaunchKernel(fn, dim3((unsigned)grid), dim3((unsigned)c->nets_block), kargs, nets_lds_bytes(c),
                               c->stream));
        return INFOLI_OK;
    }
    const void *fn = k_persist_tab[c->prec][c->exec.gather][lidx(c->exec.lanes_per_cell)];
    GridBarrier *gb = c->d_bar;
    int ns = nsteps;
    void *kargs[] = {const_cast<void *>(args), &gb, &ns};
    hipLaunchKernelGGL(k_reset_barrier, dim3(1), dim3(1), 0, c->stream, c->d_bar);
    HIPCHK(hipGetLastError());
    HIPCHK(hipLaunchKernel(fn, dim3((unsigned)grid), dim3(kBlock), kargs, dyn_lds_bytes(c), c->stream));
    return INFOLI_OK;
}


/* ------------------------------------------------------------------------ */
/* copies and fills, all on the context's stream                             */
/* ------------------------------------------------------------------------ */

/* Every copy and fill of a context is issued on its stream and waited for
 * there.  The stream is created non-blocking, so it is not ordered after the
 * null stream: a hipMemset or hipMemcpy on the null stream may still be in
 * flight (hipMemset of device memory returns before the fill completes) when
 * the stream's next kernel or upload runs. */
int h2d(infoli_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return INFOLI_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return INFOLI_OK;
}

int d2h(infoli_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return INFOLI_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return INFOLI_OK;
}

int zero_dev(infoli_ctx *c, void *dst, size_t bytes)
{
    if (bytes == 0) return INFOLI_OK;
    HIPCHK(hipMemsetAsync(dst, 0, bytes, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* sliced-ELL layout                                                         */
/* ------------------------------------------------------------------------ */

template <typename T>
int upload_sell(infoli_ctx *c, int L)
{
    const int CPW = 64 / L;
    const int64_t n = c->n, npad = c->npad;
    const int64_t nslices = npad / CPW;
    std::vector<int32_t> deg(npad, 0);
    for (int64_t i = 0; i < n; i++) deg[i] = (int32_t)(c->row_ptr[i + 1] - c->row_ptr[i]);
    std::vector<int64_t> sp(nslices + 1, 0);
    for (int64_t s = 0; s < nslices; s++) {
        int md = 0;
        for (int j = 0; j < CPW; j++) md = std::max(md, deg[s * CPW + j]);
        sp[s + 1] = sp[s] + (int64_t)((md + L - 1) / L) * 64;
    }
    const int64_t ne = sp[nslices];
    std::vector<int32_t> col(std::max<int64_t>(ne, 1), 0);
    std::vector<T> w(std::max<int64_t>(ne, 1), T(0));
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = i / CPW, cw = i % CPW, r0 = c->row_ptr[i];
        // padding slots (past the row's end, up to the slice width) name the
        // row's own cell with conductance 0: read, never added, and inside
        // the row's network (GATHER_NETS reads V_dend of that network only)
        const int64_t W = (sp[s + 1] - sp[s]) / 64;
        for (int p = 0; p < W * L; p++) {
            const int64_t e = sp[s] + (int64_t)(p / L) * 64 + cw * L + (p % L);
            col[e] = p < deg[i] ? c->col[r0 + p] : (int32_t)i;
            if (p < deg[i]) w[e] = (T)c->w[r0 + p];
        }
    }
    (void)hipFree(c->d_col); c->d_col = nullptr;
    (void)hipFree(c->d_w); c->d_w = nullptr;
    (void)hipFree(c->d_slice_ptr); c->d_slice_ptr = nullptr;
    HIPCHK(hipMalloc(&c->d_col, sizeof(int32_t) * col.size()));
    HIPCHK(hipMalloc(&c->d_w, sizeof(T) * w.size()));
    HIPCHK(hipMalloc(&c->d_slice_ptr, sizeof(int64_t) * sp.size()));
    RCHK(h2d(c, c->d_col, col.data(), sizeof(int32_t) * col.size()));
    RCHK(h2d(c, c->d_w, w.data(), sizeof(T) * w.size()));
    RCHK(h2d(c, c->d_slice_ptr, sp.data(), sizeof(int64_t) * sp.size()));
    RCHK(h2d(c, c->d_deg, deg.data(), sizeof(int32_t) * npad));
    c->sell_L = L;
    c->sell_entries = ne;
    return INFOLI_OK;
}

int bits_for(int64_t v)   // bits to hold values 0..v
{
    int b = 1;
    while (b < 63 && (v >> b) != 0) b++;
    return b;
}

void free_staged(infoli_ctx *c)
{
    void *bufs[] = {c->d_blk_cell0, c->d_blk_ncell, c->d_lrs, c->d_blk_gbase, c->d_gl32, c->d_gl64};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->d_blk_cell0 = c->d_blk_ncell = c->d_lrs = nullptr;
    c->d_blk_gbase = nullptr;
    c->d_gl32 = nullptr;
    c->d_gl64 = nullptr;
    c->staged_valid = false;
    c->staged_kb = 0;
    c->st_nblocks = c->st_elds = 0;
}

/* Staged-gather layout: cell blocks of 64..256 cells (multiples of 64, so
 * sliced-ELL slices never straddle a block) whose neighbour entries fit
 * the LDS budget; per block the list of (sender, LDS slot) sorted by
 * sender, slot = the entry's position in the block's compact row-major
 * list; and per cell its LDS row start. */
void free_tiled(infoli_ctx *c)
{
    void *bufs[] = {c->d_tiles, c->d_rb_tile0, c->d_rb_one, c->d_tsp, c->d_rlist, c->d_rdeg, c->d_tcol, c->d_tw,
                    c->d_part, c->d_dpre};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->d_dpre = nullptr;
    c->d_tiles = nullptr;
    c->d_rb_tile0 = c->d_rb_one = nullptr;
    c->d_tsp = nullptr;
    c->d_rlist = c->d_rdeg = c->d_tcol = nullptr;
    c->d_tw = c->d_part = nullptr;
    c->tiled_valid = false;
    c->tl_ntiles = c->tl_nrb = c->tl_C = 0;
    c->tl_allwc = false;
}

/* chunk of senders per tile for an LDS budget: a multiple of 64 cells */
int64_t chunk_cells(int64_t n, size_t tsize, int kb)
{
    const int64_t cap = (int64_t)kb * 1024 / (int64_t)tsize / 64 * 64;
    const int64_t nch = (n + cap - 1) / cap;
    return std::min<int64_t>(cap, round_up((n + nch - 1) / nch, 64));
}

int64_t tiled_chunk(const infoli_ctx *c, int kb) { return chunk_cells(c->n, c->tsize, kb); }

/* receiver-major CSR of a network, as the library keeps it */
struct CsrView {
    int64_t n;
    const int64_t *row_ptr;
    const int32_t *col;
    const double *w;
};

/* the host tables of the 2D-tiled layout (uploaded as they are) */
template <typename T>
struct TiledTables {
    int64_t C = 0, nrb = 0;
    std::vector<TileDesc> tiles;
    std::vector<int32_t> rbt0;
    std::vector<int64_t> tsp;
    std::vector<uint16_t> rlist, rdeg, tcol;
    std::vector<T> tw;
};

/* 2D-tiled layout: receiver blocks of R cells x sender chunks of C cells.
 * A tile exists for every (block, chunk) pair with a junction.  In a tile
 * the block's R receivers are listed by in-chunk degree, descending, and
 * their in-chunk junctions (in row order) are stored sliced-ELL with
 * 16-bit chunk-local sender ids.  Host only. */
template <typename T>
int build_tiled(const CsrView &g, size_t tsize, int L, int R, int kb, TiledTables<T> &tb)
{
    const int CPW = 64 / L;
    const int64_t n = g.n;
    const int64_t C = chunk_cells(n, tsize, kb);
    const int64_t nrb = (n + R - 1) / R;
    std::vector<TileDesc> &tiles = tb.tiles;
    std::vector<int32_t> &rbt0 = tb.rbt0;
    std::vector<int64_t> &tsp = tb.tsp;
    std::vector<uint16_t> &rlist = tb.rlist, &rdeg = tb.rdeg, &tcol = tb.tcol;
    std::vector<T> &tw = tb.tw;
    tiles.clear();
    rbt0.assign(nrb + 1, 0);
    tsp.assign(1, 0);
    rlist.clear();
    rdeg.clear();
    tcol.clear();
    tw.clear();
    for (int64_t rb = 0; rb < nrb; rb++) {
        rbt0[rb] = (int32_t)tiles.size();
        const int64_t c0 = rb * R, c1 = std::min<int64_t>(n, c0 + R);
        // the block's senders: when they span at most C cells (block-diagonal
        // batches), one tile stages exactly that range; else the sender
        // chunks [k C, (k + 1) C)
        int64_t lo = n, hi = -1;
        for (int64_t e = g.row_ptr[c0]; e < g.row_ptr[c1]; e++) {
            lo = std::min<int64_t>(lo, g.col[e]);
            hi = std::max<int64_t>(hi, g.col[e]);
        }
        if (hi >= lo) lo = lo / 64 * 64;   // the LDS copy loads 16-byte vectors from the range start
        const bool compact = hi >= lo && hi - lo + 1 <= C;
        const int64_t nchunks = compact ? 1 : (n + C - 1) / C;
        const auto chunk_of = [&](int64_t j) { return compact ? 0 : j / C; };
        const auto chunk_s0 = [&](int64_t ch) { return compact ? lo : ch * C; };
        const auto chunk_cn = [&](int64_t ch) { return compact ? hi - lo + 1 : std::min<int64_t>(C, n - ch * C); };
        // per chunk: the block's junctions into it, receivers ascending, row order
        std::vector<std::vector<int64_t>> ent(nchunks);
        std::vector<std::vector<int32_t>> entr(nchunks);
        for (int64_t i = c0; i < c1; i++)
            for (int64_t e = g.row_ptr[i]; e < g.row_ptr[i + 1]; e++) {
                ent[chunk_of(g.col[e])].push_back(e);
                entr[chunk_of(g.col[e])].push_back((int32_t)(i - c0));
            }
        for (int64_t ch = 0; ch < nchunks; ch++) {
            if (ent[ch].empty()) continue;
            std::vector<int> deg(R, 0);
            std::vector<int64_t> first(R, -1);
            for (size_t q = 0; q < ent[ch].size(); q++) {
                const int r = entr[ch][q];
                if (first[r] < 0) first[r] = (int64_t)q;
                deg[r]++;
            }
            std::vector<int> rs(R);
            for (int r = 0; r < R; r++) rs[r] = r;
            std::stable_sort(rs.begin(), rs.end(), [&](int x, int y) { return deg[x] > deg[y]; });
            TileDesc td;
            td.rb = (int32_t)rb;
            td.ch = (int32_t)ch;
            td.nsl = R / CPW;
            td.s0 = chunk_s0(ch);
            td.cn = (int32_t)chunk_cn(ch);
            td.pad = 0;
            // constant conductance over the tile (in the precision of the step)
            const T w0 = (T)g.w[ent[ch][0]];
            bool cw_all = true;
            for (int64_t e : ent[ch])
                if ((T)g.w[e] != w0) { cw_all = false; break; }
            td.wconst = cw_all ? 1 : 0;
            td.wc = (double)w0;
            td.sl0 = (int64_t)tsp.size() - 1;
            td.rl0 = (int64_t)rlist.size();
            for (int r : rs) {
                rlist.push_back((uint16_t)r);
                if (deg[r] > 65535) return infoli_fail(INFOLI_ERR_ARG, "tiled: in-chunk degree above 65535");
                rdeg.push_back((uint16_t)deg[r]);
            }
            for (int sl = 0; sl < td.nsl; sl++) {
                int md = 0;
                for (int cw = 0; cw < CPW; cw++) md = std::max(md, deg[rs[sl * CPW + cw]]);
                const int64_t base = tsp.back();
                const int64_t width = (md + L - 1) / L;
                tsp.push_back(base + width * 64);
                tcol.resize(tsp.back(), 0);
                tw.resize(tsp.back(), T(0));
                for (int cw = 0; cw < CPW; cw++) {
                    const int r = rs[sl * CPW + cw];
                    if (deg[r] == 0) continue;
                    const int64_t f = first[r];
                    for (int p = 0; p < deg[r]; p++) {
                        const int64_t e = ent[ch][f + p];
                        const int64_t slot = base + (int64_t)(p / L) * 64 + cw * L + (p % L);
                        tcol[slot] = (uint16_t)(g.col[e] - td.s0);
                        tw[slot] = (T)g.w[e];
                    }
                }
            }
            tiles.push_back(td);
        }
    }
    rbt0[nrb] = (int32_t)tiles.size();
    if (tiles.empty()) {   // no junctions at all: one empty tile keeps the kernels well-formed
        TileDesc td{0, 0, 0, 0, 0, 0, 0.0, 0, 0, 0};
        tiles.push_back(td);
        std::fill(rbt0.begin(), rbt0.end(), 0);
    }
    tb.C = C;
    tb.nrb = nrb;
    return INFOLI_OK;
}

/* partial-sum pairs the layout needs: one per receiver per tile, or one per
 * receiver per block when swept */
size_t part_slots(size_t ntiles, int64_t nrb, int R) { return std::max<size_t>(ntiles, (size_t)nrb) * (size_t)R; }

template <typename T>
int upload_tiled(infoli_ctx *c, int L, int R, int kb)
{
    TiledTables<T> tb;
    RCHK(build_tiled<T>(CsrView{c->n, c->row_ptr.data(), c->col.data(), c->w.data()}, c->tsize, L, R, kb, tb));
    const std::vector<TileDesc> &tiles = tb.tiles;
    const int64_t nrb = tb.nrb;
    free_tiled(c);
    const size_t psz = part_bytes(c);
    std::vector<int32_t> rb_one(nrb + 1);
    for (int64_t rb = 0; rb <= nrb; rb++) rb_one[rb] = (int32_t)rb;
    HIPCHK(hipMalloc(&c->d_tiles, sizeof(TileDesc) * tiles.size()));
    HIPCHK(hipMalloc(&c->d_rb_tile0, sizeof(int32_t) * tb.rbt0.size()));
    HIPCHK(hipMalloc(&c->d_rb_one, sizeof(int32_t) * rb_one.size()));
    HIPCHK(hipMalloc(&c->d_tsp, sizeof(int64_t) * tb.tsp.size()));
    HIPCHK(hipMalloc(&c->d_rlist, sizeof(uint16_t) * std::max<size_t>(tb.rlist.size(), 1)));
    HIPCHK(hipMalloc(&c->d_rdeg, sizeof(uint16_t) * std::max<size_t>(tb.rdeg.size(), 1)));
    HIPCHK(hipMalloc(&c->d_tcol, sizeof(uint16_t) * std::max<size_t>(tb.tcol.size(), 1)));
    HIPCHK(hipMalloc(&c->d_tw, sizeof(T) * std::max<size_t>(tb.tw.size(), 1)));
    // partials: one per tile, or one per block when swept
    const size_t pbytes = psz * part_slots(tiles.size(), nrb, R);
    HIPCHK(hipMalloc(&c->d_part, pbytes));
    HIPCHK(hipMalloc(&c->d_dpre, 2 * (size_t)c->npad * c->tsize));
    // every slot a step reads is written earlier in that step (the layout
    // check infoli_tiled_layout_check verifies it); zeroed all the same, so
    // the scratch never carries a freed buffer's data
    RCHK(zero_dev(c, c->d_part, pbytes));
    RCHK(zero_dev(c, c->d_dpre, 2 * (size_t)c->npad * c->tsize));
    RCHK(h2d(c, c->d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size()));
    RCHK(h2d(c, c->d_rb_tile0, tb.rbt0.data(), sizeof(int32_t) * tb.rbt0.size()));
    RCHK(h2d(c, c->d_rb_one, rb_one.data(), sizeof(int32_t) * rb_one.size()));
    RCHK(h2d(c, c->d_tsp, tb.tsp.data(), sizeof(int64_t) * tb.tsp.size()));
    RCHK(h2d(c, c->d_rlist, tb.rlist.data(), sizeof(uint16_t) * tb.rlist.size()));
    RCHK(h2d(c, c->d_rdeg, tb.rdeg.data(), sizeof(uint16_t) * tb.rdeg.size()));
    RCHK(h2d(c, c->d_tcol, tb.tcol.data(), sizeof(uint16_t) * tb.tcol.size()));
    RCHK(h2d(c, c->d_tw, tb.tw.data(), sizeof(T) * tb.tw.size()));
    c->tl_ntiles = (int64_t)tiles.size();
    c->tl_allwc = std::all_of(tiles.begin(), tiles.end(), [](const TileDesc &t) { return t.wconst != 0; });
    c->tl_nrb = nrb;
    c->tl_C = tb.C;
    c->tiled_L = L;
    c->tiled_kb = kb;
    c->tiled_R = R;
    c->tiled_valid = true;
    return INFOLI_OK;
}

int upload_staged(infoli_ctx *c, int kb)
{
    const int64_t n = c->n, npad = c->npad;
    const int64_t emax = (int64_t)kb * 1024 / (int64_t)c->tsize;
    auto deg_of = [&](int64_t i) -> int64_t { return i < n ? c->row_ptr[i + 1] - c->row_ptr[i] : 0; };
    std::vector<int32_t> b0, bn;
    std::vector<int64_t> bent;
    for (int64_t g = 0; g < npad; g += 64) {
        int64_t ge = 0;
        for (int64_t i = g; i < g + 64; i++) ge += deg_of(i);
        if (ge > emax)
            return infoli_fail(INFOLI_ERR_ARG,
                               "staged gather: cells %lld..%lld have %lld neighbour entries, more than "
                               "the %d KiB LDS budget holds (%lld)",
                               (long long)g, (long long)(g + 63), (long long)ge, kb, (long long)emax);
        if (!b0.empty() && bn.back() + 64 <= kBlock && bent.back() + ge <= emax) {
            bn.back() += 64;
            bent.back() += ge;
        } else {
            b0.push_back((int32_t)g);
            bn.push_back(64);
            bent.push_back(ge);
        }
    }
    const int64_t nb = (int64_t)b0.size();
    int64_t elds = 1;
    for (int64_t e : bent) elds = std::max(elds, e);
    std::vector<int32_t> lrs(npad, 0);
    std::vector<int64_t> gbase(nb + 1, 0);
    for (int64_t b = 0; b < nb; b++) {
        int64_t run = 0;
        for (int64_t i = b0[b]; i < b0[b] + bn[b]; i++) {
            lrs[i] = (int32_t)run;
            run += deg_of(i);
        }
        gbase[b + 1] = gbase[b] + run;
    }
    const int sb = bits_for(std::max<int64_t>(elds - 1, 1));
    const int cb = bits_for(std::max<int64_t>(n - 1, 1));
    const bool packed = sb + cb <= 32;
    const int64_t ne = gbase[nb];
    std::vector<uint64_t> list(std::max<int64_t>(ne, 1), 0);
    for (int64_t b = 0; b < nb; b++) {
        uint64_t *lb = list.data() + gbase[b];
        int64_t k = 0;
        for (int64_t i = b0[b]; i < b0[b] + bn[b] && i < n; i++)
            for (int64_t e = c->row_ptr[i]; e < c->row_ptr[i + 1]; e++)
                lb[k++] = ((uint64_t)(uint32_t)c->col[e] << 32) | (uint64_t)(lrs[i] + (e - c->row_ptr[i]));
        std::sort(lb, lb + k);
    }
    free_staged(c);
    HIPCHK(hipMalloc(&c->d_blk_cell0, sizeof(int32_t) * nb));
    HIPCHK(hipMalloc(&c->d_blk_ncell, sizeof(int32_t) * nb));
    HIPCHK(hipMalloc(&c->d_blk_gbase, sizeof(int64_t) * (nb + 1)));
    HIPCHK(hipMalloc(&c->d_lrs, sizeof(int32_t) * npad));
    RCHK(h2d(c, c->d_blk_cell0, b0.data(), sizeof(int32_t) * nb));
    RCHK(h2d(c, c->d_blk_ncell, bn.data(), sizeof(int32_t) * nb));
    RCHK(h2d(c, c->d_blk_gbase, gbase.data(), sizeof(int64_t) * (nb + 1)));
    RCHK(h2d(c, c->d_lrs, lrs.data(), sizeof(int32_t) * npad));
    if (packed) {
        std::vector<uint32_t> l32(list.size());
        for (size_t k = 0; k < list.size(); k++)
            l32[k] = ((uint32_t)(list[k] >> 32) << sb) | (uint32_t)(list[k] & 0xFFFFFFFFu);
        HIPCHK(hipMalloc(&c->d_gl32, sizeof(uint32_t) * l32.size()));
        RCHK(h2d(c, c->d_gl32, l32.data(), sizeof(uint32_t) * l32.size()));
    } else {
        HIPCHK(hipMalloc(&c->d_gl64, sizeof(uint64_t) * list.size()));
        RCHK(h2d(c, c->d_gl64, list.data(), sizeof(uint64_t) * list.size()));
    }
    c->st_nblocks = nb;
    c->st_elds = elds;
    c->st_sb = sb;
    c->staged_valid = true;
    c->staged_kb = kb;
    return INFOLI_OK;
}

/* Soma/axon workgroups beside the tiles (ode_split 2): as many as the
 * workgroup slots the tiles leave free (so they hide under the tiles), but
 * at most four cell groups each; one per group when the tiles fill every
 * slot (they then take the tail of the last round). */
int64_t soma_axon_workgroups(const infoli_ctx *c)
{
    const int64_t per = c->exec.block / 2;
    const int64_t groups = std::max<int64_t>(1, (c->n + per - 1) / per);
    const void *fg = tile_kernel(c);
    const size_t lds = (size_t)c->tl_C * c->tsize;
    allow_dyn_lds(fg, lds);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fg, c->exec.block, lds) != hipSuccess) per_cu = 0;
    const int64_t free_slots = (int64_t)per_cu * c->num_cus - c->tl_ntiles;
    if (free_slots <= 0) return groups;
    return std::min(groups, std::max(free_slots, (groups + 3) / 4));
}

/* ---- whole networks per workgroup (GATHER_NETS) ---------------------- */

/* work of a piece of the network in pair-term units: its junctions plus
 * the compartment update of each cell, which costs about as much VALU work
 * as 24 junction terms (22 exp and 34 divisions against 1 exp and 13
 * arithmetic operations) */
constexpr double kOdeTerms = 24.0;
/* Step-time model of the choice between whole networks per workgroup and
 * the graph-launched strategies, fitted to FP32 compensated runs on the
 * MI355X (sweep_256: 256 networks of 256 cells, 7.6 us per step persistent
 * against 9.9 us tiled; 560 networks of 300-700 cells, 31 against 27 us):
 * a range's work costs 0.5 ns per term unit on its CU (one workgroup per
 * CU, its latency barely hidden), the graph's whole work 0.27 ns per unit
 * spread over the CUs plus 6.5 us per step for its two dependent kernels. */
constexpr double kNetsUsPerTerm = 5.0e-4;
constexpr double kGraphUsPerTerm = 2.7e-4;
constexpr double kGraphStepUs = 6.5;

/* The cuts b that no junction crosses (every row of a cell below b names
 * only senders below b, and every row at or above b only senders at or
 * above b), from the CSR alone: [cuts[k], cuts[k+1]) are the network's
 * independent pieces, the networks of a block-diagonal batch or finer.
 * Stops early, with an empty result, once a piece would exceed `cap`
 * cells (the common case of one large connected network). */
std::vector<int64_t> independent_cuts(const infoli_ctx *c, int64_t cap)
{
    const int64_t n = c->n;
    // forward: no row below b reaches b or beyond
    std::vector<int64_t> cand{0};
    int64_t hi = -1;
    for (int64_t b = 1; b <= n; b++) {
        const int64_t i = b - 1;
        hi = std::max(hi, i);
        for (int64_t e = c->row_ptr[i]; e < c->row_ptr[i + 1]; e++) hi = std::max<int64_t>(hi, c->col[e]);
        if (b == n || hi < b) {
            if (b - cand.back() > cap) return {};
            cand.push_back(b);
        } else if (b - cand.back() > cap) {
            return {};
        }
    }
    // backward: no row at or above b reaches below b (directed networks)
    std::vector<int64_t> sufmin(n + 1, n);
    for (int64_t i = n - 1; i >= 0; i--) {
        int64_t m = i;
        for (int64_t e = c->row_ptr[i]; e < c->row_ptr[i + 1]; e++) m = std::min<int64_t>(m, c->col[e]);
        sufmin[i] = std::min(sufmin[i + 1], m);
    }
    std::vector<int64_t> cuts{0};
    for (size_t k = 1; k + 1 < cand.size(); k++)
        if (sufmin[cand[k]] >= cand[k]) cuts.push_back(cand[k]);
    cuts.push_back(n);
    for (size_t k = 0; k + 1 < cuts.size(); k++)
        if (cuts[k + 1] - cuts[k] > cap) return {};
    return cuts;
}

struct NetsPlan {
    std::vector<int32_t> rng;   // range bounds, nr + 1
    int block = 0, cpt = 1;
    bool split = false;         // one thread per compartment (k_persist_nets3)
    double max_work = 0, total_work = 0;
};

/* the most cells one workgroup holds: 1024 threads x two cells (FP32) or
 * one (FP64: 19 state values of 8 bytes per cell stay in registers) */
int64_t nets_cap(const infoli_ctx *c) { return is_f32(c) ? 2048 : 1024; }

/* Consecutive pieces packed into at most max(CUs, what the cell cap
 * forces) ranges of at most nets_cap cells, minimising the largest range's
 * work (binary search on it).  False when a piece is larger than a
 * workgroup holds. */
bool plan_nets(const infoli_ctx *c, NetsPlan &pl)
{
    const int64_t cap = nets_cap(c);
    const std::vector<int64_t> cuts = independent_cuts(c, cap);
    if (cuts.size() < 2) return false;
    const size_t np = cuts.size() - 1;
    std::vector<double> work(np);
    std::vector<int64_t> cells(np);
    pl.total_work = 0;
    double wmax = 0;
    for (size_t k = 0; k < np; k++) {
        cells[k] = cuts[k + 1] - cuts[k];
        if (cells[k] > cap) return false;
        work[k] = kOdeTerms * (double)cells[k] + (double)(c->row_ptr[cuts[k + 1]] - c->row_ptr[cuts[k]]);
        pl.total_work += work[k];
        wmax = std::max(wmax, work[k]);
    }
    const auto count = [&](double T, std::vector<int32_t> *out) {
        int64_t nr = 0, rc = 0;
        double rw = 0;
        if (out) out->assign(1, 0);
        for (size_t k = 0; k < np; k++) {
            if (rc > 0 && (rc + cells[k] > cap || rw + work[k] > T)) {
                nr++;
                if (out) out->push_back((int32_t)cuts[k]);
                rc = 0;
                rw = 0;
            }
            rc += cells[k];
            rw += work[k];
        }
        nr++;
        if (out) out->push_back((int32_t)c->n);
        return nr;
    };
    const int64_t want = std::max<int64_t>(c->num_cus, count(pl.total_work, nullptr));
    double lo = wmax, hi = pl.total_work;
    for (int it = 0; it < 60 && hi - lo > 0.5; it++) {
        const double mid = 0.5 * (lo + hi);
        if (count(mid, nullptr) <= want) hi = mid;
        else lo = mid;
    }
    count(hi, &pl.rng);
    int64_t maxc = 0;
    pl.max_work = 0;
    for (size_t g = 0, k = 0; g + 1 < pl.rng.size(); g++) {
        maxc = std::max<int64_t>(maxc, pl.rng[g + 1] - pl.rng[g]);
        double w = 0;
        for (; k < np && cuts[k] < pl.rng[g + 1]; k++) w += work[k];
        pl.max_work = std::max(pl.max_work, w);
    }
    pl.cpt = maxc > 1024 ? 2 : 1;
    pl.block = (int)round_up((maxc + pl.cpt - 1) / pl.cpt, 64);
    // ranges of at most 320 cells: one thread per compartment (three groups
    // of the range's cells rounded up to 64, at most 960 threads)
    pl.split = maxc <= 320;
    if (pl.split) pl.block = 3 * (int)round_up(maxc, 64);
    return true;
}

/* The range table of a plan and, for every range whose lists fit beside
 * its voltages in LDS, the lists in the kernel's LDS form: slot p of local
 * cell i at p * stride + i for p below the range's widest slice, 16-bit
 * range-local sender ids (padding: the cell itself), then the conductances
 * unless they are all equal.  A cell's row uses the slots below the width
 * of its 64-cell slice of the one-lane sliced-ELL layout, as the direct
 * gather does. */
template <typename T>
int upload_nets_t(infoli_ctx *c, const NetsPlan &pl)
{
    const int64_t n = c->n, nr = (int64_t)pl.rng.size() - 1;
    // lists stride: the cells a workgroup can hold; LDS before the lists:
    // V_dend in two halves, or all three voltages when split
    const int stride = pl.split ? pl.block / 3 : pl.block * pl.cpt;
    const size_t vbytes = (pl.split ? 6 : 2) * (size_t)stride * sizeof(T);
    std::vector<int32_t> deg(n);
    for (int64_t i = 0; i < n; i++) deg[i] = (int32_t)(c->row_ptr[i + 1] - c->row_ptr[i]);
    std::vector<int32_t> swid((n + 63) / 64, 0);   // slice widths of the one-lane layout
    for (int64_t i = 0; i < n; i++) swid[i / 64] = std::max(swid[i / 64], deg[i]);
    std::vector<NetsRange> rg(nr);
    std::vector<unsigned char> blob;
    size_t lds_max = vbytes;
    int64_t in_lds = 0;
    for (int64_t g = 0; g < nr; g++) {
        NetsRange &R = rg[g];
        R.r0 = pl.rng[g];
        R.r1 = pl.rng[g + 1];
        R.lds = 0;
        R.wconst = 1;
        R.off = 0;
        R.wmax = 0;
        R.pad = 0;
        R.wc = 0.0;
        bool first = true;
        for (int64_t i = R.r0; i < R.r1; i++) {
            R.wmax = std::max(R.wmax, swid[i / 64]);
            for (int64_t e = c->row_ptr[i]; e < c->row_ptr[i + 1]; e++) {
                const T wv = (T)c->w[e];
                if (first) {
                    R.wc = (double)wv;
                    first = false;
                } else if (wv != (T)R.wc) {
                    R.wconst = 0;
                }
            }
        }
        const size_t ib = ((size_t)R.wmax * stride * 2 + 15) & ~(size_t)15;
        const size_t need = vbytes + ib + (R.wconst ? 0 : (size_t)R.wmax * stride * sizeof(T));
        if (need > kLdsBytes) continue;   // the lists stay in HBM (sliced ELL)
        R.lds = 1;
        R.off = (int64_t)blob.size();
        lds_max = std::max(lds_max, need);
        in_lds++;
        blob.resize(blob.size() + (need - vbytes), 0);
        uint16_t *ids = reinterpret_cast<uint16_t *>(blob.data() + R.off);
        T *wl = reinterpret_cast<T *>(blob.data() + R.off + ib);
        for (int64_t i = R.r0; i < R.r1; i++) {
            const int64_t loc = i - R.r0, W = swid[i / 64];
            for (int64_t p = 0; p < R.wmax; p++) {
                const bool real = p < deg[i];
                ids[p * stride + loc] = (uint16_t)(real ? c->col[c->row_ptr[i] + p] - R.r0 : loc);
                if (!R.wconst && p < W) wl[p * stride + loc] = real ? (T)c->w[c->row_ptr[i] + p] : T(0);
            }
        }
    }
    (void)hipFree(c->d_rng);
    c->d_rng = nullptr;
    (void)hipFree(c->d_nets_lists);
    c->d_nets_lists = nullptr;
    HIPCHK(hipMalloc(&c->d_rng, sizeof(NetsRange) * rg.size()));
    RCHK(h2d(c, c->d_rng, rg.data(), sizeof(NetsRange) * rg.size()));
    HIPCHK(hipMalloc(&c->d_nets_lists, std::max<size_t>(blob.size(), 16)));
    RCHK(h2d(c, c->d_nets_lists, blob.data(), blob.size()));
    c->nets_nr = nr;
    c->nets_block = pl.block;
    c->nets_cpt = pl.cpt;
    c->nets_split = pl.split;
    c->nets_lds = lds_max;
    c->nets_in_lds = in_lds;
    c->nets_valid = true;
    return INFOLI_OK;
}

int upload_nets(infoli_ctx *c)
{
    NetsPlan pl;
    if (!plan_nets(c, pl))
        return infoli_fail(INFOLI_ERR_ARG, "gather NETS: a piece of the network that no junction separates has "
                                           "more than %lld cells, more than one workgroup holds",
                           (long long)nets_cap(c));
    if (pl.block * pl.cpt > 65536)
        return infoli_fail(INFOLI_ERR_ARG, "gather NETS: range of %d cells", pl.block * pl.cpt);
    return is_f32(c) ? upload_nets_t<float>(c, pl) : upload_nets_t<double>(c, pl);
}

int ensure_layout(infoli_ctx *c)
{
    const int L = c->exec.lanes_per_cell;
    if (c->sell_L != L) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        RCHK(is_f32(c) ? upload_sell<float>(c, L) : upload_sell<double>(c, L));
    }
    if (c->exec.gather == GATHER_STAGED && (!c->staged_valid || c->staged_kb != c->exec.lds_kb)) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        RCHK(upload_staged(c, c->exec.lds_kb));
    }
    if (tiled_layout(c->exec) &&
        (!c->tiled_valid || c->tiled_kb != c->exec.lds_kb || c->tiled_L != L || c->tiled_R != c->exec.tile_cells)) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        RCHK(is_f32(c) ? upload_tiled<float>(c, L, c->exec.tile_cells, c->exec.lds_kb)
                       : upload_tiled<double>(c, L, c->exec.tile_cells, c->exec.lds_kb));
    }
    if (c->exec.gather == GATHER_TILED) c->tl_nsa = soma_axon_workgroups(c);
    if (c->exec.gather == GATHER_NETS && !c->nets_valid) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        RCHK(upload_nets(c));
    }
    return INFOLI_OK;
}

void drop_graph(infoli_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gkey.mode = -1;
}

int choose_lanes(double mean_deg)
{
    if (mean_deg < 12) return 1;
    if (mean_deg < 40) return 2;
    if (mean_deg < 160) return 4;
    if (mean_deg < 640) return 8;
    return 16;
}

/* every aligned group of 64 cells fits the staged-gather LDS budget */
bool max_block_entries_fit(const infoli_ctx *c, int kb)
{
    const int64_t emax = (int64_t)kb * 1024 / (int64_t)c->tsize;
    for (int64_t g = 0; g < c->n; g += 64) {
        const int64_t hi = std::min<int64_t>(g + 64, c->n);
        if (c->row_ptr[hi] - c->row_ptr[g] > emax) return false;
    }
    return true;
}

/* mean number of sender chunks (tiles) per receiver block of the tiled layout */
double tiles_per_block(const infoli_ctx *c, int R, int64_t C)
{
    const int64_t nrb = (c->n + R - 1) / R, nch = (c->n + C - 1) / C;
    std::vector<int64_t> seen(nch, -1);
    int64_t tiles = 0;
    for (int64_t rb = 0; rb < nrb; rb++) {
        const int64_t c1 = std::min<int64_t>(c->n, (rb + 1) * R);
        for (int64_t e = c->row_ptr[rb * R]; e < c->row_ptr[c1]; e++) {
            const int64_t ch = c->col[e] / C;
            if (seen[ch] != rb) {
                seen[ch] = rb;
                tiles++;
            }
        }
    }
    return nrb > 0 ? (double)tiles / (double)nrb : 0.0;
}

/* Heuristic launch strategy (infoli_autotune measures instead):
 *  - tiled gather when the per-tile partial sums (one pair per receiver per
 *    tile) cost at most half the junction stream they replace, or when some
 *    cell needs the FP64 sums of the degree budget;
 *  - else the staged gather when a 64-cell group fits the LDS budget and
 *    the neighbour lists fit the Infinity Cache;
 *  - else direct gathers. */
void default_exec(infoli_ctx *c)
{
    const double mean_deg = c->n > 0 ? (double)c->col.size() / (double)c->n : 0;
    c->exec.mode = INFOLI_EXEC_GRAPH;
    c->exec.lanes_per_cell = choose_lanes(mean_deg);
    c->exec.block = kBlock;
    c->exec.steps_per_launch = 100;
    c->exec.lds_kb = kDefaultLdsKb;
    c->exec.tile_cells = 2048;
    c->exec.ode_split = 1;
    c->exec.fast_math = 0;
    c->exec.gather = GATHER_DIRECT;
    // whole networks per workgroup, persistent, when every piece of the
    // network that no junction separates fits one workgroup and the largest
    // range's work, advanced with no launch between steps, beats the graph's
    // estimate: the whole work spread over the CUs plus the two dependent
    // kernels of a step
    {
        NetsPlan pl;
        if (!hyb_active(c) && plan_nets(c, pl)) {
            const double persist_us = pl.max_work * kNetsUsPerTerm;
            const double graph_us = pl.total_work * kGraphUsPerTerm / std::max(c->num_cus, 1) + kGraphStepUs;
            if (persist_us < graph_us) {
                c->exec.mode = INFOLI_EXEC_PERSISTENT;
                c->exec.gather = GATHER_NETS;
                c->exec.lanes_per_cell = 1;
                c->exec.steps_per_launch = 500;
                return;
            }
        }
    }
    // a small FP32 network (at most 2,048 cells, so one sender chunk and at
    // most 32 blocks): swept tiles of 64 receivers, about one lane per 16
    // junctions, the compartments in the sweeping kernel beside the rows --
    // one kernel per step.  plain_1k: 6.97 us per step at L 8 (L 4 7.26,
    // L 16 9.33, R 128 L 4 7.42) against 9.3 for the tiled layout's four
    // tiles and its two dependent kernels (FP64: 14.7 against 14.2, kept tiled)
    if (is_f32(c) && c->n <= 2048 && mean_deg >= 1.0 && !hyb_active(c)) {
        int L = 1;
        while (L < 16 && 16.0 * L <= mean_deg) L *= 2;
        const int kb = (int)std::max<int64_t>(1, ((int64_t)c->n * (int64_t)c->tsize + 1023) / 1024);
        if (sweep_lds_bytes(c, 64, tiled_chunk(c, kb)) <= kLdsBytes) {
            c->exec.gather = GATHER_SWEEP;
            c->exec.lanes_per_cell = L;
            c->exec.block = 1024;
            c->exec.tile_cells = 64;
            c->exec.lds_kb = kb;
            c->exec.ode_split = 0;
            return;
        }
    }
    if (mean_deg < 4 && !hyb_active(c)) return;
    // tiled: receivers per tile from the mean in-tile degree of a first
    // layout guess (2048 / L0 with L0 about one lane per 16-32 junctions),
    // sender chunks small enough for one tile per CU (two when the in-tile
    // degree is high, so partial sums stay cheap) or sized for even rounds,
    // then the lanes from the in-tile degree of that layout, about one lane
    // per 32-64 junctions (longer runs per lane beat more lanes merging)
    const int64_t C0 = tiled_chunk(c, 48);
    const int R0 = 2048;
    const auto in_tile_degree = [c](int R, int64_t C) {
        const double tiles = tiles_per_block(c, R, C) * (double)((c->n + R - 1) / R);
        return tiles > 0 ? (double)c->col.size() / (tiles * R) : 0.0;
    };
    const auto lanes_for = [](double dc, double per_lane) {
        int L = 1;
        while (L < 16 && per_lane * L <= dc) L *= 2;
        return L;
    };
    // block-diagonal networks (many independent networks side by side):
    // when nearly every 512-cell block draws its senders from a range that
    // fits one 48 KiB chunk, each block gets one tile staging just that
    // range.  Then fewer, larger tiles pay (1024 lanes each, about one lane
    // per 64-128 junctions): batched_512 73.8 us per step against 78.2 us
    // with the rule below
    if (!hyb_active(c)) {
        int64_t blocks = 0, local = 0;
        for (int64_t b0 = 0; b0 < c->n; b0 += 512) {
            const int64_t b1 = std::min<int64_t>(c->n, b0 + 512);
            int64_t lo = c->n, hi = -1;
            for (int64_t e = c->row_ptr[b0]; e < c->row_ptr[b1]; e++) {
                lo = std::min<int64_t>(lo, c->col[e]);
                hi = std::max<int64_t>(hi, c->col[e]);
            }
            if (hi < lo) continue;
            blocks++;
            if (hi - lo / 64 * 64 + 1 <= C0) local++;
        }
        if (blocks > 0 && 10 * local >= 9 * blocks && (c->n + 511) / 512 > 2 * (int64_t)c->num_cus) {
            const int L = lanes_for(mean_deg, 128.0);
            const int R = std::max(256, 1024 / L);
            c->exec.gather = GATHER_TILED;
            c->exec.lds_kb = 48;
            c->exec.block = 1024;
            c->exec.lanes_per_cell = L;
            c->exec.tile_cells = R;
            // many more tiles than workgroup slots: somata and axons fill
            // the last round
            c->exec.ode_split = 2;
            return;
        }
    }
    const int L0 = lanes_for(in_tile_degree(R0, C0), 32.0);
    int R = std::max(64, 2048 / L0);
    int64_t nrb = (c->n + R - 1) / R;
    const int64_t want = (L0 >= 4 ? 2 : 1) * (int64_t)c->num_cus;
    const int64_t nch_min = std::max<int64_t>(1, (want + nrb - 1) / nrb);
    int64_t C = std::min<int64_t>(C0, round_up((c->n + nch_min - 1) / nch_min, 64));
    bool even_tiles = false;
    int64_t even_k = 0;   // tiles per CU of the even rounds
    if (nrb <= c->num_cus) {
        // chunks as large as give every CU the same number of tiles (one,
        // or two when one tile's chunk would not fit the LDS), when that is
        // no smaller than the chunk above: uneven rounds leave the CUs that
        // finish early idle, and fewer chunks mean fewer partial sums
        for (int64_t k = 1; k <= 2 && !even_tiles; k++) {
            const int64_t nch = k * c->num_cus / nrb;
            if (nch < 1) continue;
            const int64_t Cb = round_up((c->n + nch - 1) / nch, 64);
            if (Cb >= C && Cb * (int64_t)c->tsize <= 156 * 1024 && k * Cb * (int64_t)c->tsize <= 160 * 1024) {
                C = Cb;
                even_tiles = true;
                even_k = k;
            }
        }
    }
    if (even_k == 2 && 2 * R <= 16384) {
        // two tiles per CU by the LDS may still be one by the registers (the
        // FP64 tile kernel holds one 1024-thread workgroup per CU): then
        // twice the receivers per block give one even round instead of two
        const void *fk = hyb_active(c) ? k_tile_gj_hyb_tab[ntidx(1024)][lidx(L0)]
                                       : k_tile_gj_tab[c->prec][ntidx(1024)][lidx(L0)];
        const size_t lds = (size_t)C * c->tsize;
        allow_dyn_lds(fk, lds);
        int pc = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, fk, 1024, lds) != hipSuccess) pc = 1;
        const int64_t nrb2 = (c->n + 2 * R - 1) / (2 * R);
        const int64_t nch2 = c->num_cus / nrb2;
        const int64_t C2 = nch2 >= 1 ? round_up((c->n + nch2 - 1) / nch2, 64) : 0;
        if (pc < 2 && nch2 >= 1 && C2 * (int64_t)c->tsize <= 156 * 1024) {
            R *= 2;
            nrb = nrb2;
            C = C2;
        }
    }
    const int kb = (int)std::max<int64_t>(1, (C * (int64_t)c->tsize + 1023) / 1024);
    const int64_t Ck = tiled_chunk(c, kb);
    const double dc = in_tile_degree(R, Ck);
    const int L = lanes_for(dc, 64.0);
    const double psz = c->prec == INFOLI_FP32 ? 4.0 : 8.0;
    const double tpb = tiles_per_block(c, R, Ck);
    if (hyb_active(c) || tpb * psz <= 0.5 * mean_deg * (2.0 + (double)c->tsize)) {
        c->exec.gather = GATHER_TILED;
        c->exec.lds_kb = kb;
        c->exec.block = 1024;
        c->exec.lanes_per_cell = L;
        c->exec.tile_cells = R;
        // somata, axons and dendrite gates beside the tiles when the tiles
        // need more than one round of workgroup slots (they then fill the
        // last round), or, with chunks sized for even rounds, only when every
        // CU holds one tile and has a slot to spare and the tile rows are
        // short (the tiles then wait on memory and leave issue slots free);
        // in a single uneven round, in full even rounds or beside issue-bound
        // tiles they would only lengthen the launch
        const void *fg = hyb_active(c) ? k_tile_gj_hyb_tab[ntidx(1024)][lidx(L)]
                                       : k_tile_gj_tab[c->prec][ntidx(1024)][lidx(L)];
        const size_t lds = (size_t)Ck * c->tsize;
        allow_dyn_lds(fg, lds);
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fg, 1024, lds) != hipSuccess) per_cu = 1;
        per_cu = std::max(per_cu, 1);
        const double ntiles = tpb * (double)((c->n + R - 1) / R);
        const bool spare = even_tiles && per_cu >= 2 && ntiles <= (double)c->num_cus && dc < 64.0;
        c->exec.ode_split = (even_tiles ? spare : ntiles > (double)per_cu * c->num_cus) ? 2 : 1;
        return;
    }
    // swept tiles for FP32 lists beyond the Infinity Cache with at least 36
    // junctions per cell: one 1024-thread workgroup per receiver block sums
    // the block's tiles in LDS.  A block's sums and chunk fill a CU's LDS, so
    // the blocks number at most the CUs (R = n / CUs, rounded up to 64), and
    // the chunk takes the LDS the sums leave, down to 96 KiB at R 8192.
    // Measured (FP32 compensated, us per step, direct -> swept): 1M cells of
    // degree 50 372 -> 275, of degree 100 681 -> 337, of degree 40
    // 308 -> 274; 500k cells of degree 50 182 -> 114; 1.5M 599 -> 523; 2M
    // 1008 -> 909.  The direct gather stays ahead at degree 30 (1M cells:
    // 244 -> 265).  The somata, axons and dendrite gates run in a kernel
    // beside the sweep (ode_split 2): 264 us at 1M cells, against 271-278
    // with the compartments in the sweeping kernel (ode_split 0) and 272 with
    // k_tile_ode3 after it (1).  In FP64 (direct -> swept with ode_split 2):
    // 1M cells of degree 40 483-500 -> 505, of degree 50 573-587 -> 538, of
    // degree 100 979 -> 662; but 500k cells of degree 50 224 -> 235-251,
    // where V_dend (4 MB) fits one XCD's 4 MiB L2 and the direct gather's
    // loads hit it.  So FP64 sweeps from degree 45 when V_dend exceeds 4 MiB
    const bool sweep64 = !is_f32(c) && mean_deg >= 45.0 && (size_t)c->n * sizeof(double) > ((size_t)4 << 20);
    if (stream_exceeds_cache(c) && ((is_f32(c) && mean_deg >= 36.0) || sweep64)) {
        const int64_t R = round_up((c->n + c->num_cus - 1) / std::max(c->num_cus, 1), 64);
        const int kb = R <= 8192 ? sweep_max_kb(c, (int)R) : 0;
        if (kb >= 16) {
            c->exec.gather = GATHER_SWEEP;
            c->exec.lanes_per_cell = 1;
            c->exec.block = 1024;
            c->exec.tile_cells = (int)R;
            c->exec.lds_kb = kb;
            c->exec.ode_split = 2;
            return;
        }
    }
    // staged gather when a 64-cell group fits and the lists stay in the
    // Infinity Cache; beyond it both stream from HBM and the direct gather,
    // with nontemporal list loads, is the faster
    if (max_block_entries_fit(c, 40) && !stream_exceeds_cache(c)) {
        c->exec.gather = GATHER_STAGED;
        c->exec.lds_kb = 40;
    }
}

int ensure_stage(infoli_ctx *c, size_t bytes)
{
    if (c->h_stage_bytes >= bytes) return INFOLI_OK;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->h_stage_bytes = 0;
    HIPCHK(hipHostMalloc(&c->h_stage, bytes, hipHostMallocDefault));
    c->h_stage_bytes = bytes;
    return INFOLI_OK;
}

int validate_ctx(const infoli_t *c, const char *who)
{
    if (!c) return infoli_fail(INFOLI_ERR_ARG, "%s: null context", who);
    return INFOLI_OK;
}

/* device T[n] <- double[n] */
int upload_values(infoli_ctx *c, void *dst, const double *src, int64_t n)
{
    RCHK(ensure_stage(c, (size_t)n * c->tsize));
    if (is_f32(c)) {
        float *s = (float *)c->h_stage;
        for (int64_t i = 0; i < n; i++) s[i] = (float)src[i];
    } else {
        memcpy(c->h_stage, src, sizeof(double) * n);
    }
    HIPCHK(hipMemcpyAsync(dst, c->h_stage, (size_t)n * c->tsize, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return INFOLI_OK;
}

int download_values(infoli_ctx *c, const void *src, double *dst, int64_t n)
{
    RCHK(ensure_stage(c, (size_t)n * c->tsize));
    HIPCHK(hipMemcpyAsync(c->h_stage, src, (size_t)n * c->tsize, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (is_f32(c)) {
        const float *s = (const float *)c->h_stage;
        for (int64_t i = 0; i < n; i++) dst[i] = (double)s[i];
    } else {
        memcpy(dst, c->h_stage, sizeof(double) * n);
    }
    return INFOLI_OK;
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* lifecycle                                                                 */
/* ------------------------------------------------------------------------ */

extern "C" const char *infoli_version(void) { return "infoli-mi355x 0.1 (gfx950)"; }

extern "C" int infoli_create(infoli_t **out, const infoli_params *params, int64_t n_cells,
                             int precision, int device)
{
    infoli_clear_error();
    if (!out) return infoli_fail(INFOLI_ERR_ARG, "infoli_create: null output");
    *out = nullptr;
    if (!params) return infoli_fail(INFOLI_ERR_ARG, "infoli_create: null params");
    if (n_cells <= 0 || n_cells > INT32_MAX)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_create: n_cells = %lld out of range",
                           (long long)n_cells);
    if (precision < INFOLI_FP64 || precision > INFOLI_FP32_KAHAN)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_create: unknown precision %d", precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return infoli_fail(INFOLI_ERR_NODEVICE, "infoli_create: no HIP device visible");
    if (device < 0 || device >= ndev)
        return infoli_fail(INFOLI_ERR_NODEVICE, "infoli_create: device %d of %d", device, ndev);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return infoli_fail(INFOLI_ERR_NODEVICE, "infoli_create: device %d is %s, built for gfx950",
                           device, prop.gcnArchName);
    HIPCHK(hipSetDevice(device));
    RCHK(scratch_setup(device));

    infoli_ctx *c = new (std::nothrow) infoli_ctx();
    if (!c) return infoli_fail(INFOLI_ERR_NOMEM, "infoli_create: no host memory");
    c->device = device;
    c->prec = precision;
    c->n = n_cells;
    c->npad = round_up(n_cells, kBlock);
    c->tsize = precision == INFOLI_FP64 ? 8 : 4;
    c->fp64_deg = precision == INFOLI_FP32_KAHAN ? INFOLI_FP64_DEGREE_DEFAULT : 0;
    c->params = *params;
    c->num_cus = prop.multiProcessorCount;
    c->row_ptr.assign(n_cells + 1, 0);

    int rc = INFOLI_OK;
    auto fail = [&](int code) {
        infoli_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_copy[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_copy[1], hipEventDisableTiming) != hipSuccess)
        return fail(infoli_fail(INFOLI_ERR_HIP, "infoli_create: stream/event creation failed"));
    const size_t plane = (size_t)c->npad * c->tsize;
    if (hipMalloc(&c->d_state, plane * kNV) != hipSuccess ||
        hipMalloc(&c->d_vd[0], plane) != hipSuccess || hipMalloc(&c->d_vd[1], plane) != hipSuccess ||
        hipMalloc(&c->d_vs[0], plane) != hipSuccess || hipMalloc(&c->d_vs[1], plane) != hipSuccess ||
        hipMalloc(&c->d_va[0], plane) != hipSuccess || hipMalloc(&c->d_va[1], plane) != hipSuccess ||
        hipMalloc(&c->d_deg, sizeof(int32_t) * c->npad) != hipSuccess ||
        hipMalloc(&c->d_step, sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&c->d_status, sizeof(AdvanceStatus)) != hipSuccess ||
        hipHostMalloc((void **)&c->h_status, sizeof(AdvanceStatus), hipHostMallocDefault) != hipSuccess)
        return fail(infoli_fail(INFOLI_ERR_NOMEM, "infoli_create: device allocation of %lld cells failed",
                                (long long)n_cells));
    c->d_bar = &c->d_status->bar;
    c->d_flag = &c->d_status->nonfinite;
    // zeroed on the context's stream, ahead of the initial-state uploads on
    // that stream (a null-stream hipMemset is not ordered before them, and a
    // voltage plane zeroed after its upload drives the network to NaN)
    for (void *p : {c->d_vd[0], c->d_vd[1], c->d_vs[0], c->d_vs[1], c->d_va[0], c->d_va[1]}) {
        rc = zero_dev(c, p, plane);
        if (rc) return fail(rc);
    }
    rc = zero_dev(c, c->d_state, plane * kNV);
    if (rc == INFOLI_OK) rc = zero_dev(c, c->d_deg, sizeof(int32_t) * c->npad);
    if (rc) return fail(rc);

    // initial state (OMP:267-292): float readings for FP32, double for FP64
    std::vector<double> vals(n_cells);
    for (int v = 0; v < kNV; v++) {
        const double x = precision == INFOLI_FP64 ? params->v[INFOLI_P_STATE0 + v]
                                                  : (double)params->f[INFOLI_P_STATE0 + v];
        std::fill(vals.begin(), vals.end(), x);
        void *dst = (v == V_DEND) ? c->d_vd[0] : (v == V_SOMA) ? c->d_vs[0] : (v == V_AXON) ? c->d_va[0]
                                                                         : (char *)c->d_state + (size_t)v * plane;
        rc = upload_values(c, dst, vals.data(), n_cells);
        if (rc) return fail(rc);
    }
    default_exec(c);
    rc = ensure_layout(c);
    if (rc) return fail(rc);
    *out = c;
    return INFOLI_OK;
}

extern "C" void infoli_destroy(infoli_t *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    for (hipEvent_t e : c->ev_copy)
        if (e) (void)hipEventDestroy(e);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    void *bufs[] = {c->d_state, c->d_vd[0], c->d_vd[1], c->d_vs[0], c->d_vs[1], c->d_va[0], c->d_va[1], c->d_deg, c->d_slice_ptr, c->d_col, c->d_w,
                    c->d_iapp_base, c->d_mask, c->d_tab, c->d_trace, c->d_step, c->d_status};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free_staged(c);
    free_tiled(c);
    if (c->d_rng) (void)hipFree(c->d_rng);
    if (c->d_nets_lists) (void)hipFree(c->d_nets_lists);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

/* ------------------------------------------------------------------------ */
/* inputs                                                                    */
/* ------------------------------------------------------------------------ */

int check_exec(const infoli_ctx *c, infoli_exec &x);

extern "C" int infoli_set_connectivity_csr(infoli_t *c, const int64_t *row_ptr,
                                           const int32_t *col, const double *w)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_connectivity_csr"));
    if (!row_ptr) return infoli_fail(INFOLI_ERR_ARG, "infoli_set_connectivity_csr: null row_ptr");
    const int64_t n = c->n;
    if (row_ptr[0] != 0) return infoli_fail(INFOLI_ERR_ARG, "row_ptr[0] = %lld, must be 0", (long long)row_ptr[0]);
    int max_deg = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t d = row_ptr[i + 1] - row_ptr[i];
        if (d < 0 || d > INT32_MAX)
            return infoli_fail(INFOLI_ERR_ARG, "row_ptr not non-decreasing at cell %lld", (long long)i);
        max_deg = std::max<int>(max_deg, (int)d);
    }
    const int64_t nnz = row_ptr[n];
    if (nnz > 0 && (!col || !w))
        return infoli_fail(INFOLI_ERR_ARG, "infoli_set_connectivity_csr: null col or w with %lld entries",
                           (long long)nnz);
    for (int64_t e = 0; e < nnz; e++) {
        if (col[e] < 0 || col[e] >= n)
            return infoli_fail(INFOLI_ERR_ARG, "col[%lld] = %d outside [0, %lld)", (long long)e, col[e],
                               (long long)n);
        if (!std::isfinite(w[e]))
            return infoli_fail(INFOLI_ERR_ARG, "w[%lld] is not finite", (long long)e);
    }
    for (size_t k = 0; k + 1 < c->part.size(); k++) {
        for (int64_t i = c->part[k]; i < c->part[k + 1]; i++)
            for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; e++)
                if (col[e] < c->part[k] || col[e] >= c->part[k + 1])
                    return infoli_fail(INFOLI_ERR_ARG,
                                       "junction %d -> %lld crosses the partition boundary of network %zu",
                                       col[e], (long long)i, k);
    }
    c->row_ptr.assign(row_ptr, row_ptr + n + 1);
    c->col.assign(col ? col : nullptr, col ? col + nnz : nullptr);
    c->w.assign(w ? w : nullptr, w ? w + nnz : nullptr);
    c->max_deg = max_deg;
    c->staged_valid = false;
    c->tiled_valid = false;
    c->nets_valid = false;
    default_exec(c);
    // a launch mode the caller chose is kept where the new strategy has it
    // (whole networks per workgroup are persistent by construction)
    if (c->user_mode >= 0 && c->exec.gather != GATHER_NETS) {
        infoli_exec x = c->exec;
        x.mode = c->user_mode;
        if (check_exec(c, x) == INFOLI_OK) c->exec = x;
        infoli_clear_error();
    }
    c->sell_L = 0;
    drop_graph(c);
    return ensure_layout(c);
}

extern "C" int infoli_set_fp64_degree(infoli_t *c, int64_t min_degree)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_fp64_degree"));
    if (min_degree < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_set_fp64_degree: negative degree");
    if (min_degree != 0 && c->prec != INFOLI_FP32_KAHAN)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_set_fp64_degree: FP64 sums for high-degree cells apply "
                                           "to INFOLI_FP32_KAHAN only");
    if (min_degree == c->fp64_deg) return INFOLI_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fp64_deg = min_degree;
    if (hyb_active(c) && !tiled_layout(c->exec)) {
        const int keep_mode = c->exec.mode;
        default_exec(c);
        if (keep_mode != INFOLI_EXEC_PERSISTENT) c->exec.mode = keep_mode;
    }
    drop_graph(c);
    return ensure_layout(c);
}

extern "C" int64_t infoli_get_fp64_degree(const infoli_t *c)
{
    return c ? c->fp64_deg : -1;
}

extern "C" int infoli_set_partition(infoli_t *c, int n_nets, const int64_t *offsets)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_partition"));
    if (n_nets <= 0 || !offsets) return infoli_fail(INFOLI_ERR_ARG, "infoli_set_partition: bad arguments");
    if (offsets[0] != 0 || offsets[n_nets] != c->n)
        return infoli_fail(INFOLI_ERR_ARG, "partition must start at 0 and end at n_cells");
    for (int k = 0; k < n_nets; k++)
        if (offsets[k + 1] <= offsets[k])
            return infoli_fail(INFOLI_ERR_ARG, "network %d is empty or offsets decrease", k);
    for (int k = 0; k < n_nets; k++)
        for (int64_t i = offsets[k]; i < offsets[k + 1]; i++)
            for (int64_t e = c->row_ptr[i]; e < c->row_ptr[i + 1]; e++)
                if (c->col[e] < offsets[k] || c->col[e] >= offsets[k + 1])
                    return infoli_fail(INFOLI_ERR_ARG,
                                       "junction %d -> %lld crosses the boundary of network %d",
                                       c->col[e], (long long)i, k);
    c->part.assign(offsets, offsets + n_nets + 1);
    return INFOLI_OK;
}

extern "C" int infoli_set_gcal(infoli_t *c, const double *g)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_gcal"));
    return infoli_set_state(c, INFOLI_G_CAL, g);
}

extern "C" int infoli_set_iapp(infoli_t *c, const double *base)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_iapp"));
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!base) {
        if (c->d_iapp_base) HIPCHK(hipFree(c->d_iapp_base));
        c->d_iapp_base = nullptr;
        drop_graph(c);
        return INFOLI_OK;
    }
    for (int64_t i = 0; i < c->n; i++)
        if (!std::isfinite(base[i])) return infoli_fail(INFOLI_ERR_ARG, "iapp base[%lld] not finite", (long long)i);
    if (!c->d_iapp_base) HIPCHK(hipMalloc(&c->d_iapp_base, (size_t)c->n * c->tsize));
    drop_graph(c);
    return upload_values(c, c->d_iapp_base, base, c->n);
}

extern "C" int infoli_set_iapp_pulse(infoli_t *c, const uint8_t *mask, double amp, int64_t t0, int64_t t1)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_iapp_pulse"));
    if (!std::isfinite(amp)) return infoli_fail(INFOLI_ERR_ARG, "pulse amplitude not finite");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (mask) {
        if (!c->d_mask) HIPCHK(hipMalloc(&c->d_mask, (size_t)c->n));
        RCHK(h2d(c, c->d_mask, mask, (size_t)c->n));
    } else if (c->d_mask) {
        HIPCHK(hipFree(c->d_mask));
        c->d_mask = nullptr;
    }
    c->amp = amp;
    c->t0 = t0;
    c->t1 = t1;
    drop_graph(c);
    return INFOLI_OK;
}

extern "C" int infoli_set_iapp_table(infoli_t *c, int64_t t_first, int64_t n_steps, const double *table)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_iapp_table"));
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_tab) HIPCHK(hipFree(c->d_tab));
    c->d_tab = nullptr;
    c->tab_n = 0;
    drop_graph(c);
    if (n_steps <= 0) return INFOLI_OK;
    if (!table || t_first < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_set_iapp_table: bad table");
    const int64_t cnt = n_steps * c->n;
    HIPCHK(hipMalloc(&c->d_tab, (size_t)cnt * c->tsize));
    RCHK(upload_values(c, c->d_tab, table, cnt));
    c->tab_t0 = t_first;
    c->tab_n = n_steps;
    return INFOLI_OK;
}

static void *var_ptr(infoli_ctx *c, int var)
{
    if (var == V_DEND) return c->d_vd[c->step & 1];
    if (var == V_SOMA) return c->d_vs[c->step & 1];
    if (var == V_AXON) return c->d_va[c->step & 1];
    return (char *)c->d_state + (size_t)var * c->npad * c->tsize;
}

extern "C" int infoli_set_state(infoli_t *c, int var, const double *values)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_state"));
    if (var < 0 || var >= INFOLI_NVARS || !values)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_set_state: bad variable %d or null values", var);
    HIPCHK(hipSetDevice(c->device));
    return upload_values(c, var_ptr(c, var), values, c->n);
}

extern "C" int infoli_get_state(infoli_t *c, int var, double *values)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_get_state"));
    if (var < 0 || var >= INFOLI_NVARS || !values)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_get_state: bad variable %d or null values", var);
    HIPCHK(hipSetDevice(c->device));
    return download_values(c, var_ptr(c, var), values, c->n);
}

/* ------------------------------------------------------------------------ */
/* state files and the isolated warm-up                                      */
/* ------------------------------------------------------------------------ */

extern "C" int infoli_write_state(infoli_t *c, const char *path)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_write_state"));
    if (!path) return infoli_fail(INFOLI_ERR_ARG, "infoli_write_state: null path");
    // printState's column order (src/single_threaded/infoli.c:977-992)
    static const int cols[16] = {INFOLI_V_SOMA,      INFOLI_SODIUM_M,  INFOLI_POTASSIUM_N,  INFOLI_POTASSIUM_X_S,
                                 INFOLI_CALCIUM_K,   INFOLI_CALCIUM_L, INFOLI_V_DEND,       INFOLI_CALCIUM_R,
                                 INFOLI_POTASSIUM_S, INFOLI_HCURRENT_Q, INFOLI_CA2PLUS,     INFOLI_I_CAH,
                                 INFOLI_V_AXON,      INFOLI_SODIUM_M_A, INFOLI_SODIUM_H_A,  INFOLI_POTASSIUM_X_A};
    std::vector<double> v((size_t)16 * c->n);
    for (int j = 0; j < 16; j++) RCHK(infoli_get_state(c, cols[j], v.data() + (size_t)j * c->n));
    FILE *fd = fopen(path, "w");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "infoli_write_state: cannot create %s", path);
    for (int64_t i = 0; i < c->n; i++) {
        for (int j = 0; j < 16; j++) fprintf(fd, "%.8lf ", v[(size_t)j * c->n + i]);
        fprintf(fd, "\n");
    }
    if (fclose(fd) != 0) return infoli_fail(INFOLI_ERR_IO, "infoli_write_state: write to %s failed", path);
    return INFOLI_OK;
}

namespace {
struct CheckpointHeader {
    char magic[8];        // "INFOLICK"
    int32_t version;      // 1
    int32_t precision;    // INFOLI_FP64 / FP32 / FP32_KAHAN of the writer
    int64_t n_cells;
    int64_t step;
    int32_t n_vars;       // INFOLI_NVARS
    int32_t reserved;
};
}  // namespace

extern "C" int infoli_save_checkpoint(infoli_t *c, const char *path)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_save_checkpoint"));
    if (!path) return infoli_fail(INFOLI_ERR_ARG, "infoli_save_checkpoint: null path");
    std::vector<double> v((size_t)INFOLI_NVARS * c->n);
    for (int j = 0; j < INFOLI_NVARS; j++) RCHK(infoli_get_state(c, j, v.data() + (size_t)j * c->n));
    CheckpointHeader h{};
    memcpy(h.magic, "INFOLICK", 8);
    h.version = 1;
    h.precision = c->prec;
    h.n_cells = c->n;
    h.step = c->step;
    h.n_vars = INFOLI_NVARS;
    FILE *fd = fopen(path, "wb");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "infoli_save_checkpoint: cannot create %s", path);
    const bool ok = fwrite(&h, sizeof h, 1, fd) == 1 && fwrite(v.data(), sizeof(double), v.size(), fd) == v.size();
    if (fclose(fd) != 0 || !ok) return infoli_fail(INFOLI_ERR_IO, "infoli_save_checkpoint: write to %s failed", path);
    return INFOLI_OK;
}

extern "C" int infoli_load_checkpoint(infoli_t *c, const char *path)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_load_checkpoint"));
    if (!path) return infoli_fail(INFOLI_ERR_ARG, "infoli_load_checkpoint: null path");
    FILE *fd = fopen(path, "rb");
    if (!fd) return infoli_fail(INFOLI_ERR_IO, "infoli_load_checkpoint: cannot open %s", path);
    CheckpointHeader h{};
    std::vector<double> v;
    bool ok = fread(&h, sizeof h, 1, fd) == 1 && memcmp(h.magic, "INFOLICK", 8) == 0 && h.version == 1 &&
              h.n_vars == INFOLI_NVARS;
    if (ok && h.n_cells != c->n) {
        fclose(fd);
        return infoli_fail(INFOLI_ERR_ARG, "infoli_load_checkpoint: %s holds %lld cells, the network has %lld", path,
                           (long long)h.n_cells, (long long)c->n);
    }
    if (ok) {
        v.resize((size_t)INFOLI_NVARS * c->n);
        ok = fread(v.data(), sizeof(double), v.size(), fd) == v.size() && h.step >= 0;
    }
    fclose(fd);
    if (!ok) return infoli_fail(INFOLI_ERR_IO, "infoli_load_checkpoint: %s is not an infoli checkpoint", path);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = h.step;   // first: the voltages go to the buffers of this step's parity
    c->begin_done = false;
    for (int j = 0; j < INFOLI_NVARS; j++) RCHK(infoli_set_state(c, j, v.data() + (size_t)j * c->n));
    return INFOLI_OK;
}

extern "C" int infoli_isolated_steps(infoli_t *c, const int64_t *steps)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_isolated_steps"));
    if (!steps) return infoli_fail(INFOLI_ERR_ARG, "infoli_isolated_steps: null steps");
    for (int64_t i = 0; i < c->n; i++)
        if (steps[i] < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_isolated_steps: steps[%lld] < 0", (long long)i);
    HIPCHK(hipSetDevice(c->device));
    int64_t *d_steps = nullptr;
    HIPCHK(hipMalloc(&d_steps, sizeof(int64_t) * c->n));
    int rc = INFOLI_OK;
    rc = h2d(c, d_steps, steps, sizeof(int64_t) * c->n);
    if (rc == INFOLI_OK) {
        const dim3 grid((unsigned)((c->n + kBlock - 1) / kBlock));
        if (is_f32(c))
            hipLaunchKernelGGL(k_isolated<float>, grid, dim3(kBlock), 0, c->stream, make_args<float>(c, 0), c->step,
                               (const int64_t *)d_steps);
        else
            hipLaunchKernelGGL(k_isolated<double>, grid, dim3(kBlock), 0, c->stream, make_args<double>(c, 0), c->step,
                               (const int64_t *)d_steps);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = infoli_fail(INFOLI_ERR_HIP, "infoli_isolated_steps: kernel failed");
    }
    (void)hipFree(d_steps);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* execution                                                                 */
/* ------------------------------------------------------------------------ */

/* Validate an execution strategy for this network (normalising the block
 * size of the non-tiled gathers); INFOLI_OK or INFOLI_ERR_ARG with the
 * reason in the error message. */
int check_exec(const infoli_ctx *c, infoli_exec &x)
{
    if (x.mode < INFOLI_EXEC_STEP || x.mode > INFOLI_EXEC_PERSISTENT)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_set_exec: mode %d not available", x.mode);
    if (lidx(x.lanes_per_cell) < 0)
        return infoli_fail(INFOLI_ERR_ARG, "lanes_per_cell %d not in {1,2,4,8,16}", x.lanes_per_cell);
    if (x.gather == GATHER_SWEEP) x.block = 1024;       // one workgroup of 1024 threads per receiver block
    else if (x.gather != GATHER_TILED) x.block = kBlock;   // the fused step kernel has one block size
    if (x.gather == GATHER_NETS) x.lanes_per_cell = 1;    // one thread per cell (the workgroup size is planned)
    if (x.block != 256 && x.block != 512 && x.block != 1024)
        return infoli_fail(INFOLI_ERR_ARG, "block %d: the tiled gather takes 256, 512 or 1024", x.block);
    if (x.steps_per_launch < 1 || x.steps_per_launch > 100000)
        return infoli_fail(INFOLI_ERR_ARG, "steps_per_launch %d out of range", x.steps_per_launch);
    if (x.gather < GATHER_DIRECT || x.gather > GATHER_NETS)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_set_exec: gather %d unknown", x.gather);
    if (tiled_layout(x) && (x.tile_cells < 64 || x.tile_cells > 16384 || x.tile_cells % 64))
        return infoli_fail(INFOLI_ERR_ARG, "tile_cells %d: must be a multiple of 64 in [64, 16384]",
                           x.tile_cells);
    if (x.fast_math && !tiled_layout(x))
        return infoli_fail(INFOLI_ERR_ARG, "fast_math is available with the tiled gathers only");
    if (x.ode_split < 0 || x.ode_split > 2)
        return infoli_fail(INFOLI_ERR_ARG, "ode_split %d not in {0, 1, 2}", x.ode_split);
    if (hyb_active(c) && !tiled_layout(x))
        return infoli_fail(INFOLI_ERR_ARG, "cells with at least %lld neighbours sum FP64 gap-junction terms, "
                                           "which the tiled gathers implement; set the FP64 degree to 0 "
                                           "for another gather", (long long)c->fp64_deg);
    if (x.gather == GATHER_NETS) {
        if (x.mode != INFOLI_EXEC_PERSISTENT)
            return infoli_fail(INFOLI_ERR_ARG, "gather NETS (whole networks per workgroup) runs in the persistent "
                                               "launch mode only");
        if (hyb_active(c))
            return infoli_fail(INFOLI_ERR_ARG, "gather NETS sums FP32 terms only; set the FP64 degree to 0");
        if (independent_cuts(c, nets_cap(c)).empty())
            return infoli_fail(INFOLI_ERR_ARG, "gather NETS: a piece of the network that no junction separates has "
                                               "more than %lld cells, more than one workgroup holds",
                               (long long)nets_cap(c));
    }
    if (tiled_layout(x) && x.mode == INFOLI_EXEC_PERSISTENT)
        return infoli_fail(INFOLI_ERR_ARG, "the tiled gathers run as two kernels per step; "
                                           "use the step or graph launch mode");
    if (x.gather == GATHER_VLDS && (size_t)c->n * c->tsize > kVldsMaxBytes)
        return infoli_fail(INFOLI_ERR_ARG, "gather LDS: %lld cells do not fit the LDS budget", (long long)c->n);
    if (x.lds_kb < 1 || x.lds_kb > 156)
        return infoli_fail(INFOLI_ERR_ARG, "lds_kb %d outside [1, 156]", x.lds_kb);
    if (x.gather == GATHER_SWEEP && sweep_lds_bytes(c, x.tile_cells, tiled_chunk(c, x.lds_kb)) > kLdsBytes)
        return infoli_fail(INFOLI_ERR_ARG, "swept tiles: %d receivers' sums plus a %d KiB chunk exceed the "
                                           "160 KiB of LDS", x.tile_cells, x.lds_kb);
    if (x.gather == GATHER_STAGED && !max_block_entries_fit(c, x.lds_kb))
        return infoli_fail(INFOLI_ERR_ARG, "staged gather: a 64-cell group exceeds the %d KiB LDS budget",
                           x.lds_kb);
    if (x.gather == GATHER_STAGED &&
        ((int64_t)x.lds_kb * 1024 + (int64_t)(c->max_deg + 64 + kBlock) * (int64_t)c->tsize) > 160 * 1024)
        return infoli_fail(INFOLI_ERR_ARG, "staged gather: %d KiB plus the row slack of degree %d exceeds "
                                           "the 160 KiB of LDS", x.lds_kb, c->max_deg);
    return INFOLI_OK;
}

extern "C" int infoli_get_exec(const infoli_t *c, infoli_exec *out)
{
    RCHK(validate_ctx(c, "infoli_get_exec"));
    if (!out) return infoli_fail(INFOLI_ERR_ARG, "infoli_get_exec: null output");
    *out = c->exec;
    return INFOLI_OK;
}

extern "C" int infoli_set_exec(infoli_t *c, const infoli_exec *e)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_set_exec"));
    if (!e) return infoli_fail(INFOLI_ERR_ARG, "infoli_set_exec: null exec");
    infoli_exec x = *e;
    if (x.mode == INFOLI_EXEC_AUTO) {
        default_exec(c);
        x = c->exec;
    }
    RCHK(check_exec(c, x));
    const infoli_exec prev = c->exec;
    c->exec = x;
    drop_graph(c);
    int rc = ensure_layout(c);
    if (rc != INFOLI_OK) c->exec = prev;
    else c->user_mode = e->mode == INFOLI_EXEC_AUTO ? -1 : x.mode;
    return rc;
}

extern "C" int infoli_set_profiling(infoli_t *c, int on)
{
    RCHK(validate_ctx(c, "infoli_set_profiling"));
    c->profiling = on != 0;
    return INFOLI_OK;
}

extern "C" int64_t infoli_trace_records(const infoli_t *c, int64_t n_steps, int64_t stride)
{
    if (!c || n_steps <= 0 || stride <= 0) return 0;
    const int64_t a = c->step, b = c->step + n_steps;   // records: multiples of stride in [a, b)
    return (b - 1) / stride - (a + stride - 1) / stride + 1;
}

extern "C" int64_t infoli_step_count(const infoli_t *c) { return c ? c->step : -1; }

namespace {

int prof_begin(infoli_ctx *c)
{
    if (!c->profiling) return INFOLI_OK;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    c->prof_ev.push_back(a);
    c->prof_ev.push_back(b);
    HIPCHK(hipEventRecord(a, c->stream));
    return INFOLI_OK;
}

int prof_end(infoli_ctx *c)
{
    if (!c->profiling) return INFOLI_OK;
    HIPCHK(hipEventRecord(c->prof_ev.back(), c->stream));
    return INFOLI_OK;
}

template <typename T>
int capture_graph(infoli_ctx *c, const StepArgs<T> &a, int mode, int G, int grid)
{
    GraphKey key;
    key.mode = mode;
    key.L = c->exec.lanes_per_cell;
    key.gather = c->exec.gather;
    key.G = G;
    key.grid = grid;
    key.args_len = sizeof a;
    memcpy(key.args, &a, sizeof a);
    if (c->gexec && c->gkey.mode == key.mode && c->gkey.L == key.L && c->gkey.gather == key.gather &&
        c->gkey.G == key.G && c->gkey.grid == key.grid && c->gkey.args_len == key.args_len &&
        memcmp(c->gkey.args, key.args, key.args_len) == 0)
        return INFOLI_OK;
    drop_graph(c);
    c->n_captures++;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = INFOLI_OK;
    if (mode == INFOLI_EXEC_GRAPH) {
        for (int k = 0; k < G && rc == INFOLI_OK; k++) rc = launch_step(c, &a, k);
    } else {
        rc = launch_persist(c, &a, G, grid);
    }
    if (rc == INFOLI_OK) {
        hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(1), 0, c->stream, c->d_step, (int64_t)G);
        if (hipGetLastError() != hipSuccess) rc = infoli_fail(INFOLI_ERR_HIP, "bump launch failed in capture");
    }
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(c->stream, &g);
    if (rc != INFOLI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ee != hipSuccess) return infoli_fail(INFOLI_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphUpload(c->gexec, c->stream));   // its first replay then runs at full speed
    c->gkey = key;
    return INFOLI_OK;
}

/* Capture now the graph that run_steps replays for a piece of `count`
 * steps (graph and persistent modes; nothing to do for fewer steps than one
 * launch holds, or while profiling, when steps are launched one by one).
 * Cached: a later call with the same strategy and arguments reuses it. */
template <typename T>
int prepare_graph(infoli_ctx *c, StepArgs<T> a, int64_t count)
{
    const int G = c->exec.steps_per_launch;
    if (c->profiling || count < G) return INFOLI_OK;
    a.d_step = c->d_step;
    if (c->exec.mode == INFOLI_EXEC_GRAPH) return capture_graph(c, a, INFOLI_EXEC_GRAPH, G, 0);
    if (c->exec.mode == INFOLI_EXEC_PERSISTENT && !tiled_layout(c->exec)) {
        const int grid = persist_grid(c);
        if (grid <= 0) return grid < 0 ? grid : infoli_fail(INFOLI_ERR_HIP, "no persistent grid");
        return capture_graph(c, a, INFOLI_EXEC_PERSISTENT, G, grid);
    }
    return INFOLI_OK;
}

/* steps [t, t+count) in the context's execution mode */
template <typename T>
int run_steps(infoli_ctx *c, StepArgs<T> a, int64_t t, int64_t count)
{
    int mode = c->exec.mode;
    if (c->profiling && mode == INFOLI_EXEC_GRAPH) mode = INFOLI_EXEC_STEP;
    if (mode == INFOLI_EXEC_STEP) {
        a.d_step = nullptr;
        c->n_direct += count;
        for (int64_t k = 0; k < count; k++) {
            RCHK(prof_begin(c));
            RCHK(launch_step(c, &a, t + k));
            RCHK(prof_end(c));
            if (c->profiling) c->kern_launches++;
        }
        return INFOLI_OK;
    }
    // the device step counter holds t here: k_begin set it at the start of
    // the advance (or the last advance left it there), each replay adds its
    // steps, and so do the bumps after steps launched one by one
    const int G = c->exec.steps_per_launch;
    a.d_step = c->d_step;
    int64_t done = 0;
    if (mode == INFOLI_EXEC_GRAPH) {
        RCHK(prepare_graph(c, a, count));
        for (; done + G <= count; done += G) {
            HIPCHK(hipGraphLaunch(c->gexec, c->stream));
            c->n_replays++;
        }
        StepArgs<T> b = a;
        b.d_step = nullptr;
        const int64_t rem = count - done;
        c->n_direct += rem;
        for (; done < count; done++) RCHK(launch_step(c, &b, t + done));
        if (rem > 0) {
            hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(1), 0, c->stream, c->d_step, rem);
            HIPCHK(hipGetLastError());
        }
        return INFOLI_OK;
    }
    // persistent
    if (tiled_layout(c->exec))
        return infoli_fail(INFOLI_ERR_ARG, "tiled gather has no persistent form");
    int grid = persist_grid(c);
    if (grid <= 0) return grid < 0 ? grid : infoli_fail(INFOLI_ERR_HIP, "no persistent grid");
    RCHK(prepare_graph(c, a, count));
    while (done < count) {
        const int piece = (int)std::min<int64_t>(G, count - done);
        if (!c->profiling && piece == G && c->gexec) {
            HIPCHK(hipGraphLaunch(c->gexec, c->stream));
            c->n_replays++;
        } else {
            c->n_direct += piece;
            RCHK(prof_begin(c));
            RCHK(launch_persist(c, &a, piece, grid));
            RCHK(prof_end(c));
            if (c->profiling) c->kern_launches++;
            hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(1), 0, c->stream, c->d_step, (int64_t)piece);
            HIPCHK(hipGetLastError());
        }
        done += piece;
    }
    return INFOLI_OK;
}

/* Enqueue the copy of the records of steps [t, t + count) from the device
 * ring into pinned staging buffer `buf` (of two), after the steps' kernels
 * on the same stream; returns the number of records. */
template <typename T>
int64_t enqueue_records(infoli_ctx *c, int64_t t, int64_t count, int64_t stride, int buf, int64_t per_buf)
{
    const int64_t first = (t + stride - 1) / stride, last = (t + count - 1) / stride;
    if (last < first) return 0;
    const int64_t nrec = last - first + 1, n = c->n;
    T *dst = (T *)c->h_stage + (size_t)buf * per_buf * n;
    for (int64_t r = 0; r < nrec;) {
        const int64_t slot = (first + r) % c->trace_cap;
        const int64_t run = std::min(nrec - r, c->trace_cap - slot);
        if (hipMemcpyAsync(dst + r * n, (T *)c->d_trace + slot * n, sizeof(T) * run * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess)
            return infoli_fail(INFOLI_ERR_HIP, "trace copy failed");
        r += run;
    }
    if (hipEventRecord(c->ev_copy[buf], c->stream) != hipSuccess) return infoli_fail(INFOLI_ERR_HIP, "event");
    return nrec;
}

/* Wait for staging buffer `buf` and widen its records into the caller's trace. */
template <typename T>
int drain_records(infoli_ctx *c, int buf, int64_t per_buf, int64_t nrec, double *trace, int64_t *rec)
{
    if (nrec <= 0) return INFOLI_OK;
    HIPCHK(hipEventSynchronize(c->ev_copy[buf]));
    const int64_t n = c->n;
    const T *s = (const T *)c->h_stage + (size_t)buf * per_buf * n;
    double *dst = trace + (*rec) * n;
    for (int64_t i = 0; i < nrec * n; i++) dst[i] = (double)s[i];
    *rec += nrec;
    return INFOLI_OK;
}

/* The device trace ring and the pinned staging for an advance of n_steps
 * with a record every stride_eff steps (0: no trace); *per_piece = records
 * per piece (the pieces' copies and widening overlap the next piece).  A
 * call that records at most one step needs one slot; any other gets a whole
 * piece of kTracePieceSteps steps' records, so calls of different lengths
 * share one ring and the graph captured with it.  The ring only grows. */
template <typename T>
int prepare_trace(infoli_ctx *c, int64_t n_steps, int64_t stride_eff, int64_t *per_piece)
{
    *per_piece = 1;
    if (stride_eff <= 0) return INFOLI_OK;
    const int64_t n = c->n;
    const int64_t want = infoli_trace_records(c, n_steps, stride_eff);
    const int64_t budget = std::max<int64_t>(1, (int64_t)(kTraceRingBytes / ((size_t)n * sizeof(T))));
    const int64_t piece = std::max<int64_t>(1, (kTracePieceSteps + stride_eff - 1) / stride_eff);
    const int64_t cap = want <= 1 ? 1 : std::min(piece, budget);
    RCHK(ensure_stage(c, (size_t)2 * cap * n * sizeof(T)));
    if (cap > c->trace_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->d_trace) HIPCHK(hipFree(c->d_trace));
        c->d_trace = nullptr;
        c->trace_cap = 0;
        HIPCHK(hipMalloc(&c->d_trace, (size_t)cap * n * sizeof(T)));
        c->trace_cap = cap;
        drop_graph(c);
    }
    *per_piece = cap;
    return INFOLI_OK;
}

template <typename T>
int prepare_impl(infoli_ctx *c, int64_t n_steps, int64_t stride_eff)
{
    int64_t per_piece = 1;
    RCHK(prepare_trace<T>(c, n_steps, stride_eff, &per_piece));
    const int64_t first = stride_eff > 0 ? std::min<int64_t>(n_steps, per_piece * stride_eff) : n_steps;
    return prepare_graph<T>(c, make_args<T>(c, stride_eff), first);
}

template <typename T>
int advance_impl(infoli_ctx *c, int64_t n_steps, int64_t stride, double *trace)
{
    const int64_t n = c->n;
    const int64_t stride_eff = trace ? stride : 0;
    int64_t per_piece = 1;
    RCHK(prepare_trace<T>(c, n_steps, stride_eff, &per_piece));
    StepArgs<T> a = make_args<T>(c, stride_eff);
    c->kern_ms = 0;
    c->kern_launches = 0;
    c->n_replays = c->n_direct = c->n_captures = 0;
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.clear();

    // k_begin sets the device step counter and clears the status, unless the
    // last advance left both so (it ran on the counter and ended clean): one
    // launch less in front of the steps of back-to-back advances
    const bool need_begin = !c->begin_done;
    c->begin_done = false;
    if (need_begin) {
        hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, c->stream, c->d_step, c->step, c->d_status);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int64_t t = c->step, left = n_steps, rec = 0;
    int64_t pending = 0;   // records of the previous piece still in staging buffer (piece - 1) & 1
    for (int piece = 0; left > 0; piece++) {
        int64_t chunk = left;
        if (stride_eff > 0) chunk = std::min<int64_t>(chunk, per_piece * stride_eff);
        RCHK(run_steps<T>(c, a, t, chunk));
        if (stride_eff > 0) {
            const int64_t got = enqueue_records<T>(c, t, chunk, stride_eff, piece & 1, per_piece);
            if (got < 0) return (int)got;
            // the GPU runs this piece while the previous one is widened
            RCHK(drain_records<T>(c, (piece - 1) & 1, per_piece, pending, trace, &rec));
            pending = got;
        }
        t += chunk;
        left -= chunk;
        c->step = t;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    if (stride_eff > 0) {
        const int last_buf = (int)(((n_steps + per_piece * stride_eff - 1) / (per_piece * stride_eff) - 1) & 1);
        RCHK(drain_records<T>(c, last_buf, per_piece, pending, trace, &rec));
    }
    hipLaunchKernelGGL(k_check_finite<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (const T *)c->d_vd[c->step & 1], (const T *)c->d_vs[c->step & 1],
                       (const T *)c->d_va[c->step & 1], n, c->d_flag);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_status, c->d_status, sizeof(AdvanceStatus), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int flag = c->h_status->nonfinite;
    const GridBarrier gb = c->h_status->bar;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (c->profiling) {
        for (size_t i = 0; i + 1 < c->prof_ev.size(); i += 2) {
            float e = 0;
            HIPCHK(hipEventElapsedTime(&e, c->prof_ev[i], c->prof_ev[i + 1]));
            c->kern_ms += e;
        }
    }
    if (gb.error) return infoli_fail(INFOLI_ERR_BARRIER, "persistent kernel grid barrier timed out");
    if (flag) return infoli_fail(INFOLI_ERR_NONFINITE, "a voltage is NaN or Inf after step %lld", (long long)c->step);
    // steps launched one by one (launch mode STEP, or graph mode while
    // profiling) take their index as an argument and leave the counter behind
    const bool on_counter = c->exec.mode != INFOLI_EXEC_STEP && !(c->profiling && c->exec.mode == INFOLI_EXEC_GRAPH);
    c->begin_done = on_counter && gb.count == 0;
    return INFOLI_OK;
}

}  // namespace

extern "C" int infoli_advance(infoli_t *c, int64_t n_steps, int64_t trace_stride, double *trace)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_advance"));
    if (n_steps < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_advance: n_steps < 0");
    if (trace && trace_stride <= 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_advance: trace_stride must be > 0");
    if (n_steps == 0) return INFOLI_OK;
    HIPCHK(hipSetDevice(c->device));
    RCHK(ensure_layout(c));
    return is_f32(c) ? advance_impl<float>(c, n_steps, trace_stride, trace)
                     : advance_impl<double>(c, n_steps, trace_stride, trace);
}

extern "C" int infoli_prepare(infoli_t *c, int64_t n_steps, int64_t trace_stride)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_prepare"));
    if (n_steps < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_prepare: n_steps < 0");
    if (trace_stride < 0) return infoli_fail(INFOLI_ERR_ARG, "infoli_prepare: trace_stride < 0");
    if (n_steps == 0) return INFOLI_OK;
    HIPCHK(hipSetDevice(c->device));
    RCHK(ensure_layout(c));
    return is_f32(c) ? prepare_impl<float>(c, n_steps, trace_stride) : prepare_impl<double>(c, n_steps, trace_stride);
}

namespace {

/* the gap-current kernel of the context's layout, writing n values of the
 * context's precision to device buffer `out` */
int launch_gap_currents(infoli_ctx *c, void *out)
{
    const int64_t nblk = c->npad / kBlock;
    const void *fn = hyb_active(c) ? k_gap_hyb_tab[lidx(c->exec.lanes_per_cell)]
                                   : k_gap_tab[c->prec][lidx(c->exec.lanes_per_cell)];
    int64_t t = c->step;
    hipError_t e;
    if (is_f32(c)) {
        StepArgs<float> a = make_args<float>(c, 0);
        void *kargs[] = {&a, &t, &out};
        e = hipLaunchKernel(fn, dim3((unsigned)nblk), dim3(kBlock), kargs, 0, c->stream);
    } else {
        StepArgs<double> a = make_args<double>(c, 0);
        void *kargs[] = {&a, &t, &out};
        e = hipLaunchKernel(fn, dim3((unsigned)nblk), dim3(kBlock), kargs, 0, c->stream);
    }
    if (e != hipSuccess) return infoli_fail(INFOLI_ERR_HIP, "gap-current kernel launch failed: %s", hipGetErrorString(e));
    return INFOLI_OK;
}

}  // namespace

extern "C" int infoli_gap_currents(infoli_t *c, double *i_c)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_gap_currents"));
    if (!i_c) return infoli_fail(INFOLI_ERR_ARG, "infoli_gap_currents: null output");
    HIPCHK(hipSetDevice(c->device));
    RCHK(ensure_layout(c));
    void *out = nullptr;
    HIPCHK(hipMalloc(&out, (size_t)c->n * c->tsize));
    int rc = launch_gap_currents(c, out);
    if (rc == INFOLI_OK) rc = download_values(c, out, i_c, c->n);
    (void)hipFree(out);
    return rc;
}

extern "C" int infoli_gap_currents_noalloc(infoli_t *c, double *i_c)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_gap_currents_noalloc"));
    if (!i_c) return infoli_fail(INFOLI_ERR_ARG, "infoli_gap_currents_noalloc: null output");
    const size_t bytes = (size_t)c->n * c->tsize;
    if (bytes > (size_t)INFOLI_NOALLOC_SCRATCH_BYTES)
        return infoli_fail(INFOLI_ERR_ARG,
                           "infoli_gap_currents_noalloc: %lld cells need %zu bytes of scratch, the static "
                           "buffer holds %zu (%zu cells in this precision); use infoli_gap_currents",
                           (long long)c->n, bytes, (size_t)INFOLI_NOALLOC_SCRATCH_BYTES,
                           (size_t)INFOLI_NOALLOC_SCRATCH_BYTES / c->tsize);
    void *out = c->device < kMaxDevices ? g_scratch_ptr[c->device] : nullptr;
    if (!out) return infoli_fail(INFOLI_ERR_STATE, "infoli_gap_currents_noalloc: no scratch on device %d", c->device);
    // the sliced-ELL lists the kernel reads must already be on the device:
    // building them would allocate
    if (c->sell_L != c->exec.lanes_per_cell)
        return infoli_fail(INFOLI_ERR_STATE, "infoli_gap_currents_noalloc: the neighbour lists of the current "
                                             "strategy are not built (infoli_advance or infoli_gap_currents "
                                             "builds them)");
    HIPCHK(hipSetDevice(c->device));
    RCHK(launch_gap_currents(c, out));
    // straight into the caller's array: no staging buffer
    HIPCHK(hipMemcpyAsync(i_c, out, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (is_f32(c)) {
        // widen in place from the end: value i's double overwrites floats
        // 2i and 2i + 1, both already widened
        unsigned char *b = reinterpret_cast<unsigned char *>(i_c);
        for (int64_t i = c->n - 1; i >= 0; i--) {
            float v;
            memcpy(&v, b + (size_t)i * sizeof(float), sizeof v);
            const double d = (double)v;
            memcpy(b + (size_t)i * sizeof(double), &d, sizeof d);
        }
    }
    return INFOLI_OK;
}

extern "C" int infoli_last_timing(const infoli_t *c, double *advance_ms, double *kernel_ms,
                                  int64_t *kernel_launches)
{
    RCHK(validate_ctx(c, "infoli_last_timing"));
    if (advance_ms) *advance_ms = c->last_ms;
    if (kernel_ms) *kernel_ms = c->kern_ms;
    if (kernel_launches) *kernel_launches = c->kern_launches;
    return INFOLI_OK;
}

extern "C" int infoli_last_launches(const infoli_t *c, int64_t *graph_replays, int64_t *direct_steps,
                                    int64_t *captures)
{
    RCHK(validate_ctx(c, "infoli_last_launches"));
    if (graph_replays) *graph_replays = c->n_replays;
    if (direct_steps) *direct_steps = c->n_direct;
    if (captures) *captures = c->n_captures;
    return INFOLI_OK;
}

extern "C" int infoli_synchronize(infoli_t *c)
{
    RCHK(validate_ctx(c, "infoli_synchronize"));
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* autotuner                                                                 */
/* ------------------------------------------------------------------------ */

extern "C" int infoli_autotune(infoli_t *c, int64_t trial_steps, infoli_exec *chosen)
{
    infoli_clear_error();
    RCHK(validate_ctx(c, "infoli_autotune"));
    if (trial_steps < 1) return infoli_fail(INFOLI_ERR_ARG, "infoli_autotune: trial_steps < 1");
    HIPCHK(hipSetDevice(c->device));
    // save the full state so the trials leave no trace
    const size_t plane = (size_t)c->npad * c->tsize;
    std::vector<unsigned char> saved(plane * kNV);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < kNV; v++)
        RCHK(d2h(c, saved.data() + plane * v, var_ptr(c, v), plane));
    const int64_t step0 = c->step;
    const infoli_exec orig = c->exec;
    const bool prof = c->profiling;
    c->profiling = false;

    // candidates: the heuristic's choice first, then tiled shapes around it
    // (lanes, receivers per tile and chunk size halved and doubled, both
    // compartment placements), swept tiles, whole networks per workgroup,
    // then the other gathers unless some cell needs the FP64 sums only the
    // tiled gathers have
    std::vector<infoli_exec> cands;
    const int spl = (int)std::min<int64_t>(trial_steps, 500);
    default_exec(c);
    infoli_exec h = c->exec;
    c->exec = orig;
    h.steps_per_launch = spl;
    cands.push_back(h);
    const int L0 = h.gather == GATHER_TILED ? h.lanes_per_cell : 1;
    const int R0 = h.gather == GATHER_TILED ? h.tile_cells : 2048;
    const int kb0 = h.gather == GATHER_TILED ? h.lds_kb : 48;
    for (int L : {L0 / 2, L0, L0 * 2}) {
        if (lidx(L) < 0) continue;
        for (int R : {R0 / 2, R0, R0 * 2}) {
            if (R < 64 || R > 16384 || R % 64) continue;
            const int kbs[3] = {std::max(1, kb0 / 2), kb0, std::min(156, kb0 * 2)};
            for (int i = 0; i < 3; i++) {
                const int kb = kbs[i];
                if (kb > 156 || (i > 0 && kb == kbs[i - 1])) continue;
                for (int split : {1, 2}) {
                    const infoli_exec e{INFOLI_EXEC_GRAPH, L, 1024, spl, GATHER_TILED, kb, R, split, 0};
                    if (memcmp(&e, &h, sizeof e) != 0) cands.push_back(e);
                }
            }
        }
    }
    // swept tiles, one lane per receiver, about one block per CU (and half
    // and twice that many) with the largest chunk that fits beside the
    // blocks' sums, for networks of at least 16 such blocks
    {
        const int Rs0 = h.gather == GATHER_SWEEP
                            ? h.tile_cells
                            : (int)std::min<int64_t>(16384, round_up((c->n + c->num_cus - 1) /
                                                                         std::max(c->num_cus, 1), 64));
        for (int L : {1}) {
            for (int R : {Rs0 / 2, Rs0, Rs0 * 2}) {
                if (R < 64 || R > 16384 || R % 64 || c->n < (int64_t)16 * R) continue;
                const int kb = sweep_max_kb(c, R);
                if (kb < 1) continue;
                for (int split : {0, 2}) {
                    const infoli_exec e{INFOLI_EXEC_GRAPH, L, 1024, spl, GATHER_SWEEP, kb, R, split, 0};
                    if (memcmp(&e, &h, sizeof e) != 0) cands.push_back(e);
                }
            }
        }
    }
    // whole networks per workgroup, when every piece of the network fits one
    if (!hyb_active(c) && !independent_cuts(c, nets_cap(c)).empty()) {
        const infoli_exec e{INFOLI_EXEC_PERSISTENT, 1, kBlock, spl, GATHER_NETS, kDefaultLdsKb, 1024, 1, 0};
        if (memcmp(&e, &h, sizeof e) != 0) cands.push_back(e);
    }
    if (!hyb_active(c)) {
        const int modes[] = {INFOLI_EXEC_GRAPH, INFOLI_EXEC_PERSISTENT};
        for (int L : {1, 2, 4, 8, 16})
            for (int mode : modes) {
                if (mode == INFOLI_EXEC_PERSISTENT && L != 4) continue;
                cands.push_back(infoli_exec{mode, L, kBlock, spl, GATHER_DIRECT, kDefaultLdsKb, 1024, 1, 0});
                if ((size_t)c->n * c->tsize <= kVldsMaxBytes)
                    cands.push_back(infoli_exec{mode, L, kBlock, spl, GATHER_VLDS, kDefaultLdsKb, 1024, 1, 0});
                for (int kb : {40, 76, 150})
                    if (max_block_entries_fit(c, kb))
                        cands.push_back(infoli_exec{mode, L, kBlock, spl, GATHER_STAGED, kb, 1024, 1, 0});
            }
    }
    double best = 1e300;
    infoli_exec best_e = orig;
    int rc = INFOLI_OK;
    for (infoli_exec e : cands) {
        if (check_exec(c, e) != INFOLI_OK) continue;
        c->exec = e;
        drop_graph(c);
        rc = ensure_layout(c);
        if (rc) break;
        // warm-up launch (graph capture, code load), then the timed trial
        int r1 = infoli_advance(c, e.steps_per_launch, 0, nullptr);
        int r2 = (r1 == INFOLI_OK || r1 == INFOLI_ERR_NONFINITE) ? infoli_advance(c, trial_steps, 0, nullptr) : r1;
        if ((r2 == INFOLI_OK || r2 == INFOLI_ERR_NONFINITE) && c->last_ms < best) {
            best = c->last_ms;
            best_e = e;
        }
        if (r1 == INFOLI_ERR_BARRIER || r2 == INFOLI_ERR_BARRIER) {
            rc = INFOLI_ERR_BARRIER;
            break;
        }
    }
    // restore
    c->exec = best_e;
    drop_graph(c);
    c->step = step0;
    c->begin_done = false;
    for (int v = 0; v < kNV; v++)
        RCHK(h2d(c, var_ptr(c, v), saved.data() + plane * v, plane));
    c->profiling = prof;
    RCHK(ensure_layout(c));
    if (rc != INFOLI_OK) return infoli_fail(rc, "infoli_autotune: a trial failed (%s)", infoli_strerror(rc));
    infoli_clear_error();   // the messages of candidates that did not apply
    if (chosen) *chosen = best_e;
    return INFOLI_OK;
}

/* ------------------------------------------------------------------------ */
/* host check of the tiled layout                                            */
/* ------------------------------------------------------------------------ */

namespace {

/* The tables of build_tiled against the network, and one step's scratch
 * traffic as the kernels index it: which partial-sum and dendrite slots the
 * first launch writes (k_tile_gj with its soma/axon workgroups, or
 * k_tile_sweep) and which the second reads (k_tile_ode, k_tile_ode3 or
 * k_tile_dend). */
template <typename T>
int check_tiled(const CsrView &g, size_t tsize, size_t psz, int gather, int L, int R, int kb, int split, int block,
                int64_t nsa)
{
    TiledTables<T> tb;
    RCHK(build_tiled<T>(g, tsize, L, R, kb, tb));
    const int64_t n = g.n, npad = round_up(n, kBlock), nnz = g.row_ptr[n];
    const int CPW = 64 / L;
    const int64_t ntiles = (int64_t)tb.tiles.size(), nrb = tb.nrb, C = tb.C;
    const bool any = nnz > 0;
    if ((int64_t)tb.rbt0.size() != nrb + 1 || tb.rbt0[nrb] != (any ? ntiles : 0))
        return infoli_fail(INFOLI_ERR_STATE, "layout: block table does not end at the tile count");
    if (gather == GATHER_SWEEP && sweep_lds_bytes_of(R, C, tsize, psz) > kLdsBytes)
        return infoli_fail(INFOLI_ERR_STATE, "layout: swept sums and chunk exceed the LDS");
    int64_t covered = 0;
    for (int64_t rb = 0; rb < nrb; rb++) {
        int64_t prev_end = -1;
        for (int32_t ti = tb.rbt0[rb]; ti < tb.rbt0[rb + 1]; ti++) {
            const TileDesc &td = tb.tiles[ti];
            if (td.rb != rb) return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d lists block %d, not %lld", ti, td.rb, (long long)rb);
            if (td.nsl * CPW != R) return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d has %d slices", ti, td.nsl);
            if (td.s0 < 0 || td.s0 % 64 || td.cn < 1 || td.s0 + td.cn > n || td.cn > C)
                return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d stages senders [%lld, %lld) of %lld (chunk %lld)", ti,
                                   (long long)td.s0, (long long)(td.s0 + td.cn), (long long)n, (long long)C);
            // sender ranges ascend within the block: partials merge in ascending sender order
            if (td.s0 < prev_end) return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d's senders overlap or precede the last tile's", ti);
            prev_end = td.s0 + td.cn;
            if (td.rl0 < 0 || td.rl0 + R > (int64_t)tb.rlist.size() || td.sl0 < 0 ||
                td.sl0 + td.nsl + 1 > (int64_t)tb.tsp.size())
                return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d indexes past its tables", ti);
            std::vector<char> seen(R, 0);
            for (int k = 0; k < R; k++) {
                const int r = tb.rlist[td.rl0 + k];
                if (r >= R || seen[r]) return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d's receiver list is not a permutation", ti);
                seen[r] = 1;
                // the kernels take a slice's fewest valid slots from its last row
                if (k > 0 && tb.rdeg[td.rl0 + k] > tb.rdeg[td.rl0 + k - 1])
                    return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d's receivers are not by degree, descending", ti);
            }
            for (int sl = 0; sl < td.nsl; sl++) {
                const int64_t base = tb.tsp[td.sl0 + sl], span = tb.tsp[td.sl0 + sl + 1] - base;
                if (span < 0 || span % 64 || base + span > (int64_t)tb.tcol.size())
                    return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d slice %d has a bad extent", ti, sl);
                const int W = (int)(span / 64);
                int md = 0;
                for (int cw = 0; cw < CPW; cw++) {
                    const int64_t pos = td.rl0 + (int64_t)sl * CPW + cw;
                    const int r = tb.rlist[pos], d = tb.rdeg[pos];
                    md = std::max(md, d);
                    const int64_t cell = rb * R + r;
                    // the row's junctions from this tile's senders, in row order
                    std::vector<int64_t> es;
                    if (cell < n)
                        for (int64_t e = g.row_ptr[cell]; e < g.row_ptr[cell + 1]; e++)
                            if (g.col[e] >= td.s0 && g.col[e] < td.s0 + td.cn) es.push_back(e);
                    if ((int64_t)es.size() != d)
                        return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d receiver %lld lists %d junctions, the network %zu", ti,
                                           (long long)cell, d, es.size());
                    for (int p = 0; p < W * L; p++) {
                        const int64_t slot = base + (int64_t)(p / L) * 64 + cw * L + (p % L);
                        // every slot is read (padding too: its term is dropped), so every id stays in the chunk
                        if (tb.tcol[slot] >= td.cn)
                            return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d slot %lld reads sender %d of a %d-cell chunk", ti,
                                               (long long)slot, tb.tcol[slot], td.cn);
                        if (p < d) {
                            const int64_t e = es[p];
                            if (td.s0 + tb.tcol[slot] != g.col[e] || tb.tw[slot] != (T)g.w[e] ||
                                (td.wconst && (T)g.w[e] != (T)td.wc))
                                return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d receiver %lld entry %d differs from the network",
                                                   ti, (long long)cell, p);
                        }
                    }
                    covered += d;
                }
                if (W != (md + L - 1) / L) return infoli_fail(INFOLI_ERR_STATE, "layout: tile %d slice %d is %d wide for degree %d", ti, sl, W, md);
            }
        }
    }
    if (covered != nnz)
        return infoli_fail(INFOLI_ERR_STATE, "layout: the tiles hold %lld junctions, the network %lld", (long long)covered, (long long)nnz);

    // one step's scratch traffic
    const int64_t slots = (int64_t)part_slots((size_t)ntiles, nrb, R);
    std::vector<char> pw(slots, 0), dw(2 * npad, 0);
    std::vector<int> soma(npad, 0), axon(npad, 0);
    const auto wpart = [&](int64_t i) -> bool {
        if (i < 0 || i >= slots) return false;
        pw[i] = 1;
        return true;
    };
    if (gather == GATHER_TILED) {
        if (any)
            for (int64_t ti = 0; ti < ntiles; ti++)
                for (int k = 0; k < R; k++)
                    if (!wpart(ti * R + tb.rlist[tb.tiles[ti].rl0 + k]))
                        return infoli_fail(INFOLI_ERR_STATE, "layout: tile %lld writes past the partials", (long long)ti);
        if (split == 2) {
            // the workgroups past the tiles (k_tile_gj): group g of NT/2 cells,
            // workgroup (tiles + j) takes groups j, j + nsa, ...
            for (int64_t j = 0; j < nsa; j++)
                for (int64_t gi = j; gi * (block / 2) < n; gi += nsa)
                    for (int wv = 0; wv < block / 64; wv++)
                        for (int lane = 0; lane < 64; lane++) {
                            const int64_t cell = (gi * (block / 128) + (wv >> 1)) * 64 + lane;
                            if (cell >= n) continue;
                            if (cell >= npad) return infoli_fail(INFOLI_ERR_STATE, "layout: cell %lld past the planes", (long long)cell);
                            if (wv & 1) {
                                axon[cell]++;
                                dw[cell] = dw[npad + cell] = 1;
                            } else {
                                soma[cell]++;
                            }
                        }
            for (int64_t i = 0; i < n; i++)
                if (soma[i] != 1 || axon[i] != 1)
                    return infoli_fail(INFOLI_ERR_STATE, "layout: cell %lld's soma runs %d times, its axon %d (nsa %lld)",
                                       (long long)i, soma[i], axon[i], (long long)nsa);
        }
    } else if (split >= 1) {
        for (int64_t rb = 0; rb < nrb; rb++)
            for (int r = 0; r < R; r++)
                if (!wpart(rb * R + r)) return infoli_fail(INFOLI_ERR_STATE, "layout: block %lld writes past the partials", (long long)rb);
        if (split == 2) {
            // k_tile_beside: wave 0 of workgroup b writes the dendrite sums
            // of cells 64 b .. 64 b + 63
            for (int64_t b = 0; b < npad / 64; b++)
                for (int lane = 0; lane < 64; lane++) {
                    const int64_t cell = b * 64 + lane;
                    if (cell < n) dw[cell] = dw[npad + cell] = 1;
                }
        }
    }
    const bool reads_parts = gather == GATHER_TILED || split >= 1;
    for (int64_t cell = 0; cell < n && reads_parts; cell++) {
        const int64_t rb = cell / R, rloc = cell - rb * R;
        const int64_t t0 = gather == GATHER_TILED ? (any ? tb.rbt0[rb] : 0) : rb;
        const int64_t t1 = gather == GATHER_TILED ? (any ? tb.rbt0[rb + 1] : 0) : rb + 1;
        for (int64_t tt = t0; tt < t1; tt++) {
            const int64_t i = tt * R + rloc;
            if (i >= slots || !pw[i])
                return infoli_fail(INFOLI_ERR_STATE, "layout: cell %lld reads partial %lld before any kernel of the step writes it",
                                   (long long)cell, (long long)i);
        }
        if (split == 2 && (!dw[cell] || !dw[npad + cell]))
            return infoli_fail(INFOLI_ERR_STATE, "layout: cell %lld reads its dendrite sums before they are written", (long long)cell);
    }
    return INFOLI_OK;
}

}  // namespace

extern "C" int infoli_tiled_layout_check(int64_t n, const int64_t *row_ptr, const int32_t *col, const double *w,
                                         int precision, int gather, int lanes_per_cell, int tile_cells, int lds_kb,
                                         int ode_split, int block, int64_t nsa)
{
    infoli_clear_error();
    if (n <= 0 || n > INT32_MAX || !row_ptr || row_ptr[0] != 0)
        return infoli_fail(INFOLI_ERR_ARG, "infoli_tiled_layout_check: bad network");
    if (precision < INFOLI_FP64 || precision > INFOLI_FP32_KAHAN || (gather != GATHER_TILED && gather != GATHER_SWEEP) ||
        lidx(lanes_per_cell) < 0 || tile_cells < 64 || tile_cells > 16384 || tile_cells % 64 || lds_kb < 1 ||
        lds_kb > 156 || ode_split < 0 || ode_split > 2 ||
        (block != 256 && block != 512 && block != 1024) || (gather == GATHER_SWEEP && block != 1024) ||
        (ode_split == 2 && nsa < 1))
        return infoli_fail(INFOLI_ERR_ARG, "infoli_tiled_layout_check: bad shape");
    for (int64_t i = 0; i < n; i++)
        if (row_ptr[i + 1] < row_ptr[i]) return infoli_fail(INFOLI_ERR_ARG, "infoli_tiled_layout_check: row_ptr decreases");
    const int64_t nnz = row_ptr[n];
    if (nnz > 0 && (!col || !w)) return infoli_fail(INFOLI_ERR_ARG, "infoli_tiled_layout_check: null col or w");
    for (int64_t e = 0; e < nnz; e++)
        if (col[e] < 0 || col[e] >= n) return infoli_fail(INFOLI_ERR_ARG, "infoli_tiled_layout_check: col[%lld] out of range", (long long)e);
    const CsrView g{n, row_ptr, col, w};
    const size_t tsize = precision == INFOLI_FP64 ? 8 : 4;
    const size_t psz = precision == INFOLI_FP32_KAHAN ? 8 : tsize;
    return precision == INFOLI_FP64
               ? check_tiled<double>(g, tsize, psz, gather, lanes_per_cell, tile_cells, lds_kb, ode_split, block, nsa)
               : check_tiled<float>(g, tsize, psz, gather, lanes_per_cell, tile_cells, lds_kb, ode_split, block, nsa);
}
