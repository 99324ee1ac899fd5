/*
 * infoli_kernels.h -- device code of the InfOli network step for gfx950.
 *
 * A step is the gap-junction currents (src/openmp/infoli.c:496-520) then the
 * three-compartment update (src/openmp/infoli.c:548-679; FP64 twin
 * src/single_threaded/infoli.c:371-858).  V_dend, V_soma and V_axon live in
 * ping-pong buffers indexed by the step's parity, so every compartment reads
 * the start-of-step voltages whichever kernel or wavefront runs it.
 *
 * Two families of kernels:
 *  - fused step kernels (k_step, k_persist): a 256-thread workgroup owns a
 *    block of cells, sums their neighbour lists (sliced ELL: entry p of cell
 *    c in slice s at slice_ptr[s] + (p / L) * 64 + c * L + (p % L), so a load
 *    instruction of a wavefront reads 64 consecutive entries) from V_dend in
 *    memory, in LDS, or from a staged per-block copy, then updates the block;
 *  - the tiled path: k_tile_gj sums the junctions of one receiver block from
 *    one sender chunk (the chunk's V_dend in LDS, 16-bit chunk-local ids)
 *    into per-tile partial sums, and a compartment kernel (k_tile_ode3,
 *    k_tile_ode or k_tile_dend) merges them in chunk order and updates the
 *    cells; somata and axons can also run in k_tile_gj beside the tiles.
 *    k_tile_sweep walks the same tiles with one workgroup per receiver
 *    block and merges the partials in LDS instead.
 * Each row's terms are summed in list (ascending sender) order in every path.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace infoli {

constexpr int kBlock = 256;          // threads per workgroup = cells per workgroup
constexpr int kNV = 19;              // per-cell variables (default.conf 24-42 order)

enum Var {
    V_DEND = 0, CALCIUM_R, POTASSIUM_S, HCURRENT_Q, CA2PLUS, I_CAH, G_CAL,
    V_SOMA, SODIUM_M, SODIUM_H, POTASSIUM_N, POTASSIUM_P, POTASSIUM_X_S,
    CALCIUM_K, CALCIUM_L, V_AXON, SODIUM_M_A, SODIUM_H_A, POTASSIUM_X_A
};

/* Model constants in the step's arithmetic.  The four coupling
 * conductances are the reference's G_INT/(1-P1), G_INT/P1, G_INT/(1-P2)
 * and G_INT/P2, evaluated once on the host in the same precision. */
template <typename T>
struct Model {
    T DELTA, C_M, G_NA_S, G_KDR_S, G_K_S, G_LS, G_K_CA, G_CAH, G_LD, G_H,
      G_NA_A, G_K_A, G_LA, V_NA, V_K, V_CA, V_H, V_L;
    T G_sd, G_ds, G_as, G_sa;
};

/* One tile of the 2D-tiled gap-junction kernel: the junctions from sender
 * chunk `ch` into receiver block `rb`. */
struct TileDesc {
    int32_t rb, ch, nsl;
    int32_t wconst;            // 1: every junction of the tile has conductance wc (tw not read)
    int64_t sl0;               // first slice in tsp
    int64_t rl0;               // first receiver slot in rlist / rdeg
    double wc;
    int64_t s0;                // senders [s0, s0 + cn) staged in LDS; ids are relative to s0
    int32_t cn;
    int32_t pad;
};

template <typename T>
struct StepArgs {
    int64_t n;            // cells
    int64_t npad;         // plane stride (multiple of kBlock)
    T *st;                // kNV planes of npad values
    T *vd[2];             // V_dend ping-pong: vd[t & 1] holds V_dend at the start of step t
    T *vs[2];             // V_soma ping-pong (the dendrite and axon read the old value)
    T *va[2];             // V_axon ping-pong
    const int32_t *deg;   // neighbour count per cell
    const int64_t *slice_ptr;
    const int32_t *col;   // sender ids, sliced-ELL
    const T *w;           // conductances, sliced-ELL
    const T *iapp_base;   // per-cell constant Iapp (nullable)
    const uint8_t *mask;  // pulse mask (nullable = every cell)
    T amp;
    int64_t t0, t1;       // pulse on steps [t0, t1)
    const T *tab;         // per-step Iapp rows (nullable)
    int64_t tab_t0, tab_n;
    T *trace;             // ring of trace_cap records of n values (nullable)
    int64_t stride, trace_cap;
    const int64_t *d_step;// device step counter (nullable: step index is absolute)
    // cell blocks: uniform 256-cell blocks (direct / whole-V-in-LDS gathers)
    // or the staged-gather table
    int64_t nblocks;
    const int32_t *blk_cell0;  // first cell of each block (multiple of 64)
    const int32_t *blk_ncell;  // cells per block (multiple of 64, <= 256)
    const int64_t *blk_gbase;  // gather-list start of each block, nblocks + 1
    const int32_t *lrs;        // LDS row start of each cell within its block
    const uint32_t *gl32;      // (sender << sb) | slot, sorted by sender within a block
    const uint64_t *gl64;      // (sender << 32) | slot, when the packed form does not fit
    int sb;                    // slot bits of gl32
    int stream_nt;             // neighbour lists larger than the caches: nontemporal loads
    // 2D-tiled gap junctions (receiver block x sender chunk)
    const TileDesc *tiles;     // tiles, receiver-block major, chunks ascending
    const int32_t *rb_tile0;   // first tile of each receiver block, nrb + 1
    const int32_t *rb_part0;   // first partial of each receiver block, nrb + 1: rb_tile0,
                               // or rb itself when one workgroup sweeps a block's tiles
    const int64_t *tsp;        // per tile slice: first entry (64 entries per column)
    const uint16_t *rlist;     // per tile: receivers (local to the block), in-chunk degree descending
    const uint16_t *rdeg;      // in-chunk degree of each listed receiver
    const uint16_t *tcol;      // sender index local to the chunk
    const T *tw;               // conductance
    void *part;                // partial sums, [tile][receiver block size] (sum, comp) pairs,
                               // or [block][receiver block size] when swept
    int32_t fp64_deg;          // FP32 compensated: cells with at least this many
                               // neighbours sum FP64 terms (0 = none)
    int32_t ntiles;            // tiles (workgroups past them run somata and axons)
    int32_t nsa;               // soma/axon workgroups (each loops over cell groups)
    int32_t fast_math;         // FP32 somata and axons of those workgroups with the
                               // hardware exp2 and reciprocal
    T *dpre;                   // ode_split 2: [2][npad] per cell the dendrite's current
                               // sum before I_c, and I_h (written beside the tiles)
    int32_t rcells;            // receivers per block (multiple of 64)
    int32_t sweep_ode;         // swept tiles: each workgroup updates its block's cells
                               // after their sums (no partials, no compartment kernel)
    Model<T> m;
};

/* how a step reads its neighbours' V_dend */
enum Gather {
    GATHER_DIRECT = 0,   // random loads from the V_dend array (L2-resident)
    GATHER_VLDS = 1,     // the whole V_dend array staged in LDS (small networks)
    GATHER_STAGED = 2,   // per block: senders sorted, copied to LDS, then read in row order
    GATHER_TILED = 3,    // receiver block x sender chunk tiles, chunk of V_dend in LDS,
                         // partial sums merged by the compartment kernel
    GATHER_SWEEP = 4,    // the tiled layout, one workgroup per receiver block walking its
                         // chunks, the partial sums merged in LDS
    GATHER_NETS = 5      // whole networks per workgroup, V_dend in LDS and the state in
                         // registers for many steps per launch (k_persist_nets)
};

struct GridBarrier {
    unsigned int count;
    unsigned int error;
};

/* What one infoli_advance reports back: the persistent kernels' barrier
 * (its timeout flag stays set until the next advance) and the NaN/Inf
 * watchdog's flag, copied to the host in one piece. */
struct AdvanceStatus {
    GridBarrier bar;
    int nonfinite;
    int pad;
};

/* ---------------------------------------------------------------------- */
/* compartment update, FP32: src/openmp/infoli.c:556-676                   */
/*                                                                          */
/* The three compartments are independent given the start-of-step voltages */
/* (the dendrite reads V_s, the soma V_d and V_a, the axon V_s, all old),   */
/* so each is a function of its own; `ode` runs all three.  Each returns    */
/* its new voltage, committed by the caller (OMP:673-675).                  */
/* ---------------------------------------------------------------------- */

struct CellState32 {
    float vd, r, s, q, ca, icah, gcal, vs, m, h, n, p, xs, k, l, va, ma, ha, xa;
};
struct CellState64 {
    double vd, r, s, q, ca, icah, gcal, vs, m, h, n, p, xs, k, l, va, ma, ha, xa;
};

__device__ __forceinline__ float fmin_ref(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ double fmin_ref(double a, double b) { return (a < b) ? a : b; }

/* FM = false: the reference's float arithmetic (libm-accurate expf, IEEE
 * division).  FM = true: the hardware exp2 (x * log2 e first) and division
 * as a multiply by the hardware reciprocal, each within ~2 ulp. */
template <bool FM>
__device__ __forceinline__ float E(float x)
{
    if constexpr (FM) return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    else return expf(x);
}

template <bool FM>
__device__ __forceinline__ float D(float a, float b)
{
    if constexpr (FM) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}

/* dendrite, OMP:561-593, up to its voltage: updates q, r, s, ca, icah and
 * returns the sum of the currents that precede I_c (I_h apart) */
template <bool FM = false>
__device__ __forceinline__ float ode_dend_pre(const Model<float> &P, CellState32 &c, float Vd, float Vs,
                                              float &I_h_out)
{
    // Dend H current OMP:561-564
    float q_inf = D<FM>(1, 1 + E<FM>((Vd + 80) / 4));
    float tau_q = D<FM>(1, E<FM>(-0.086f * Vd - 14.6f) + E<FM>(0.070f * Vd - 1.87f));
    float dq_dt = D<FM>(q_inf - c.q, tau_q);
    c.q = P.DELTA * dq_dt + c.q;
    // Dend Ca current OMP:567-572
    float alpha_r = D<FM>(1.7f, 1 + E<FM>(D<FM>(-(Vd - 5), 13.9f)));
    // 0/0 at Vd = -8.5 exactly (the reference computes NaN there): the limit 0.02 * 5
    const float xr = Vd + 8.5f;
    float beta_r = xr == 0 ? 0.1f : D<FM>(0.02f * xr, E<FM>(D<FM>(xr, 5)) - 1);
    float r_inf = D<FM>(alpha_r, alpha_r + beta_r);
    float tau_r = D<FM>(5, alpha_r + beta_r);
    float dr_dt = D<FM>(r_inf - c.r, tau_r);
    c.r = P.DELTA * dr_dt + c.r;
    // Dend K current OMP:575-580 (previous Ca2Plus)
    float alpha_s = fmin_ref((0.00002f * c.ca), 0.01f);
    float beta_s = 0.015f;
    float s_inf = D<FM>(alpha_s, alpha_s + beta_s);
    float tau_s = D<FM>(1, alpha_s + beta_s);
    float ds_dt = D<FM>(s_inf - c.s, tau_s);
    c.s = P.DELTA * ds_dt + c.s;
    // Dend Cal OMP:583-584 (previous I_CaH)
    float dCa_dt = -3 * c.icah - 0.075f * c.ca;
    c.ca = P.DELTA * dCa_dt + c.ca;
    // dendritic currents OMP:587-593 (new gates, old voltages)
    float I_sd = P.G_sd * (Vd - Vs);
    float I_CaH = P.G_CAH * (c.r * c.r) * (Vd - P.V_CA);
    float I_K_Ca = P.G_K_CA * c.s * (Vd - P.V_K);
    float I_ld = P.G_LD * (Vd - P.V_L);
    float I_h = P.G_H * c.q * (Vd - P.V_H);
    c.icah = I_CaH;
    I_h_out = I_h;
    // the reference sums left to right: ((((I_CaH + I_sd) + I_ld) + I_K_Ca) + I_c) + I_h,
    // so the first four terms form one partial sum and I_c joins later
    return I_CaH + I_sd + I_ld + I_K_Ca;
}

/* the dendrite voltage from the partial current sum of ode_dend_pre */
template <bool FM = false>
__device__ __forceinline__ float ode_dend_post(const Model<float> &P, float Vd, float pre, float I_h, float I_c,
                                               float iApp)
{
    float dVd_dt = D<FM>(-(pre + I_c + I_h) + iApp, P.C_M);
    return P.DELTA * dVd_dt + Vd;
}

template <bool FM = false>
__device__ __forceinline__ float ode_dend(const Model<float> &P, CellState32 &c, float Vd, float Vs,
                                          float I_c, float iApp)
{
    float I_h;
    const float pre = ode_dend_pre<FM>(P, c, Vd, Vs, I_h);
    return ode_dend_post<FM>(P, Vd, pre, I_h, I_c, iApp);
}

/* soma, OMP:599-642; updates k, l, m, h, n, p, xs */
template <bool FM = false>
__device__ __forceinline__ float ode_soma(const Model<float> &P, CellState32 &c, float Vd, float Vs, float Va)
{
    // Soma calcium OMP:599-606
    float k_inf = D<FM>(1, 1 + E<FM>(D<FM>(-1 * (Vs + 61), 4.2f)));
    float l_inf = D<FM>(1, 1 + E<FM>(D<FM>(Vs + 85.5f, 8.5f)));
    float tau_l = D<FM>(20 * E<FM>(D<FM>(Vs + 160, 30)), 1 + E<FM>(D<FM>(Vs + 84, 7.3f))) + 35;
    float dk_dt = (k_inf - c.k) / 1.0f;
    float dl_dt = D<FM>(l_inf - c.l, tau_l);
    c.k = P.DELTA * dk_dt + c.k;
    c.l = P.DELTA * dl_dt + c.l;
    // Soma sodium OMP:609-614
    float m_inf = D<FM>(1, 1 + E<FM>(D<FM>(-30 - Vs, 5.5f)));
    float h_inf = D<FM>(1, 1 + E<FM>(D<FM>(-70 - Vs, -5.8f)));
    float tau_h = 3 * E<FM>(D<FM>(-40 - Vs, 33));
    float dh_dt = D<FM>(h_inf - c.h, tau_h);
    c.m = m_inf;
    c.h = c.h + P.DELTA * dh_dt;
    // Soma potassium OMP:617-624
    float n_inf = D<FM>(1, 1 + E<FM>(D<FM>(-3 - Vs, 10)));
    float p_inf = D<FM>(1, 1 + E<FM>(D<FM>(-51 - Vs, -12)));
    float tau_n = 5 + (47 * E<FM>(D<FM>(-(-50 - Vs), 900)));
    float dn_dt = D<FM>(n_inf - c.n, tau_n);
    float dp_dt = D<FM>(p_inf - c.p, tau_n);
    c.n = P.DELTA * dn_dt + c.n;
    c.p = P.DELTA * dp_dt + c.p;
    // Soma potassium X OMP:627-632 (literal 0.05f)
    // 0/0 at Vs = -25 exactly (NaN in the reference): the limit 0.13 * 10
    const float xs = Vs + 25;
    float alpha_x_s = xs == 0 ? 1.3f : D<FM>(0.13f * xs, 1 - E<FM>(D<FM>(-xs, 10)));
    float beta_x_s = 1.69f * E<FM>(-0.0125f * (Vs + 35));
    float x_inf_s = D<FM>(alpha_x_s, alpha_x_s + beta_x_s);
    float tau_x_s = D<FM>(1, alpha_x_s + beta_x_s);
    float dx_dt_s = D<FM>(x_inf_s - c.xs, tau_x_s);
    c.xs = 0.05f * dx_dt_s + c.xs;
    // somatic currents OMP:635-642
    float I_ds = P.G_ds * (Vs - Vd);
    float k3 = c.k * c.k * c.k, m3 = c.m * c.m * c.m;
    float n2 = c.n * c.n, xs2 = c.xs * c.xs;
    float I_CaL = c.gcal * k3 * c.l * (Vs - P.V_CA);
    float I_Na_s = P.G_NA_S * m3 * c.h * (Vs - P.V_NA);
    float I_ls = P.G_LS * (Vs - P.V_L);
    float I_Kdr_s = P.G_KDR_S * (n2 * n2) * (Vs - P.V_K);
    float I_K_s = P.G_K_S * (xs2 * xs2) * (Vs - P.V_K);
    float I_as = P.G_as * (Vs - Va);
    float dVs_dt = D<FM>(-(I_CaL + I_ds + I_as + I_Na_s + I_ls + I_Kdr_s + I_K_s), P.C_M);
    return P.DELTA * dVs_dt + Vs;
}

/* axon, OMP:648-668; updates ma, ha, xa */
template <bool FM = false>
__device__ __forceinline__ float ode_axon(const Model<float> &P, CellState32 &c, float Vs, float Va)
{
    // Axon sodium OMP:648-653
    float m_inf_a = D<FM>(1, 1 + E<FM>(D<FM>(-30 - Va, 5.5f)));
    float h_inf_a = D<FM>(1, 1 + E<FM>(D<FM>(-60 - Va, -5.8f)));
    float tau_h_a = 1.5f * E<FM>(D<FM>(-40 - Va, 33));
    float dh_dt_a = D<FM>(h_inf_a - c.ha, tau_h_a);
    c.ma = m_inf_a;
    c.ha = c.ha + P.DELTA * dh_dt_a;
    // Axon potassium OMP:656-661
    const float xa = Va + 25;   // 0/0 at -25 exactly: the limit
    float alpha_x_a = xa == 0 ? 1.3f : D<FM>(0.13f * xa, 1 - E<FM>(D<FM>(-xa, 10)));
    float beta_x_a = 1.69f * E<FM>(-0.0125f * (Va + 35));
    float x_inf_a = D<FM>(alpha_x_a, alpha_x_a + beta_x_a);
    float tau_x_a = D<FM>(1, alpha_x_a + beta_x_a);
    float dx_dt_a = D<FM>(x_inf_a - c.xa, tau_x_a);
    c.xa = 0.05f * dx_dt_a + c.xa;
    // axonal currents OMP:664-668
    float ma3 = c.ma * c.ma * c.ma, xa2 = c.xa * c.xa;
    float I_Na_a = P.G_NA_A * ma3 * c.ha * (Va - P.V_NA);
    float I_la = P.G_LA * (Va - P.V_L);
    float I_sa = P.G_sa * (Va - Vs);
    float I_K_a = P.G_K_A * (xa2 * xa2) * (Va - P.V_K);
    float dVa_dt = D<FM>(-(I_K_a + I_sa + I_la + I_Na_a), P.C_M);
    return P.DELTA * dVa_dt + Va;
}

/* ---------------------------------------------------------------------- */
/* compartment update, FP64: src/single_threaded/infoli.c:371-858          */
/* ---------------------------------------------------------------------- */

/* CompDend ST:371-428, split as the FP32 form above (DendHCurr, DendCaCurr, DendKCurr, DendCal,
 * DendCurrVolt) */
__device__ __forceinline__ double ode_dend_pre(const Model<double> &P, CellState64 &c, double prevV_dend,
                                               double prevV_soma, double &I_h_out)
{
    // DendHCurr ST:430-447
    double q_inf = 1 / (1 + exp((prevV_dend + 80) / 4));
    double tau_q = 1 / (exp(-0.086 * prevV_dend - 14.6) + exp(0.070 * prevV_dend - 1.87));
    double dq_dt = (q_inf - c.q) / tau_q;
    double q = P.DELTA * dq_dt + c.q;
    // DendCaCurr ST:449-468
    double alpha_r = 1.7 / (1 + exp(-(prevV_dend - 5) / 13.9));
    const double xr = prevV_dend + 8.5;   // 0/0 at -8.5 exactly: the limit 0.02 * 5
    double beta_r = xr == 0 ? 0.1 : 0.02 * xr / (exp(xr / 5) - 1);
    double r_inf = alpha_r / (alpha_r + beta_r);
    double tau_r = 5 / (alpha_r + beta_r);
    double dr_dt = (r_inf - c.r) / tau_r;
    double r = P.DELTA * dr_dt + c.r;
    // DendKCurr ST:470-489
    double alpha_s = fmin_ref((0.00002 * c.ca), 0.01);
    double beta_s = 0.015;
    double s_inf = alpha_s / (alpha_s + beta_s);
    double tau_s = 1 / (alpha_s + beta_s);
    double ds_dt = (s_inf - c.s) / tau_s;
    double sK = P.DELTA * ds_dt + c.s;
    // DendCal ST:492-507
    double dCa_dt = -3 * c.icah - 0.075 * c.ca;
    double Ca = P.DELTA * dCa_dt + c.ca;
    // DendCurrVolt ST:509-542
    double I_sd = P.G_sd * (prevV_dend - prevV_soma);
    double I_CaH = P.G_CAH * r * r * (prevV_dend - P.V_CA);
    double I_K_Ca = P.G_K_CA * sK * (prevV_dend - P.V_K);
    double I_ld = P.G_LD * (prevV_dend - P.V_L);
    double I_h = P.G_H * q * (prevV_dend - P.V_H);
    c.q = q; c.r = r; c.s = sK; c.ca = Ca; c.icah = I_CaH;
    I_h_out = I_h;
    // left to right as the reference: ((((I_CaH + I_sd) + I_ld) + I_K_Ca) + I_c) + I_h
    return I_CaH + I_sd + I_ld + I_K_Ca;
}

__device__ __forceinline__ double ode_dend_post(const Model<double> &P, double prevV_dend, double pre, double I_h,
                                                double I_c, double I_app)
{
    double dVd_dt = (-(pre + I_c + I_h) + I_app) / P.C_M;
    return P.DELTA * dVd_dt + prevV_dend;
}

__device__ __forceinline__ double ode_dend(const Model<double> &P, CellState64 &c, double prevV_dend,
                                           double prevV_soma, double I_c, double I_app)
{
    double I_h;
    const double pre = ode_dend_pre(P, c, prevV_dend, prevV_soma, I_h);
    return ode_dend_post(P, prevV_dend, pre, I_h, I_c, I_app);
}

/* CompSoma ST:560-616 */
__device__ __forceinline__ double ode_soma(const Model<double> &P, CellState64 &c, double prevV_dend,
                                           double prevV_soma, double prevV_axon)
{
    // SomaCalcium ST:618-640
    double k_inf = (1 / (1 + exp(-1 * (prevV_soma + 61) / 4.2)));
    double l_inf = (1 / (1 + exp((prevV_soma + 85.5) / 8.5)));
    double tau_l = ((20 * exp((prevV_soma + 160) / 30) / (1 + exp((prevV_soma + 84) / 7.3))) + 35);
    double dk_dt = (k_inf - c.k) / 1.0;
    double dl_dt = (l_inf - c.l) / tau_l;
    double k = P.DELTA * dk_dt + c.k;
    double l = P.DELTA * dl_dt + c.l;
    // SomaSodium ST:642-663
    double m_inf = 1 / (1 + (exp((-30 - prevV_soma) / 5.5)));
    double h_inf = 1 / (1 + (exp((-70 - prevV_soma) / -5.8)));
    double tau_h = 3 * exp((-40 - prevV_soma) / 33);
    double dh_dt = (h_inf - c.h) / tau_h;
    double m = m_inf;
    double h = c.h + P.DELTA * dh_dt;
    // SomaPotassium ST:665-688
    double n_inf = 1 / (1 + exp((-3 - prevV_soma) / 10));
    double p_inf = 1 / (1 + exp((-51 - prevV_soma) / -12));
    double tau_n = 5 + (47 * exp(-(-50 - prevV_soma) / 900));
    double dn_dt = (n_inf - c.n) / tau_n;
    double dp_dt = (p_inf - c.p) / tau_n;
    double n = P.DELTA * dn_dt + c.n;
    double p = P.DELTA * dp_dt + c.p;
    // SomaPotassiumX ST:690-709 (literal 0.05)
    const double xs = prevV_soma + 25;    // 0/0 at -25 exactly: the limit 0.13 * 10
    double alpha_x_s = xs == 0 ? 1.3 : 0.13 * xs / (1 - exp(-xs / 10));
    double beta_x_s = 1.69 * exp(-0.0125 * (prevV_soma + 35));
    double x_inf_s = alpha_x_s / (alpha_x_s + beta_x_s);
    double tau_x_s = 1 / (alpha_x_s + beta_x_s);
    double dx_dt_s = (x_inf_s - c.xs) / tau_x_s;
    double x_s = 0.05 * dx_dt_s + c.xs;
    // SomaCurrVolt ST:711-749
    double I_ds = P.G_ds * (prevV_soma - prevV_dend);
    double I_CaL = c.gcal * k * k * k * l * (prevV_soma - P.V_CA);
    double I_Na_s = P.G_NA_S * m * m * m * h * (prevV_soma - P.V_NA);
    double I_ls = P.G_LS * (prevV_soma - P.V_L);
    double I_Kdr_s = P.G_KDR_S * n * n * n * n * (prevV_soma - P.V_K);
    double xs2 = x_s * x_s;
    double I_K_s = P.G_K_S * (xs2 * xs2) * (prevV_soma - P.V_K);
    double I_as = P.G_as * (prevV_soma - prevV_axon);
    double dVs_dt = (-(I_CaL + I_ds + I_as + I_Na_s + I_ls + I_Kdr_s + I_K_s)) / P.C_M;
    c.k = k; c.l = l; c.m = m; c.h = h; c.n = n; c.p = p; c.xs = x_s;
    return P.DELTA * dVs_dt + prevV_soma;
}

/* CompAxon ST:753-786 */
__device__ __forceinline__ double ode_axon(const Model<double> &P, CellState64 &c, double prevV_soma,
                                           double prevV_axon)
{
    // AxonSodium ST:788-810
    double m_inf_a = 1 / (1 + (exp((-30 - prevV_axon) / 5.5)));
    double h_inf_a = 1 / (1 + (exp((-60 - prevV_axon) / -5.8)));
    double tau_h_a = 1.5 * exp((-40 - prevV_axon) / 33);
    double dh_dt_a = (h_inf_a - c.ha) / tau_h_a;
    double m_a = m_inf_a;
    double h_a = c.ha + P.DELTA * dh_dt_a;
    // AxonPotassium ST:812-831
    const double xa = prevV_axon + 25;
    double alpha_x_a = xa == 0 ? 1.3 : 0.13 * xa / (1 - exp(-xa / 10));
    double beta_x_a = 1.69 * exp(-0.0125 * (prevV_axon + 35));
    double x_inf_a = alpha_x_a / (alpha_x_a + beta_x_a);
    double tau_x_a = 1 / (alpha_x_a + beta_x_a);
    double dx_dt_a = (x_inf_a - c.xa) / tau_x_a;
    double x_a = 0.05 * dx_dt_a + c.xa;
    // AxonCurrVolt ST:833-858
    double I_Na_a = P.G_NA_A * m_a * m_a * m_a * h_a * (prevV_axon - P.V_NA);
    double I_la = P.G_LA * (prevV_axon - P.V_L);
    double I_sa = P.G_sa * (prevV_axon - prevV_soma);
    double xa2 = x_a * x_a;
    double I_K_a = P.G_K_A * (xa2 * xa2) * (prevV_axon - P.V_K);
    double dVa_dt = (-(I_K_a + I_sa + I_la + I_Na_a)) / P.C_M;
    c.ma = m_a; c.ha = h_a; c.xa = x_a;
    return P.DELTA * dVa_dt + prevV_axon;
}

/* the compartments with the arithmetic switch (FP64 has one arithmetic) */
template <bool FM>
__device__ __forceinline__ double dend(const Model<double> &P, CellState64 &c, double Vd, double Vs, double I_c,
                                       double iApp)
{
    return ode_dend(P, c, Vd, Vs, I_c, iApp);
}
template <bool FM>
__device__ __forceinline__ float dend(const Model<float> &P, CellState32 &c, float Vd, float Vs, float I_c,
                                      float iApp)
{
    return ode_dend<FM>(P, c, Vd, Vs, I_c, iApp);
}
template <bool FM>
__device__ __forceinline__ double dend_pre(const Model<double> &P, CellState64 &c, double Vd, double Vs, double &I_h)
{
    return ode_dend_pre(P, c, Vd, Vs, I_h);
}
template <bool FM>
__device__ __forceinline__ float dend_pre(const Model<float> &P, CellState32 &c, float Vd, float Vs, float &I_h)
{
    return ode_dend_pre<FM>(P, c, Vd, Vs, I_h);
}
template <bool FM>
__device__ __forceinline__ double dend_post(const Model<double> &P, double Vd, double pre, double I_h, double I_c,
                                            double iApp)
{
    return ode_dend_post(P, Vd, pre, I_h, I_c, iApp);
}
template <bool FM>
__device__ __forceinline__ float dend_post(const Model<float> &P, float Vd, float pre, float I_h, float I_c,
                                           float iApp)
{
    return ode_dend_post<FM>(P, Vd, pre, I_h, I_c, iApp);
}
template <bool FM>
__device__ __forceinline__ double soma(const Model<double> &P, CellState64 &c, double Vd, double Vs, double Va)
{
    return ode_soma(P, c, Vd, Vs, Va);
}
template <bool FM>
__device__ __forceinline__ float soma(const Model<float> &P, CellState32 &c, float Vd, float Vs, float Va)
{
    return ode_soma<FM>(P, c, Vd, Vs, Va);
}
template <bool FM>
__device__ __forceinline__ double axon(const Model<double> &P, CellState64 &c, double Vs, double Va)
{
    return ode_axon(P, c, Vs, Va);
}
template <bool FM>
__device__ __forceinline__ float axon(const Model<float> &P, CellState32 &c, float Vs, float Va)
{
    return ode_axon<FM>(P, c, Vs, Va);
}

/* all three compartments, voltages committed last */
template <bool FM = false, typename T, typename C>
__device__ __forceinline__ void ode(const Model<T> &P, C &c, T I_c, T iApp)
{
    const T Vd = c.vd, Vs = c.vs, Va = c.va;
    const T vd = dend<FM>(P, c, Vd, Vs, I_c, iApp);
    const T vs = soma<FM>(P, c, Vd, Vs, Va);
    const T va = axon<FM>(P, c, Vs, Va);
    c.vd = vd;
    c.vs = vs;
    c.va = va;
}

template <typename T> struct CellOf;
template <> struct CellOf<float> { using type = CellState32; };
template <> struct CellOf<double> { using type = CellState64; };

/* ---------------------------------------------------------------------- */
/* gap-junction term and accumulators                                       */
/* ---------------------------------------------------------------------- */

/* w * f(V) * V with f(V) = 0.8 exp(-V^2/100) + 0.2 (OMP:514-517, ST:549-552).
 * FP32: exp(-V^2/100) = exp2(V^2 * (-log2(e)/100)) on the hardware exp2
 * unit (one multiply, one v_exp_f32), and 0.8 e + 0.2 as one fused
 * multiply-add, then times w and V.  Of the forms measured against FP64 on
 * the 10,000-cell dense network this one keeps the compensated sums closest
 * (7.2e-4 mV over 100 ms, against 8.2e-4 with a separate multiply and add);
 * folding a constant w into the constants (0.8 w, 0.2 w) gives every term
 * of a cell the same rounding bias, which the compensated sum then
 * accumulates faithfully (1.3e-3 mV).  FP64: libm-accurate exp. */
__device__ __forceinline__ float gj_term(float vi, float vj, float w)
{
    const float K = -0.014426950408889634f; /* -log2(e) / 100 */
    float V = vi - vj;
    float e = __builtin_amdgcn_exp2f((V * V) * K);
    float f = __builtin_fmaf(0.8f, e, 0.2f);
    return w * f * V;
}

__device__ __forceinline__ double gj_term(double vi, double vj, double w)
{
    double V = vi - vj;
    double f = 0.8 * exp(-1 * (V * V) / 100) + 0.2;
    return w * f * V;
}

template <typename T, bool KAHAN>
struct Acc {
    T s = 0, c = 0;
    __device__ __forceinline__ void add(T x)
    {
        if constexpr (KAHAN) {
            T t = s + x;                       // TwoSum (Knuth): exact error of s + x
            T bp = t - s;
            c += (s - (t - bp)) + (x - bp);
            s = t;
        } else {
            s += x;
        }
    }
    __device__ __forceinline__ void merge(T os, T oc)
    {
        if constexpr (KAHAN) {
            T t = s + os;
            T bp = t - s;
            T e = (s - (t - bp)) + (os - bp);
            c = c + oc + e;
            s = t;
        } else {
            s += os;
        }
    }
    __device__ __forceinline__ T value() const
    {
        if constexpr (KAHAN) return s + c;
        else return s;
    }
};

__device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }

/* ---------------------------------------------------------------------- */
/* cell state in HBM                                                        */
/* ---------------------------------------------------------------------- */

template <typename T>
__device__ __forceinline__ T iapp_of(const StepArgs<T> &a, int64_t cell, int64_t t)
{
    if (a.tab && t >= a.tab_t0 && t < a.tab_t0 + a.tab_n)
        return a.tab[(t - a.tab_t0) * a.n + cell];
    T b = a.iapp_base ? a.iapp_base[cell] : T(0);
    bool on = (t >= a.t0) && (t < a.t1) && (!a.mask || a.mask[cell]);
    return on ? b + a.amp : b;
}

/* the start-of-step state of a cell: voltages from the ping-pong halves
 * of step t, gates from the in-place planes */
template <typename T, typename C>
__device__ __forceinline__ void load_cell(const StepArgs<T> &a, int64_t cell, int64_t t, C &c)
{
    const int64_t np = a.npad;
    const T *st = a.st;
    c.vd = a.vd[t & 1][cell];
    c.vs = a.vs[t & 1][cell];
    c.va = a.va[t & 1][cell];
    c.r = st[CALCIUM_R * np + cell];
    c.s = st[POTASSIUM_S * np + cell];
    c.q = st[HCURRENT_Q * np + cell];
    c.ca = st[CA2PLUS * np + cell];
    c.icah = st[I_CAH * np + cell];
    c.gcal = st[G_CAL * np + cell];
    c.h = st[SODIUM_H * np + cell];
    c.n = st[POTASSIUM_N * np + cell];
    c.p = st[POTASSIUM_P * np + cell];
    c.xs = st[POTASSIUM_X_S * np + cell];
    c.k = st[CALCIUM_K * np + cell];
    c.l = st[CALCIUM_L * np + cell];
    c.ha = st[SODIUM_H_A * np + cell];
    c.xa = st[POTASSIUM_X_A * np + cell];
    c.m = 0;
    c.ma = 0;
}

template <typename T, typename C>
__device__ __forceinline__ void store_cell(const StepArgs<T> &a, int64_t cell, int64_t t, const C &c)
{
    const int64_t np = a.npad;
    T *st = a.st;
    a.vd[(t & 1) ^ 1][cell] = c.vd;
    a.vs[(t & 1) ^ 1][cell] = c.vs;
    a.va[(t & 1) ^ 1][cell] = c.va;
    st[CALCIUM_R * np + cell] = c.r;
    st[POTASSIUM_S * np + cell] = c.s;
    st[HCURRENT_Q * np + cell] = c.q;
    st[CA2PLUS * np + cell] = c.ca;
    st[I_CAH * np + cell] = c.icah;
    st[SODIUM_M * np + cell] = c.m;
    st[SODIUM_H * np + cell] = c.h;
    st[POTASSIUM_N * np + cell] = c.n;
    st[POTASSIUM_P * np + cell] = c.p;
    st[POTASSIUM_X_S * np + cell] = c.xs;
    st[CALCIUM_K * np + cell] = c.k;
    st[CALCIUM_L * np + cell] = c.l;
    st[SODIUM_M_A * np + cell] = c.ma;
    st[SODIUM_H_A * np + cell] = c.ha;
    st[POTASSIUM_X_A * np + cell] = c.xa;
    if (a.trace && (t % a.stride) == 0)
        a.trace[((t / a.stride) % a.trace_cap) * a.n + cell] = c.va;
}

/* ---------------------------------------------------------------------- */
/* one lane's share of a sliced-ELL row                                     */
/* ---------------------------------------------------------------------- */

constexpr int kUnroll = 8;   // row entries in flight per lane

/* The lane owns row positions p = kk*L + l (kk = 0, 1, ...) of a row of d
 * entries; entry kk sits at base + kk*64.  W = the slice width, uniform
 * across the wavefront, so every slot kk < W is allocated: slots past the
 * row's end are padding (sender = the row's own cell, conductance 0) and are
 * loaded like the
 * others -- no divergent branches, kUnroll loads in flight -- while their
 * term is replaced by an exact +0 (a no-op for the plain and the TwoSum
 * accumulator).  Terms are accumulated in position order. */
template <typename T, bool KAHAN, int L, typename VOf>
__device__ __forceinline__ void row_terms(Acc<T, KAHAN> &acc, T vi, int d, int l, int W, int64_t base,
                                          const T *__restrict__ w, VOf v_of)
{
    for (int kk0 = 0; kk0 < W; kk0 += kUnroll) {
        T wv[kUnroll], vv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (kk0 + u < W) {
                wv[u] = w[base + (int64_t)(kk0 + u) * 64];
                vv[u] = v_of(kk0 + u);
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (kk0 + u < W) {
                const T term = gj_term(vi, vv[u], wv[u]);
                acc.add(((kk0 + u) * L + l < d) ? term : T(0));
            }
        }
    }
}

/* the same with the sender index loaded first (direct gather) */
template <typename T, bool KAHAN, int L, bool NT, typename IdxT, typename VOf>
__device__ __forceinline__ void row_terms_idx(Acc<T, KAHAN> &acc, T vi, int d, int l, int W, int64_t base,
                                              const T *__restrict__ w, const IdxT *__restrict__ idx, VOf v_of)
{
    for (int kk0 = 0; kk0 < W; kk0 += kUnroll) {
        int jv[kUnroll];
        T wv[kUnroll], vv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (kk0 + u < W) {
                const int64_t e = base + (int64_t)(kk0 + u) * 64;
                if constexpr (NT) {   // streamed once: keep it out of the caches
                    jv[u] = (int)__builtin_nontemporal_load(&idx[e]);
                    wv[u] = __builtin_nontemporal_load(&w[e]);
                } else {
                    jv[u] = (int)idx[e];
                    wv[u] = w[e];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
            if (kk0 + u < W) vv[u] = v_of(jv[u]);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (kk0 + u < W) {
                const T term = gj_term(vi, vv[u], wv[u]);
                acc.add(((kk0 + u) * L + l < d) ? term : T(0));
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* one workgroup, one block of cells, one step                              */
/* ---------------------------------------------------------------------- */

/* Phase 1 of the staged gather: the block's neighbour entries, sorted by
 * sender id, copy V_dend[sender] into LDS slot `slot` (the entry's position
 * in the block's compact row-major list).  Sorted senders make the 64 lanes
 * of a load touch a few cache lines instead of 64. */
template <typename T>
__device__ __forceinline__ void stage_gather(const StepArgs<T> &a, int64_t blk, const T *__restrict__ vin,
                                             T *__restrict__ vj)
{
    const int tid = threadIdx.x;
    const int64_t g0 = a.blk_gbase[blk], g1 = a.blk_gbase[blk + 1];
    if (a.gl32) {
        const uint32_t *__restrict__ gl = a.gl32;
        const unsigned sb = (unsigned)a.sb;
        const uint32_t mask = (1u << sb) - 1u;
        int64_t e = g0 + tid;
        constexpr int GU = 8;
        for (; e + (GU - 1) * kBlock < g1; e += GU * kBlock) {
            uint32_t u[GU];
            T x[GU];
#pragma unroll
            for (int q = 0; q < GU; q++) u[q] = gl[e + q * kBlock];
#pragma unroll
            for (int q = 0; q < GU; q++) x[q] = vin[u[q] >> sb];
#pragma unroll
            for (int q = 0; q < GU; q++) vj[u[q] & mask] = x[q];
        }
        for (; e < g1; e += kBlock) {
            const uint32_t u = gl[e];
            vj[u & mask] = vin[u >> sb];
        }
    } else {
        const uint64_t *__restrict__ gl = a.gl64;
        for (int64_t e = g0 + tid; e < g1; e += kBlock) {
            const uint64_t u = gl[e];
            vj[(uint32_t)u] = vin[(int64_t)(u >> 32)];
        }
    }
}

template <typename T, int L, bool KAHAN, int G>
__device__ __forceinline__ void step_block(const StepArgs<T> &a, int64_t blk, int64_t t,
                                           T *ic_lds, T *dyn)
{
    constexpr int CPW = 64 / L;             // cells per wavefront per round
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int cw = lane / L, l = lane % L;
    const T *__restrict__ vin = a.vd[t & 1];
    int64_t cell0;
    int ncell;
    if constexpr (G == GATHER_STAGED) {
        cell0 = a.blk_cell0[blk];
        ncell = a.blk_ncell[blk];
    } else {
        cell0 = blk * kBlock;
        ncell = kBlock;
    }

    if constexpr (G == GATHER_VLDS) {
        for (int64_t i = tid; i < a.n; i += kBlock) dyn[i] = vin[i];
        __syncthreads();
    } else if constexpr (G == GATHER_STAGED) {
        stage_gather(a, blk, vin, dyn);
        __syncthreads();
    }

    // (1) gap-junction currents: slices of CPW cells, one per wavefront per round
    const int nsl = ncell / CPW;
    const int rounds = (nsl + 3) >> 2;
    for (int r = 0; r < rounds; r++) {
        const int sl = r * 4 + wave;                 // block-local slice (wave-uniform)
        const int local = sl * CPW + cw;
        const int64_t cell = cell0 + local;
        Acc<T, KAHAN> acc;
        if (sl < nsl && cell < a.n) {
            const int d = a.deg[cell];
            if (d > 0) {
                const T vi = (G == GATHER_VLDS) ? dyn[cell] : vin[cell];
                const int64_t s_ = cell0 / CPW + sl;
                const int64_t base0 = a.slice_ptr[s_];
                const int W = (int)((a.slice_ptr[s_ + 1] - base0) >> 6);
                const int64_t base = base0 + cw * L + l;
                if constexpr (G == GATHER_STAGED) {
                    const T *__restrict__ vj = dyn + a.lrs[cell] + l;
                    row_terms<T, KAHAN, L>(acc, vi, d, l, W, base, a.w,
                                           [&](int kk) { return vj[kk * L]; });
                } else {
                    const T *__restrict__ vsrc = (G == GATHER_VLDS) ? dyn : vin;
                    if (a.stream_nt)
                        row_terms_idx<T, KAHAN, L, true>(acc, vi, d, l, W, base, a.w, a.col,
                                                         [&](int j) { return vsrc[j]; });
                    else
                        row_terms_idx<T, KAHAN, L, false>(acc, vi, d, l, W, base, a.w, a.col,
                                                          [&](int j) { return vsrc[j]; });
                }
            }
        }
        if constexpr (L > 1) {
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1) {
                T os = shfl_xor(acc.s, off);
                T oc = KAHAN ? shfl_xor(acc.c, off) : T(0);
                acc.merge(os, oc);
            }
        }
        if (l == 0 && sl < nsl) ic_lds[local] = acc.value();
    }
    __syncthreads();

    // (2) compartment update, one thread per cell
    const int64_t cell = cell0 + tid;
    if (tid < ncell && cell < a.n) {
        typename CellOf<T>::type c;
        load_cell(a, cell, t, c);
        ode(a.m, c, ic_lds[tid], iapp_of(a, cell, t));
        store_cell(a, cell, t, c);
    }
    __syncthreads();   // ic_lds / dyn are reused by the next block or step
}

/* ---------------------------------------------------------------------- */
/* kernels                                                                  */
/* ---------------------------------------------------------------------- */

/* One step: step index = (*d_step if given) + k; one workgroup per block. */
template <typename T, int L, bool KAHAN, int G>
__global__ __launch_bounds__(kBlock) void k_step(StepArgs<T> a, int64_t k)
{
    __shared__ T ic_lds[kBlock];
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    step_block<T, L, KAHAN, G>(a, blockIdx.x, t, ic_lds, reinterpret_cast<T *>(dyn_lds));
}

/* Bounded grid barrier: every workgroup arrives once per barrier; waits at
 * most kMaxSpins polls, then records a timeout that makes every later
 * barrier return at once, so the grid always drains. */
constexpr unsigned kMaxSpins = 1u << 22;

__device__ __forceinline__ bool grid_barrier(GridBarrier *gb, unsigned target)
{
    __shared__ int s_ok;
    // The workgroup barrier orders every wave's stores before thread 0
    // (workgroup-scope release and acquire), and thread 0's agent-scope
    // release publishes them with one L2 write-back for the whole workgroup
    // (measured on a 256-workgroup grid: 10.8 us per barrier, against 49.5
    // with a release in every wave).  The wait spins on relaxed loads (an
    // acquire load per spin would invalidate the L2 on every iteration) and
    // acquires once, for the whole workgroup.
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        if (__hip_atomic_load(&gb->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
        } else {
            __hip_atomic_fetch_add(&gb->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            while (__hip_atomic_load(&gb->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (__hip_atomic_load(&gb->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
                if (++spins > kMaxSpins) {
                    __hip_atomic_store(&gb->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

/* `nsteps` steps starting at *d_step; workgroups loop over cell blocks. */
template <typename T, int L, bool KAHAN, int G>
__global__ __launch_bounds__(kBlock) void k_persist(StepArgs<T> a, GridBarrier *gb, int nsteps)
{
    __shared__ T ic_lds[kBlock];
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int64_t t0 = *a.d_step;
    for (int k = 0; k < nsteps; k++) {
        for (int64_t b = blockIdx.x; b < a.nblocks; b += gridDim.x)
            step_block<T, L, KAHAN, G>(a, b, t0 + k, ic_lds, reinterpret_cast<T *>(dyn_lds));
        if (k + 1 < nsteps && !grid_barrier(gb, (unsigned)(k + 1) * gridDim.x)) return;
    }
}

/* One workgroup's cell range of GATHER_NETS: cells [r0, r1), whole
 * networks.  With lds, the range's neighbour lists live in LDS for the whole
 * launch, copied from `off` in the lists buffer: slot p of local cell i at
 * p * stride + i, wmax slots, 16-bit range-local sender ids, then the
 * conductances (none when every conductance of the range is wc). */
struct NetsRange {
    int32_t r0, r1;
    int32_t lds, wconst;
    int64_t off;
    int32_t wmax, pad;
    double wc;
};

/* Whole networks per workgroup, many steps per launch (block-diagonal
 * batches of small networks).  Workgroup g owns rg[g]: whole networks, so
 * every neighbour of its cells is its own.  It keeps their V_dend in LDS
 * (two halves, by step parity), their state in registers and, when they
 * fit, their neighbour lists in LDS, and advances nsteps steps with nothing
 * but workgroup barriers between them: no other workgroup's data is read, so
 * no cross-workgroup hand-over (which costs more than a kernel boundary on
 * this GPU) is needed.  Thread i holds local cells i + q * blockDim
 * (q < CPT); stride = blockDim * CPT.  Each row is summed in list order up to
 * the width of its 64-cell slice of the one-lane sliced-ELL layout, padding
 * slots adding +0, exactly as the direct gather sums it (row_terms_idx with
 * L = 1), and each cell is updated by `ode`, so the results are those of the
 * other strategies at one lane per cell, bit for bit.  The state returns to
 * HBM after the last step; the V_axon trace is written on the trace steps as
 * usual. */
template <typename T, bool KAHAN, int CPT>
__global__ __launch_bounds__(1024) void k_persist_nets(StepArgs<T> a, const NetsRange *__restrict__ rg,
                                                       const unsigned char *__restrict__ lists, int nsteps)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int nt = blockDim.x, tid = threadIdx.x, stride = nt * CPT;
    const NetsRange R = rg[blockIdx.x];
    const int64_t r0 = R.r0;
    const int nc = R.r1 - R.r0;
    T *v0 = reinterpret_cast<T *>(dyn_lds);
    T *v1 = v0 + stride;
    const uint16_t *ids = reinterpret_cast<const uint16_t *>(v1 + stride);
    const T *wl = reinterpret_cast<const T *>(dyn_lds + 2 * (size_t)stride * sizeof(T) +
                                              (((size_t)R.wmax * stride * 2 + 15) & ~(size_t)15));
    if (R.lds) {
        // the range's lists into LDS, 16 bytes per thread and load
        const size_t bytes = (((size_t)R.wmax * stride * 2 + 15) & ~(size_t)15) +
                             (R.wconst ? 0 : (size_t)R.wmax * stride * sizeof(T));
        const uint4 *src = reinterpret_cast<const uint4 *>(lists + R.off);
        uint4 *dst = reinterpret_cast<uint4 *>(v1 + stride);
        for (size_t i = tid; i < bytes / 16; i += nt) dst[i] = src[i];
    }
    const T wc = (T)R.wc;
    const int64_t t0 = *a.d_step;
    typename CellOf<T>::type cs[CPT];
    T ib[CPT];
    bool mk[CPT];
    int dg[CPT], W[CPT];
    int64_t base[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int loc = tid + q * nt;
        const int64_t cell = r0 + loc;
        dg[q] = 0;
        W[q] = 0;
        base[q] = 0;
        ib[q] = T(0);
        mk[q] = false;
        if (loc < nc) {
            load_cell(a, cell, t0, cs[q]);
            v0[loc] = cs[q].vd;
            ib[q] = a.iapp_base ? a.iapp_base[cell] : T(0);
            mk[q] = !a.mask || a.mask[cell];
            dg[q] = a.deg[cell];
            const int64_t s_ = cell >> 6;
            const int64_t b0 = a.slice_ptr[s_];
            W[q] = (int)((a.slice_ptr[s_ + 1] - b0) >> 6);
            base[q] = b0 + (cell & 63);
        }
    }
    __syncthreads();
    for (int k = 0; k < nsteps; k++) {
        const int64_t t = t0 + k;
        const T *__restrict__ vc = (k & 1) ? v1 : v0;
        T *__restrict__ vn = (k & 1) ? v0 : v1;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int loc = tid + q * nt;
            if (loc < nc) {
                const int64_t cell = r0 + loc;
                Acc<T, KAHAN> acc;
                if (dg[q] > 0) {
                    const T vi = vc[loc];
                    if (R.lds) {
                        // the row from LDS: batches of kUnroll slots, as row_terms_idx
                        const int d = dg[q], Wq = W[q];
                        for (int kk0 = 0; kk0 < Wq; kk0 += kUnroll) {
                            int jv[kUnroll];
                            T wv[kUnroll], vv[kUnroll];
#pragma unroll
                            for (int u = 0; u < kUnroll; u++) {
                                if (kk0 + u < Wq) {
                                    jv[u] = ids[(kk0 + u) * stride + loc];
                                    wv[u] = R.wconst ? wc : wl[(kk0 + u) * stride + loc];
                                }
                            }
#pragma unroll
                            for (int u = 0; u < kUnroll; u++)
                                if (kk0 + u < Wq) vv[u] = vc[jv[u]];
#pragma unroll
                            for (int u = 0; u < kUnroll; u++) {
                                if (kk0 + u < Wq) {
                                    const T term = gj_term(vi, vv[u], wv[u]);
                                    acc.add((kk0 + u) < d ? term : T(0));
                                }
                            }
                        }
                    } else {
                        row_terms_idx<T, KAHAN, 1, false>(acc, vi, dg[q], 0, W[q], base[q], a.w, a.col,
                                                          [&](int j) { return vc[j - r0]; });
                    }
                }
                // iapp_of, with this cell's constant and mask in registers
                T iapp;
                if (a.tab && t >= a.tab_t0 && t < a.tab_t0 + a.tab_n)
                    iapp = a.tab[(t - a.tab_t0) * a.n + cell];
                else
                    iapp = ((t >= a.t0) && (t < a.t1) && mk[q]) ? ib[q] + a.amp : ib[q];
                ode(a.m, cs[q], acc.value(), iapp);
                vn[loc] = cs[q].vd;
                if (a.trace && (t % a.stride) == 0)
                    a.trace[((t / a.stride) % a.trace_cap) * a.n + cell] = cs[q].va;
            }
        }
        __syncthreads();
    }
    // the state after the last step, as store_cell leaves it (no trace)
    const int64_t tl = t0 + nsteps - 1;
    const int nx = (int)((tl & 1) ^ 1);
    const int64_t np = a.npad;
    T *st = a.st;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int loc = tid + q * nt;
        if (loc < nc && nsteps > 0) {
            const int64_t cell = r0 + loc;
            const typename CellOf<T>::type &c = cs[q];
            a.vd[nx][cell] = c.vd;
            a.vs[nx][cell] = c.vs;
            a.va[nx][cell] = c.va;
            st[CALCIUM_R * np + cell] = c.r;
            st[POTASSIUM_S * np + cell] = c.s;
            st[HCURRENT_Q * np + cell] = c.q;
            st[CA2PLUS * np + cell] = c.ca;
            st[I_CAH * np + cell] = c.icah;
            st[SODIUM_M * np + cell] = c.m;
            st[SODIUM_H * np + cell] = c.h;
            st[POTASSIUM_N * np + cell] = c.n;
            st[POTASSIUM_P * np + cell] = c.p;
            st[POTASSIUM_X_S * np + cell] = c.xs;
            st[CALCIUM_K * np + cell] = c.k;
            st[CALCIUM_L * np + cell] = c.l;
            st[SODIUM_M_A * np + cell] = c.ma;
            st[SODIUM_H_A * np + cell] = c.ha;
            st[POTASSIUM_X_A * np + cell] = c.xa;
        }
    }
}

/* The same with one thread per compartment (ranges of at most 320 cells,
 * three groups of ncp = blockDim / 3 threads): threads [0, ncp) hold the
 * dendrites and sum the rows, [ncp, 2 ncp) the somata, [2 ncp, 3 ncp) the
 * axons, each with its compartment's state in registers and the three
 * voltages of every cell in LDS (two halves each, by step parity).  A
 * compartment reads only start-of-step voltages, so the three groups run
 * side by side and the step is the arithmetic of `ode` (and of k_tile_ode3),
 * bit for bit; each thread's dependent chain is a third of a whole cell's,
 * which the few wavefronts of a small range cannot otherwise hide.  Lists:
 * as k_persist_nets, with stride ncp. */
template <typename T, bool KAHAN>
__global__ __launch_bounds__(1024) void k_persist_nets3(StepArgs<T> a, const NetsRange *__restrict__ rg,
                                                        const unsigned char *__restrict__ lists, int nsteps)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int ncp = blockDim.x / 3, tid = threadIdx.x;
    const int part = tid / ncp, loc = tid - part * ncp;   // part: 0 dendrite, 1 soma, 2 axon
    const NetsRange R = rg[blockIdx.x];
    const int64_t r0 = R.r0;
    const int nc = R.r1 - R.r0;
    T *vdl = reinterpret_cast<T *>(dyn_lds);   // [2][ncp]
    T *vsl = vdl + 2 * ncp;                    // [2][ncp]
    T *val = vsl + 2 * ncp;                    // [2][ncp]
    const uint16_t *ids = reinterpret_cast<const uint16_t *>(val + 2 * ncp);
    const T *wl = reinterpret_cast<const T *>(dyn_lds + 6 * (size_t)ncp * sizeof(T) +
                                              (((size_t)R.wmax * ncp * 2 + 15) & ~(size_t)15));
    if (R.lds) {
        const size_t bytes = (((size_t)R.wmax * ncp * 2 + 15) & ~(size_t)15) +
                             (R.wconst ? 0 : (size_t)R.wmax * ncp * sizeof(T));
        const uint4 *src = reinterpret_cast<const uint4 *>(lists + R.off);
        uint4 *dst = reinterpret_cast<uint4 *>(val + 2 * ncp);
        for (size_t i = tid; i < bytes / 16; i += blockDim.x) dst[i] = src[i];
    }
    const T wc = (T)R.wc;
    const int64_t t0 = *a.d_step;
    const int64_t np = a.npad;
    const T *st = a.st;
    const int64_t cell = r0 + loc;
    const bool live = loc < nc;
    typename CellOf<T>::type c;
    T ib = T(0);
    bool mk = false;
    int dg = 0, W = 0;
    int64_t base = 0;
    if (live) {
        const int p = (int)(t0 & 1);
        if (part == 0) {
            c.r = st[CALCIUM_R * np + cell];
            c.s = st[POTASSIUM_S * np + cell];
            c.q = st[HCURRENT_Q * np + cell];
            c.ca = st[CA2PLUS * np + cell];
            c.icah = st[I_CAH * np + cell];
            vdl[loc] = a.vd[p][cell];
            ib = a.iapp_base ? a.iapp_base[cell] : T(0);
            mk = !a.mask || a.mask[cell];
            dg = a.deg[cell];
            const int64_t s_ = cell >> 6;
            const int64_t b0 = a.slice_ptr[s_];
            W = (int)((a.slice_ptr[s_ + 1] - b0) >> 6);
            base = b0 + (cell & 63);
        } else if (part == 1) {
            c.gcal = st[G_CAL * np + cell];
            c.h = st[SODIUM_H * np + cell];
            c.n = st[POTASSIUM_N * np + cell];
            c.p = st[POTASSIUM_P * np + cell];
            c.xs = st[POTASSIUM_X_S * np + cell];
            c.k = st[CALCIUM_K * np + cell];
            c.l = st[CALCIUM_L * np + cell];
            c.m = 0;
            vsl[loc] = a.vs[p][cell];
        } else {
            c.ha = st[SODIUM_H_A * np + cell];
            c.xa = st[POTASSIUM_X_A * np + cell];
            c.ma = 0;
            val[loc] = a.va[p][cell];
        }
    }
    __syncthreads();
    for (int k = 0; k < nsteps; k++) {
        const int64_t t = t0 + k;
        const int cu = k & 1, nx = cu ^ 1;
        const T *__restrict__ vc = vdl + cu * ncp;
        if (live) {
            if (part == 0) {
                Acc<T, KAHAN> acc;
                const T vi = vc[loc];
                if (dg > 0) {
                    if (R.lds) {
                        for (int kk0 = 0; kk0 < W; kk0 += kUnroll) {
                            int jv[kUnroll];
                            T wv[kUnroll], vv[kUnroll];
#pragma unroll
                            for (int u = 0; u < kUnroll; u++) {
                                if (kk0 + u < W) {
                                    jv[u] = ids[(kk0 + u) * ncp + loc];
                                    wv[u] = R.wconst ? wc : wl[(kk0 + u) * ncp + loc];
                                }
                            }
#pragma unroll
                            for (int u = 0; u < kUnroll; u++)
                                if (kk0 + u < W) vv[u] = vc[jv[u]];
#pragma unroll
                            for (int u = 0; u < kUnroll; u++) {
                                if (kk0 + u < W) {
                                    const T term = gj_term(vi, vv[u], wv[u]);
                                    acc.add((kk0 + u) < dg ? term : T(0));
                                }
                            }
                        }
                    } else {
                        row_terms_idx<T, KAHAN, 1, false>(acc, vi, dg, 0, W, base, a.w, a.col,
                                                          [&](int j) { return vc[j - r0]; });
                    }
                }
                T iapp;
                if (a.tab && t >= a.tab_t0 && t < a.tab_t0 + a.tab_n)
                    iapp = a.tab[(t - a.tab_t0) * a.n + cell];
                else
                    iapp = ((t >= a.t0) && (t < a.t1) && mk) ? ib + a.amp : ib;
                vdl[nx * ncp + loc] = dend<false>(a.m, c, vi, vsl[cu * ncp + loc], acc.value(), iapp);
            } else if (part == 1) {
                vsl[nx * ncp + loc] = soma<false>(a.m, c, vc[loc], vsl[cu * ncp + loc], val[cu * ncp + loc]);
            } else {
                const T va = axon<false>(a.m, c, vsl[cu * ncp + loc], val[cu * ncp + loc]);
                val[nx * ncp + loc] = va;
                if (a.trace && (t % a.stride) == 0) a.trace[((t / a.stride) % a.trace_cap) * a.n + cell] = va;
            }
        }
        __syncthreads();
    }
    // the state after the last step, as store_cell leaves it (no trace)
    if (live && nsteps > 0) {
        const int64_t tl = t0 + nsteps - 1;
        const int nx = (int)((tl & 1) ^ 1), lx = nsteps & 1;   // LDS half holding the last step's voltages
        T *sw = a.st;
        if (part == 0) {
            a.vd[nx][cell] = vdl[lx * ncp + loc];
            sw[CALCIUM_R * np + cell] = c.r;
            sw[POTASSIUM_S * np + cell] = c.s;
            sw[HCURRENT_Q * np + cell] = c.q;
            sw[CA2PLUS * np + cell] = c.ca;
            sw[I_CAH * np + cell] = c.icah;
        } else if (part == 1) {
            a.vs[nx][cell] = vsl[lx * ncp + loc];
            sw[SODIUM_M * np + cell] = c.m;
            sw[SODIUM_H * np + cell] = c.h;
            sw[POTASSIUM_N * np + cell] = c.n;
            sw[POTASSIUM_P * np + cell] = c.p;
            sw[POTASSIUM_X_S * np + cell] = c.xs;
            sw[CALCIUM_K * np + cell] = c.k;
            sw[CALCIUM_L * np + cell] = c.l;
        } else {
            a.va[nx][cell] = val[lx * ncp + loc];
            sw[SODIUM_M_A * np + cell] = c.ma;
            sw[SODIUM_H_A * np + cell] = c.ha;
            sw[POTASSIUM_X_A * np + cell] = c.xa;
        }
    }
}

/* I_c of every cell from V_dend at step t (no state change); direct gathers
 * over the uniform 256-cell blocks of the sliced-ELL layout. */
template <typename T, int L, bool KAHAN, bool HYB = false>
__global__ __launch_bounds__(kBlock) void k_gap_currents(StepArgs<T> a, int64_t t, T *out)
{
    constexpr int CPW = 64 / L;
    constexpr int CPR = kBlock / L;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int cw = lane / L, l = lane % L;
    const T *vin = a.vd[t & 1];
    const int64_t cell0 = (int64_t)blockIdx.x * kBlock;
    for (int r = 0; r < L; r++) {
        const int local = r * CPR + wave * CPW + cw;
        const int64_t cell = cell0 + local;
        const int64_t slice = (cell0 + r * CPR + wave * CPW) / CPW;
        Acc<T, KAHAN> acc;
        if (cell < a.n) {
            const int d = a.deg[cell];
            const T vi = vin[cell];
            const int64_t base = a.slice_ptr[slice] + cw * L + l;
            if (HYB && d >= a.fp64_deg) {
                double sd = 0.0;
                for (int kk = 0; kk * L + l < d; kk++) {
                    const int64_t e = base + (int64_t)kk * 64;
                    sd += gj_term((double)vi, (double)vin[a.col[e]], (double)a.w[e]);
                }
                acc.s = (T)sd;
                acc.c = (T)(sd - (double)acc.s);
            } else {
                for (int kk = 0; kk * L + l < d; kk++) {
                    const int64_t e = base + (int64_t)kk * 64;
                    acc.add(gj_term(vi, vin[a.col[e]], a.w[e]));
                }
            }
        }
        if constexpr (L > 1) {
            for (int off = L / 2; off > 0; off >>= 1) {
                T os = shfl_xor(acc.s, off);
                T oc = KAHAN ? shfl_xor(acc.c, off) : T(0);
                acc.merge(os, oc);
            }
        }
        if (l == 0 && cell < a.n) out[cell] = acc.value();
    }
}

/* ---------------------------------------------------------------------- */
/* 2D-tiled gap junctions + separate compartment kernel                     */
/* ---------------------------------------------------------------------- */

template <typename T, bool KAHAN> struct PartOf { using type = T; };
template <> struct PartOf<float, true> { using type = float2; };

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

/* V_dend[s0 .. s0+cn) into LDS with the direct global-to-LDS loads of
 * gfx950 (global_load_lds_dwordx4): wavefront w copies the 1 KiB blocks
 * w, w + NW, ..., lane l's 16 bytes landing at the block's LDS address
 * (M0) + 16 l, so the data never passes through registers and the loads
 * are issued back to back.  Lanes past the last whole vector stay off; the
 * cn % 4 (FP32) or cn % 2 (FP64) remaining values are copied by ordinary
 * loads and stores.  The source is 16-byte aligned (chunks start at
 * multiples of 64 cells of a 256-byte aligned array).  Returns once the
 * copy has landed in LDS (the caller's barrier then publishes it). */
template <typename T, int NT>
__device__ __forceinline__ void tile_chunk_to_lds(const T *__restrict__ src, T *__restrict__ dst, int64_t cn,
                                                  int tid)
{
    constexpr int E = 16 / (int)sizeof(T), NW = NT / 64;
    const int nv = (int)(cn / E);
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const char *sb = reinterpret_cast<const char *>(src);
    char *db = reinterpret_cast<char *>(dst);
    const int nb = (nv + 63) >> 6;   // blocks of 64 vectors
#pragma unroll 4
    for (int blk = wave; blk < nb; blk += NW) {
        const int i = blk * 64 + lane;
        if (i < nv)
            __builtin_amdgcn_global_load_lds(sb + (size_t)i * 16,
                                             (__attribute__((address_space(3))) void *)(db + (size_t)blk * 1024), 16, 0,
                                             0);
    }
    for (int64_t i = (int64_t)nv * E + tid; i < cn; i += NT) dst[i] = src[i];
    __builtin_amdgcn_s_waitcnt(0);   // every load, the LDS ones included, has completed
}

/* load from a wavefront-uniform pointer plus a per-lane 32-bit byte offset
 * (the scalar-base + vector-offset form of the global load) */
template <typename X>
__device__ __forceinline__ X ld_off(const X *__restrict__ p, uint32_t byte_off)
{
    return *reinterpret_cast<const X *>(reinterpret_cast<const char *>(p) + byte_off);
}

/* One lane's share of a tile row: entries at irow/wrow[(kk*64) + loff],
 * irow and wrow uniform across the wavefront, W the slice width.  Same
 * arithmetic and order as row_terms_idx; WC: every junction has
 * conductance wc.  TAG gives each calling kernel its own instantiation:
 * the batch lambda is inlined by the compiler's cost model, which counts
 * its call sites, so a caller added elsewhere would otherwise move the
 * register allocation of k_tile_gj. */
template <typename T, bool KAHAN, int L, bool WC, bool HYB, int TAG = 0>
__device__ __forceinline__ void tile_row_terms(Acc<T, KAHAN> &acc, double &accd, bool wide, T vi, int d, int l,
                                               int W, int nvm, int kstart, int loff,
                                               const uint16_t *__restrict__ irow, const T *__restrict__ wrow, T wc,
                                               const T *__restrict__ vch)
{
    const uint32_t oi = (uint32_t)loff * (uint32_t)sizeof(uint16_t);
    const uint32_t ow = (uint32_t)loff * (uint32_t)sizeof(T);
    // one batch of kUnroll slots from kk0.  CHW: check the slice width W
    // (slots past it do not exist); CHD: check the row's degree (slots past
    // it are padding, read but not added)
    const auto batch = [&](auto chw, auto chd, int kk0) {
        constexpr bool CHW = decltype(chw)::value, CHD = decltype(chd)::value;
        const uint16_t *ib = irow + (int64_t)kk0 * 64;
        const T *wb_ = wrow + (int64_t)kk0 * 64;
        int jv[kUnroll];
        T wv[kUnroll], vv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (!CHW || kk0 + u < W) {
                jv[u] = (int)ld_off(ib + u * 64, oi);
                if constexpr (!WC) wv[u] = ld_off(wb_ + u * 64, ow);
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
            if (!CHW || kk0 + u < W) vv[u] = vch[jv[u]];
        if (HYB && wide) {
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (!CHW || kk0 + u < W) {
                    const double term = gj_term((double)vi, (double)vv[u], (double)(WC ? wc : wv[u]));
                    accd += (!CHD || (kk0 + u) * L + l < d) ? term : 0.0;
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (!CHW || kk0 + u < W) {
                    const T term = gj_term(vi, vv[u], WC ? wc : wv[u]);
                    acc.add((!CHD || (kk0 + u) * L + l < d) ? term : T(0));
                }
            }
        }
    };
    // Per-entry conductances (not WC): a padding slot holds conductance 0,
    // so its term is an exact zero (a finite product with 0), and adding it
    // leaves the sum and its compensation as the degree check's +0 would,
    // bit for bit; no slot needs the check.  The tile's constant conductance
    // (WC) and the FP64 terms of wide rows (HYB) keep it.
    constexpr bool PADZ = !WC && !HYB;
    int kk0 = kstart;
    if constexpr (sizeof(T) == 4) {   // FP64: the extra copies of the loop cost more registers than they save
        // slots below nvm (the fewest valid slots of any lane of the wave)
        // need no check; then whole batches inside the width
        const int nv = PADZ ? W : (nvm < W ? nvm : W);
        for (; kk0 + kUnroll <= nv; kk0 += kUnroll) batch(std::false_type{}, std::false_type{}, kk0);
        if constexpr (!PADZ)
            for (; kk0 + kUnroll <= W; kk0 += kUnroll) batch(std::false_type{}, std::true_type{}, kk0);
    }
    for (; kk0 < W; kk0 += kUnroll) batch(std::true_type{}, std::bool_constant<!PADZ>{}, kk0);
}

/* FP32 tile rows of constant conductance wc, the ids of the next batch
 * loaded while the current one is summed (the unchecked slots below nvm),
 * then the checked tail as above. */
template <bool KAHAN, int L>
__device__ __forceinline__ void tile_row_terms_pf(Acc<float, KAHAN> &acc, float vi, int d, int l, int W, int nvm,
                                                  int loff, const uint16_t *__restrict__ irow, float wc,
                                                  const float *__restrict__ vch)
{
    const uint32_t oi = (uint32_t)loff * (uint32_t)sizeof(uint16_t);
    const int nv = nvm < W ? nvm : W;
    int kk0 = 0;
    if (kUnroll <= nv) {
        int jv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) jv[u] = (int)ld_off(irow + u * 64, oi);
        for (; kk0 + kUnroll <= nv; kk0 += kUnroll) {
            const bool more = kk0 + 2 * kUnroll <= nv;
            int jn[kUnroll];
            if (more) {
#pragma unroll
                for (int u = 0; u < kUnroll; u++) jn[u] = (int)ld_off(irow + (int64_t)(kk0 + kUnroll + u) * 64, oi);
            }
            float vv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) vv[u] = vch[jv[u]];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) acc.add(gj_term(vi, vv[u], wc));
            if (more) {
#pragma unroll
                for (int u = 0; u < kUnroll; u++) jv[u] = jn[u];
            }
        }
    }
    double unused = 0.0;
    tile_row_terms<float, KAHAN, L, true, false>(acc, unused, false, vi, d, l, W, kk0, kk0, loff, irow, vch, wc,
                                                 vch);   // (no conductances are read: wc)
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void part_store(void *p, int64_t i, const Acc<T, KAHAN> &a)
{
    if constexpr (KAHAN) reinterpret_cast<float2 *>(p)[i] = make_float2(a.s, a.c);
    else reinterpret_cast<T *>(p)[i] = a.s;
}

template <typename T, bool KAHAN>
__device__ __forceinline__ void part_merge(const void *p, int64_t i, Acc<T, KAHAN> &a)
{
    if constexpr (KAHAN) {
        const float2 v = reinterpret_cast<const float2 *>(p)[i];
        a.merge(v.x, v.y);
    } else {
        a.merge(reinterpret_cast<const T *>(p)[i], T(0));
    }
}

/* I_c of a cell: its partials over the tiles of its receiver block, merged
 * in ascending sender-chunk order (one partial, already merged, when the
 * block's tiles were swept by one workgroup). */
template <typename T, bool KAHAN>
__device__ __forceinline__ T tiled_ic(const StepArgs<T> &a, int64_t cell)
{
    using P = typename PartOf<T, KAHAN>::type;
    const P *__restrict__ part = reinterpret_cast<const P *>(a.part);
    const int64_t rb = cell / a.rcells, rloc = cell - rb * a.rcells;
    const int32_t t0 = a.rb_part0[rb], t1 = a.rb_part0[rb + 1];   // uniform per wavefront
    Acc<T, KAHAN> acc;
    constexpr int B = 12;   // partials in flight: one batch for up to 12 sender chunks
    for (int32_t tt = t0; tt < t1; tt += B) {
        P v[B];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (tt + q < t1) v[q] = part[(int64_t)(tt + q) * a.rcells + rloc];
#pragma unroll
        for (int q = 0; q < B; q++) {
            if (tt + q < t1) {
                if constexpr (KAHAN) acc.merge(v[q].x, v[q].y);
                else acc.merge(v[q], T(0));
            }
        }
    }
    return acc.value();
}

/* compartment update of every cell, I_c from the tile partials */
template <typename T, bool KAHAN, bool FM>
__global__ __launch_bounds__(kBlock) void k_tile_ode(StepArgs<T> a, int64_t k)
{
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= a.n) return;
    typename CellOf<T>::type c;
    load_cell(a, cell, t, c);
    ode<FM>(a.m, c, tiled_ic<T, KAHAN>(a, cell), iapp_of(a, cell, t));
    store_cell(a, cell, t, c);
}

/* One compartment of one cell for step t, from the start-of-step voltages
 * (ping-pong buffers): the dendrite needs this step's gap-junction current
 * (merged from the tile partials), the soma and the axon do not. */
template <typename T, bool KAHAN, bool FM>
__device__ __forceinline__ void tile_dend_cell(const StepArgs<T> &a, int64_t cell, int64_t t)
{
    const int64_t np = a.npad;
    T *st = a.st;
    const int cur = (int)(t & 1), nxt = cur ^ 1;
    typename CellOf<T>::type c;
    c.q = st[HCURRENT_Q * np + cell];
    c.r = st[CALCIUM_R * np + cell];
    c.s = st[POTASSIUM_S * np + cell];
    c.ca = st[CA2PLUS * np + cell];
    c.icah = st[I_CAH * np + cell];
    const T vd = dend<FM>(a.m, c, a.vd[cur][cell], a.vs[cur][cell], tiled_ic<T, KAHAN>(a, cell),
                          iapp_of(a, cell, t));
    st[HCURRENT_Q * np + cell] = c.q;
    st[CALCIUM_R * np + cell] = c.r;
    st[POTASSIUM_S * np + cell] = c.s;
    st[CA2PLUS * np + cell] = c.ca;
    st[I_CAH * np + cell] = c.icah;
    a.vd[nxt][cell] = vd;
}

/* The dendrite up to its voltage, beside the tiles (no I_c needed yet):
 * gates, Ca2+ and I_CaH in place, the partial current sum and I_h to dpre. */
template <typename T, bool FM>
__device__ __forceinline__ void tile_dend_pre_cell(const StepArgs<T> &a, int64_t cell, int64_t t)
{
    const int64_t np = a.npad;
    T *st = a.st;
    const int cur = (int)(t & 1);
    typename CellOf<T>::type c;
    c.q = st[HCURRENT_Q * np + cell];
    c.r = st[CALCIUM_R * np + cell];
    c.s = st[POTASSIUM_S * np + cell];
    c.ca = st[CA2PLUS * np + cell];
    c.icah = st[I_CAH * np + cell];
    T I_h;
    const T pre = dend_pre<FM>(a.m, c, a.vd[cur][cell], a.vs[cur][cell], I_h);
    st[HCURRENT_Q * np + cell] = c.q;
    st[CALCIUM_R * np + cell] = c.r;
    st[POTASSIUM_S * np + cell] = c.s;
    st[CA2PLUS * np + cell] = c.ca;
    st[I_CAH * np + cell] = c.icah;
    a.dpre[cell] = pre;
    a.dpre[np + cell] = I_h;
}

template <typename T, bool FM>
__device__ __forceinline__ void tile_soma_cell(const StepArgs<T> &a, int64_t cell, int64_t t)
{
    const int64_t np = a.npad;
    T *st = a.st;
    const int cur = (int)(t & 1), nxt = cur ^ 1;
    typename CellOf<T>::type c;
    c.k = st[CALCIUM_K * np + cell];
    c.l = st[CALCIUM_L * np + cell];
    c.h = st[SODIUM_H * np + cell];
    c.n = st[POTASSIUM_N * np + cell];
    c.p = st[POTASSIUM_P * np + cell];
    c.xs = st[POTASSIUM_X_S * np + cell];
    c.gcal = st[G_CAL * np + cell];
    const T vs = soma<FM>(a.m, c, a.vd[cur][cell], a.vs[cur][cell], a.va[cur][cell]);
    st[CALCIUM_K * np + cell] = c.k;
    st[CALCIUM_L * np + cell] = c.l;
    st[SODIUM_M * np + cell] = c.m;
    st[SODIUM_H * np + cell] = c.h;
    st[POTASSIUM_N * np + cell] = c.n;
    st[POTASSIUM_P * np + cell] = c.p;
    st[POTASSIUM_X_S * np + cell] = c.xs;
    a.vs[nxt][cell] = vs;
}

template <typename T, bool FM>
__device__ __forceinline__ void tile_axon_cell(const StepArgs<T> &a, int64_t cell, int64_t t)
{
    const int64_t np = a.npad;
    T *st = a.st;
    const int cur = (int)(t & 1), nxt = cur ^ 1;
    typename CellOf<T>::type c;
    c.ha = st[SODIUM_H_A * np + cell];
    c.xa = st[POTASSIUM_X_A * np + cell];
    const T va = axon<FM>(a.m, c, a.vs[cur][cell], a.va[cur][cell]);
    st[SODIUM_M_A * np + cell] = c.ma;
    st[SODIUM_H_A * np + cell] = c.ha;
    st[POTASSIUM_X_A * np + cell] = c.xa;
    a.va[nxt][cell] = va;
    if (a.trace && (t % a.stride) == 0)
        a.trace[((t / a.stride) % a.trace_cap) * a.n + cell] = va;
}

/* The same update with one wavefront per compartment: a workgroup of 192
 * threads covers 64 cells; wave 0 runs the dendrites, wave 1 the somata,
 * wave 2 the axons, each from the start-of-step voltages. */
template <typename T, bool KAHAN, bool FM>
__global__ __launch_bounds__(192) void k_tile_ode3(StepArgs<T> a, int64_t k)
{
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    const int wave = threadIdx.x >> 6;
    const int64_t cell = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    if (cell >= a.n) return;
    if (wave == 0) tile_dend_cell<T, KAHAN, FM>(a, cell, t);
    else if (wave == 1) tile_soma_cell<T, FM>(a, cell, t);
    else tile_axon_cell<T, FM>(a, cell, t);
}

/* Every cell's soma, axon and dendrite up to its voltage (none needs this
 * step's I_c), in a kernel of its own that runs beside the swept tiles
 * (k_tile_sweep with ode_split 2).  The sweep fills each CU's LDS, so these
 * workgroups, which use none, share its CUs through the registers and
 * wavefront slots it leaves, in a second branch of the step's graph; the
 * dendrite voltages follow both (k_tile_dend).  192 threads cover 64 cells:
 * wave 0 the dendrite gates and partial current sum, wave 1 the somata, wave
 * 2 the axons, from the start-of-step voltages. */
template <typename T, bool FM>
__global__ __launch_bounds__(192) void k_tile_beside(StepArgs<T> a, int64_t k)
{
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    const int wave = threadIdx.x >> 6;
    const int64_t cell = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    if (cell >= a.n) return;
    if (wave == 0) tile_dend_pre_cell<T, FM>(a, cell, t);
    else if (wave == 1) tile_soma_cell<T, FM>(a, cell, t);
    else tile_axon_cell<T, FM>(a, cell, t);
}

/* Dendrite voltages (the rest of every compartment ran beside the tiles),
 * one thread per cell: this step's I_c from the tile partials, then the
 * voltage from the partial current sum. */
template <typename T, bool KAHAN, bool FM>
__global__ __launch_bounds__(kBlock) void k_tile_dend(StepArgs<T> a, int64_t k)
{
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= a.n) return;
    const int cur = (int)(t & 1), nxt = cur ^ 1;
    a.vd[nxt][cell] = dend_post<FM>(a.m, a.vd[cur][cell], a.dpre[cell], a.dpre[a.npad + cell],
                                    tiled_ic<T, KAHAN>(a, cell), iapp_of(a, cell, t));
}

/* One tile: copy the sender chunk's V_dend to LDS, then each receiver's
 * junctions from that chunk (L lanes per receiver, list order) into a
 * partial (sum, compensation) pair, which the receiver's first lane hands to
 * sink(receiver within the block, pair).  Every receiver of the block is
 * listed once per tile, so the sink sees each receiver once.  AWC: every
 * tile has a constant conductance (FP32 only), and the rows prefetch their
 * next ids. */
template <typename T, int L, bool KAHAN, int NT, bool HYB, bool AWC, typename Sink>
__device__ __forceinline__ void tile_rows(const StepArgs<T> &a, const TileDesc td, const T *__restrict__ vin,
                                          T *__restrict__ vch, Sink &&sink)
{
    constexpr int CPW = 64 / L;
    // the wavefront index is uniform: make it scalar so that the slice
    // pointers and widths below live in SGPRs and the row loads use a
    // scalar base with a 32-bit lane offset
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int cw = lane / L, l = lane % L;
    const int64_t s0 = td.s0;
    const int64_t cn = td.cn;
    const int64_t cellb = (int64_t)td.rb * a.rcells;
    constexpr int NW = NT / 64;
    const int loff = cw * L + l;
    // wavefront w takes slices w, 2NW-1-w, 2NW+w, ...: the tile's slices are
    // ordered by width, so the snake evens out the waves' work.  The next
    // slice's receiver entries are loaded while the current one is summed;
    // the first slice's while the chunk is copied.
    const auto snake = [&](int r) { return r * NW + ((r & 1) ? NW - 1 - wave : wave); };
    int sl = snake(0), rloc = 0, d = 0;
    if (sl < td.nsl) {
        const int64_t slot = td.rl0 + (int64_t)sl * CPW + cw;
        rloc = a.rlist[slot];
        d = a.rdeg[slot];
    }
    tile_chunk_to_lds<T, NT>(vin + s0, vch, cn, tid);
    __syncthreads();
    for (int r = 0; r * NW < td.nsl; r++) {
        const int sn = snake(r + 1);
        int rloc_n = 0, d_n = 0;
        if (sn < td.nsl) {
            const int64_t slot = td.rl0 + (int64_t)sn * CPW + cw;
            rloc_n = a.rlist[slot];
            d_n = a.rdeg[slot];
        }
        if (sl < td.nsl) {
            Acc<T, KAHAN> acc;
            // the fewest valid slots of any lane in the wave: below it no
            // slot needs the degree check.  A tile lists its receivers by
            // in-chunk degree, descending, so the slice's last row (lane 63
            // holds its last position) has the smallest degree dmin, and the
            // fewest valid slots are dmin / L
            int nvm = 0;
            if constexpr (sizeof(T) == 4) nvm = __builtin_amdgcn_readlane(d, 63) / L;
            if (d > 0) {
                const T vi = vin[cellb + rloc];
                const int64_t base0 = a.tsp[td.sl0 + sl];
                const int W = (int)((a.tsp[td.sl0 + sl + 1] - base0) >> 6);
                // high-degree receivers: FP64 terms and sum, returned as a
                // (sum, compensation) pair of floats
                const bool wide = HYB && a.deg[cellb + rloc] >= a.fp64_deg;
                double accd = 0.0;
                if constexpr (AWC && sizeof(T) == 4 && !HYB)
                    tile_row_terms_pf<KAHAN, L>(acc, vi, d, l, W, nvm, loff, a.tcol + base0, (float)td.wc, vch);
                else if (td.wconst)
                    tile_row_terms<T, KAHAN, L, true, HYB>(acc, accd, wide, vi, d, l, W, nvm, 0, loff,
                                                           a.tcol + base0, a.tw + base0, (T)td.wc, vch);
                else
                    tile_row_terms<T, KAHAN, L, false, HYB>(acc, accd, wide, vi, d, l, W, nvm, 0, loff,
                                                            a.tcol + base0, a.tw + base0, (T)td.wc, vch);
                if (HYB && wide) {
                    acc.s = (T)accd;
                    acc.c = (T)(accd - (double)acc.s);
                }
            }
            if constexpr (L > 1) {
#pragma unroll
                for (int off = L / 2; off > 0; off >>= 1) {
                    T os = shfl_xor(acc.s, off);
                    T oc = KAHAN ? shfl_xor(acc.c, off) : T(0);
                    acc.merge(os, oc);
                }
            }
            if (l == 0) sink(rloc, acc);
        }
        sl = sn;
        rloc = rloc_n;
        d = d_n;
    }
}

/* One workgroup per tile (tile_rows), the partial pairs stored to
 * part[tile][receiver]; workgroups past the tiles update somata, axons and
 * dendrite gates.  The AWC variant's launch bounds leave room for the
 * prefetched ids (four waves per SIMD rather than eight; it compiles to
 * 60 VGPRs, still eight). */
template <typename T, int L, bool KAHAN, int NT, bool HYB = false, bool AWC = false>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 && !HYB) ? (AWC ? 4 : 8) : 1) void k_tile_gj(StepArgs<T> a, int64_t k)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    T *vch = reinterpret_cast<T *>(dyn_lds);
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    if ((int)blockIdx.x >= a.ntiles) {
        // workgroups past the tiles: somata (even waves), axons and the
        // dendrites up to their voltage (odd waves) of groups of NT/2 cells,
        // which need no gap-junction current, so they run beside the tiles;
        // workgroup g takes groups g, g + nsa, ...
        const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
        for (int64_t g = (int64_t)blockIdx.x - a.ntiles; g * (NT / 2) < a.n; g += a.nsa) {
            const int64_t cell = (g * (NT / 128) + (wv >> 1)) * 64 + (threadIdx.x & 63);
            if (cell < a.n) {
                // even waves: somata; odd waves: axons, then the dendrites up
                // to their voltage (about 12 exp each way)
                if (a.fast_math) {
                    if (wv & 1) tile_axon_cell<T, true>(a, cell, t);
                    else tile_soma_cell<T, true>(a, cell, t);
                    if (wv & 1) tile_dend_pre_cell<T, true>(a, cell, t);
                } else {
                    if (wv & 1) tile_axon_cell<T, false>(a, cell, t);
                    else tile_soma_cell<T, false>(a, cell, t);
                    if (wv & 1) tile_dend_pre_cell<T, false>(a, cell, t);
                }
            }
        }
        return;
    }
    const int64_t ti = blockIdx.x;
    tile_rows<T, L, KAHAN, NT, HYB, AWC>(a, a.tiles[ti], a.vd[t & 1], vch,
                                         [&](int rloc, const Acc<T, KAHAN> &acc) {
                                             part_store<T, KAHAN>(a.part, ti * a.rcells + rloc, acc);
                                         });
}

/* The rows of one swept tile.  A swept tile's rows are short (a block's
 * junctions spread over every chunk: 1.6 per receiver per chunk at 1M
 * cells of degree 50), so per-slice work and load latency dominate.  Each
 * wavefront runs its slices as a pipeline, loading the receiver's V_dend
 * with the first K = 4 id / conductance slots DB = 1 slice ahead, and its
 * receiver, degree and slice bounds one slice before that; longer rows load
 * the rest in place (tile_row_terms from slot K), and only they, in tiles
 * of constant conductance or rows of FP64 terms, need the wavefront's
 * fewest valid slots.  Measured at 1M cells of degree 50, with
 * the compartments in a kernel after the sweep: 275 us per step; 6 slots
 * ahead 277, 8 slots 285 (305 with the minimum on every slice), two slices
 * ahead 297, four slices' loads at once before summing them 336 (8 slots).
 * The terms are added in slot order with the checks of
 * tile_row_terms, so each row's pair is the one tile_rows makes, and it is
 * merged into the receiver's accumulator in LDS. */
template <typename T, int L, bool KAHAN, int NT, bool HYB, typename Idle>
__device__ __forceinline__ void sweep_tile_rows(const StepArgs<T> &a, const TileDesc td, const T *__restrict__ vin,
                                                T *__restrict__ vch, typename PartOf<T, KAHAN>::type *accl,
                                                Idle &&idle)
{
    constexpr int CPW = 64 / L, NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int cw = lane / L, l = lane % L;
    const int loff = cw * L + l;
    const int64_t cellb = (int64_t)td.rb * a.rcells;
    const int nsl = td.nsl;
    const bool wct = td.wconst != 0;
    const T wc = (T)td.wc;
    const auto snake = [&](int r) { return r * NW + ((r & 1) ? NW - 1 - wave : wave); };
    constexpr int DB = 1, K = 4;   // K: slots loaded ahead
    // stage A (DB + 1 slices ahead): receiver, degree, first entry, width
    struct SA {
        int rloc = 0, d = 0, W = 0;
        int64_t b0 = 0;
    };
    // stage B (DB slices ahead): the receiver's V_dend, the first slots
    struct SB {
        T vi = 0;
        int jv[K];
        T wv[K];
    };
    const auto load_a = [&](int sl, SA &x) {
        if (sl < nsl) {
            const int64_t slot = td.rl0 + (int64_t)sl * CPW + cw;
            x.rloc = a.rlist[slot];
            x.d = a.rdeg[slot];
            x.b0 = a.tsp[td.sl0 + sl];
            x.W = (int)((a.tsp[td.sl0 + sl + 1] - x.b0) >> 6);
        }
    };
    const auto load_b = [&](int sl, const SA &x, SB &y) {
        if (sl < nsl) {
            // receivers past the network's end (last block) have degree 0
            // and read nothing
            if (x.d > 0) y.vi = vin[cellb + x.rloc];
            const uint16_t *ib = a.tcol + x.b0;
            const T *wb = a.tw + x.b0;
#pragma unroll
            for (int u = 0; u < K; u++) {
                if (u < x.W) {
                    y.jv[u] = (int)ld_off(ib + u * 64, (uint32_t)loff * (uint32_t)sizeof(uint16_t));
                    if (!wct) y.wv[u] = ld_off(wb + u * 64, (uint32_t)loff * (uint32_t)sizeof(T));
                }
            }
        }
    };
    SA av[DB + 2];
    SB bv[DB + 1];
#pragma unroll
    for (int i = 0; i <= DB; i++) load_a(snake(i), av[i]);
#pragma unroll
    for (int i = 0; i < DB; i++) load_b(snake(i), av[i], bv[i]);
    tile_chunk_to_lds<T, NT>(vin + td.s0, vch, td.cn, tid);
    __syncthreads();
    // wavefronts without a slice of this tile (a block of few slices)
    // take other work while the rows are summed
    if (nsl < NW && wave >= nsl) idle(wave - nsl, NW - nsl);
    for (int r = 0; r * NW < nsl; r++) {
        const int sl = snake(r);
        load_a(snake(r + DB + 1), av[DB + 1]);
        load_b(snake(r + DB), av[DB], bv[DB]);
        const SA &a0 = av[0];
        const SB &b0 = bv[0];
        if (sl < nsl) {
            const int d = a0.d, W = a0.W;
            Acc<T, KAHAN> acc;
            // the fewest valid slots of any lane: the tile lists receivers by
            // in-chunk degree, descending, so the slice's last row (lane 63)
            // has the smallest degree, as in tile_rows
            int nvm = 0;
            if (sizeof(T) == 4 && W > K && (wct || HYB)) nvm = __builtin_amdgcn_readlane(d, 63) / L;
            if (d > 0) {
                const T vi = b0.vi;
                const bool wide = HYB && a.deg[cellb + a0.rloc] >= a.fp64_deg;
                double accd = 0.0;
                T vv[K];
#pragma unroll
                for (int u = 0; u < K; u++)
                    if (u < W) vv[u] = vch[b0.jv[u]];
                if (!wct && !(HYB && wide)) {
                    // per-entry conductances: padding slots add exact zeros
                    // (tile_row_terms), no degree check
#pragma unroll
                    for (int u = 0; u < K; u++)
                        if (u < W) acc.add(gj_term(vi, vv[u], b0.wv[u]));
                } else {
#pragma unroll
                    for (int u = 0; u < K; u++) {
                        if (u < W) {
                            const bool ok = u * L + l < d;
                            if (HYB && wide) {
                                const double term =
                                    gj_term((double)vi, (double)vv[u], (double)(wct ? wc : b0.wv[u]));
                                accd += ok ? term : 0.0;
                            } else {
                                const T term = gj_term(vi, vv[u], wct ? wc : b0.wv[u]);
                                acc.add(ok ? term : T(0));
                            }
                        }
                    }
                }
                if (W > K) {
                    if (wct)
                        tile_row_terms<T, KAHAN, L, true, HYB, 1>(acc, accd, wide, vi, d, l, W, nvm, K, loff,
                                                                  a.tcol + a0.b0, a.tw + a0.b0, wc, vch);
                    else
                        tile_row_terms<T, KAHAN, L, false, HYB, 1>(acc, accd, wide, vi, d, l, W, nvm, K, loff,
                                                                   a.tcol + a0.b0, a.tw + a0.b0, wc, vch);
                }
                if (HYB && wide) {
                    acc.s = (T)accd;
                    acc.c = (T)(accd - (double)acc.s);
                }
            }
            if constexpr (L > 1) {
#pragma unroll
                for (int off = L / 2; off > 0; off >>= 1) {
                    T os = shfl_xor(acc.s, off);
                    T oc = KAHAN ? shfl_xor(acc.c, off) : T(0);
                    acc.merge(os, oc);
                }
            }
            if (l == 0) {
                if constexpr (KAHAN) {
                    const float2 v = accl[a0.rloc];
                    Acc<T, KAHAN> m;
                    m.s = v.x;
                    m.c = v.y;
                    m.merge(acc.s, acc.c);
                    accl[a0.rloc] = make_float2(m.s, m.c);
                } else {
                    accl[a0.rloc] += acc.s;   // Acc::merge without compensation
                }
            }
        }
#pragma unroll
        for (int i = 0; i <= DB; i++) av[i] = av[i + 1];
#pragma unroll
        for (int i = 0; i < DB; i++) bv[i] = bv[i + 1];
    }
}

/* The tiled layout swept by one workgroup per receiver block: the workgroup
 * walks the block's tiles in ascending chunk order, and each receiver's tile
 * pair is merged at once into the receiver's accumulator in LDS, in the order
 * tiled_ic merges stored partials, so the sums are the same bit for bit.  One
 * pair per receiver is stored at the end (part[block][receiver]), or, with
 * sweep_ode, the workgroup updates its block's cells itself.  For
 * networks whose blocks draw on many chunks: the per-tile partials would
 * cross HBM twice, here they stay in LDS.  LDS: R accumulators, then the
 * chunk. */
template <typename T, int L, bool KAHAN, int NT, bool HYB = false, bool SMALL = false>
__global__ __launch_bounds__(NT, 1) void k_tile_sweep(StepArgs<T> a, int64_t k)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    using P = typename PartOf<T, KAHAN>::type;
    P *accl = reinterpret_cast<P *>(dyn_lds);
    T *vch = reinterpret_cast<T *>(dyn_lds + (size_t)a.rcells * sizeof(P));   // R is a multiple of 64
    const int64_t t = (a.d_step ? *a.d_step : 0) + k;
    const T *__restrict__ vin = a.vd[t & 1];
    const int64_t rb = blockIdx.x;
    for (int r = threadIdx.x; r < a.rcells; r += NT) {
        if constexpr (KAHAN) accl[r] = make_float2(0.0f, 0.0f);
        else accl[r] = T(0);
    }
    const int32_t t0 = a.rb_tile0[rb], t1 = a.rb_tile0[rb + 1];
    // A block of fewer slices than wavefronts (small networks; the SMALL
    // instantiation) leaves wavefronts idle during the rows of its first
    // tile: with sweep_ode they
    // run the block's somata, axons and dendrites up to their voltage there
    // (none needs this step's I_c), and only the dendrite voltages follow the
    // sums, as after the stored tiles with ode_split 2.
    const int64_t cell0 = rb * a.rcells;
    const int64_t ncell = a.n - cell0 < a.rcells ? a.n - cell0 : a.rcells;
    const bool beside = SMALL && a.sweep_ode && a.rcells / (64 / L) < NT / 64 && t1 > t0;
    const auto idle = [&](int iw, int nw) {
        if (!beside) return;
        const int lane = threadIdx.x & 63;
        const int ng = (int)((ncell + 63) >> 6);   // task g < ng: somata of group g; ng + g: axons, dendrite gates
        for (int task = iw; task < 2 * ng; task += nw) {
            const int g = task < ng ? task : task - ng;
            const int64_t cell = cell0 + g * 64 + lane;
            if (cell < a.n) {
                if (a.fast_math) {
                    if (task < ng) tile_soma_cell<T, true>(a, cell, t);
                    else {
                        tile_axon_cell<T, true>(a, cell, t);
                        tile_dend_pre_cell<T, true>(a, cell, t);
                    }
                } else {
                    if (task < ng) tile_soma_cell<T, false>(a, cell, t);
                    else {
                        tile_axon_cell<T, false>(a, cell, t);
                        tile_dend_pre_cell<T, false>(a, cell, t);
                    }
                }
            }
        }
    };
    for (int32_t ti = t0; ti < t1; ti++) {
        // the last tile's rows are done with the chunk buffer (and, before
        // the first tile, the accumulators are cleared)
        __syncthreads();
        if constexpr (SMALL) {
            if (ti == t0) sweep_tile_rows<T, L, KAHAN, NT, HYB>(a, a.tiles[ti], vin, vch, accl, idle);
            else sweep_tile_rows<T, L, KAHAN, NT, HYB>(a, a.tiles[ti], vin, vch, accl, [](int, int) {});
        } else {
            sweep_tile_rows<T, L, KAHAN, NT, HYB>(a, a.tiles[ti], vin, vch, accl, [](int, int) {});
        }
    }
    __syncthreads();
    if (beside) {
        // the dendrite voltages from the merged sums (I_c as below)
        const int cur = (int)(t & 1), nxt = cur ^ 1;
        for (int r = threadIdx.x; r < ncell; r += NT) {
            const int64_t cell = cell0 + r;
            T ic;
            if constexpr (KAHAN) {
                const float2 v = accl[r];
                ic = v.x + v.y;
            } else {
                ic = accl[r];
            }
            if (a.fast_math)
                a.vd[nxt][cell] = dend_post<true>(a.m, a.vd[cur][cell], a.dpre[cell], a.dpre[a.npad + cell], ic,
                                                  iapp_of(a, cell, t));
            else
                a.vd[nxt][cell] = dend_post<false>(a.m, a.vd[cur][cell], a.dpre[cell], a.dpre[a.npad + cell], ic,
                                                   iapp_of(a, cell, t));
        }
        return;
    }
    if (a.sweep_ode) {
        // the block's cells, one thread each, with I_c = the merged pair's
        // value (what tiled_ic returns for the stored pair)
        for (int r = threadIdx.x; r < a.rcells; r += NT) {
            const int64_t cell = rb * a.rcells + r;
            if (cell >= a.n) break;
            T ic;
            if constexpr (KAHAN) {
                const float2 v = accl[r];
                ic = v.x + v.y;
            } else {
                ic = accl[r];
            }
            typename CellOf<T>::type cs;
            load_cell(a, cell, t, cs);
            if (a.fast_math) ode<true>(a.m, cs, ic, iapp_of(a, cell, t));
            else ode<false>(a.m, cs, ic, iapp_of(a, cell, t));
            store_cell(a, cell, t, cs);
        }
        return;
    }
    P *__restrict__ part = reinterpret_cast<P *>(a.part);
    for (int r = threadIdx.x; r < a.rcells; r += NT) part[rb * a.rcells + r] = accl[r];
}

/* Every cell alone for steps[cell] steps: no gap-junction current, no
 * applied current (the closed-circuit warm-up of
 * src/single_threaded/infoli.c:219-236, CompDend with randomness 1).  The
 * state stays in the buffers of step t (the warm-up precedes step t). */
template <typename T>
__global__ __launch_bounds__(kBlock) void k_isolated(StepArgs<T> a, int64_t t, const int64_t *steps)
{
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= a.n) return;
    const int64_t s = steps[cell];
    if (s <= 0) return;
    typename CellOf<T>::type c;
    load_cell(a, cell, t, c);
    for (int64_t k = 0; k < s; k++) ode(a.m, c, T(0), T(0));
    store_cell(a, cell, t + 1, c);   // step t + 1's "new" buffers are step t's
}

/* before each persistent launch: the arrival count only (a timeout of an
 * earlier launch of the same advance stays reported, and makes every later
 * barrier return at once) */
__global__ void k_reset_barrier(GridBarrier *gb) { gb->count = 0; }

/* start of an advance: the device step counter, the barrier and the flags */
__global__ void k_begin(int64_t *d_step, int64_t t, AdvanceStatus *st)
{
    *d_step = t;
    st->bar.count = 0;
    st->bar.error = 0;
    st->nonfinite = 0;
}

__global__ void k_bump_step(int64_t *d_step, int64_t by) { *d_step += by; }

/* flag[0] = 1 if any voltage of the first n cells is NaN/Inf. */
template <typename T>
__global__ void k_check_finite(const T *vd, const T *vs, const T *va, int64_t n, int *flag)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        if (!isfinite(vd[i]) || !isfinite(vs[i]) || !isfinite(va[i]))
            __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}


}  // namespace infoli
