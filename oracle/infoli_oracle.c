/*
 * infoli_oracle.c -- CPU restatement of the reference InfOli network step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against and the CPU baseline the benchmark times.  Nothing in the product
 * (infoli_amd/, include/) links, loads or calls it.
 *
 * Three arithmetics, each a restatement of reference code (paths relative to
 * the reference repository):
 *   ORACLE_FP64  -- src/single_threaded/infoli.c (mod_prec = double,
 *                   src/single_threaded/infoli.h:68): CompDend/CompSoma/
 *                   CompAxon :371-858 and IcNeighbors :544-556, including
 *                   ST's own association (G_CAH*r*r, g*k*k*k*l, pow(x,4)).
 *   ORACLE_FP32  -- src/openmp/infoli.c (mod_prec = float,
 *                   src/openmp/infoli.h:39-40): gap-junction loop :496-520
 *                   and compartment loop :548-679 with expf/powf and float
 *                   literals, in place, voltages committed last (:673-675).
 *   ORACLE_FP32K -- ORACLE_FP32 with the per-cell neighbour sum compensated:
 *                   each term enters a TwoSum (Knuth) accumulator in list
 *                   order; the result is sum + compensation.  Cells at or
 *                   above oracle_set_fp64_degree()'s budget (default: none)
 *                   take ST's FP64 term on the FP32 values, summed in FP64.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off (no FMA contraction, no
 * fast-math) so the operation order below is the order executed.
 *
 * Parity pins (tests/test_oracle_golden.py):
 *   - FP64 equals the compiled reference ST binary (oracle/_ref/st.x)
 *     bit-for-bit on the committed golden traces (10-cell complete graph,
 *     20-cell random graph with per-cell per-step Iapp);
 *   - FP32 single-cell step 1 gives V_axon = -59.97596741 (OpenMP reference,
 *     measured) and tracks src/regression_test/stable_version.c's baseline
 *     (float state, double exp) within the reference checker's 1% criterion;
 *   - gap-junction sums match a long-double brute force, are odd in
 *     V_i - V_j and conserve sum_i I_c = 0 for symmetric w;
 *   - under a degree budget, FP32K cells at or above it equal the FP64
 *     sum (float)gj64 bit for bit on float-exact inputs;
 *   - at the removable singularities of beta_r (V_dend = -8.5) and alpha_x
 *     (V = -25), where the reference computes 0/0 = NaN, the limit (0.1 and
 *     1.3) is used instead: the one-step result at the singular voltage lies
 *     between those at the neighbouring floats/doubles (continuity).
 * Every exported function is pinned by at least one of these.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORACLE_FP64  0
#define ORACLE_FP32  1
#define ORACLE_FP32K 2

#define NV 19   /* per-cell variables, bin/default.conf lines 24-42 order */
enum { V_DEND = 0, CAL_R, POT_S, HCUR_Q, CA2P, I_CAH, G_CAL, V_SOMA, NA_M,
       NA_H, POT_N, POT_P, POT_XS, CAL_K, CAL_L, V_AXON, NA_MA, NA_HA, POT_XA };

/* ------------------------------------------------------------------------ */
/* FP64: src/single_threaded/infoli.c                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    double DELTA, C_M, G_NA_S, G_KDR_S, G_K_S, G_LS, G_K_CA, G_CAH, G_LD, G_H,
           G_NA_A, G_K_A, G_LA, P1, P2, G_INT, V_NA, V_K, V_CA, V_H, V_L;
} p64;

static void load_p64(const double *c, p64 *p)
{
    /* parameter order of bin/default.conf / src/openmp/infoli.c:240-262 */
    p->DELTA = c[0]; p->C_M = c[2]; p->G_NA_S = c[3]; p->G_KDR_S = c[4];
    p->G_K_S = c[5]; p->G_LS = c[6]; p->G_K_CA = c[7]; p->G_CAH = c[8];
    p->G_LD = c[9]; p->G_H = c[10]; p->G_NA_A = c[11]; p->G_K_A = c[13];
    p->G_LA = c[14]; p->P1 = c[15]; p->P2 = c[16]; p->G_INT = c[17];
    p->V_NA = c[18]; p->V_K = c[19]; p->V_CA = c[20]; p->V_H = c[21];
    p->V_L = c[22];
}

static inline double min64(double a, double b) { return (a < b) ? a : b; } /* ST:1007-1010 */

/* IcNeighbors, ST:544-556 */
static double gj64(const double *vd, int64_t i, const int64_t *rp,
                   const int32_t *col, const double *w)
{
    double prevV = vd[i], I_c = 0;
    for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
        double V = prevV - vd[col[e]];
        double f = 0.8 * exp(-1 * pow(V, 2) / 100) + 0.2;
        double Cond = w[e];
        I_c = I_c + (Cond * f * V);
    }
    return I_c;
}

/* One cell, ST:371-858.  Reads the cell's start-of-step values, writes the
 * new ones in place (ST double-buffers prev/new states, ST:316-317; the
 * cell's update reads only its own previous state, so this is equal). */
static void cell64(const p64 *P, double **s, int64_t i, double I_c, double I_app)
{
    const double prevV_dend = s[V_DEND][i], prevV_soma = s[V_SOMA][i],
                 prevV_axon = s[V_AXON][i];

    /* DendHCurr ST:430-447 */
    double q_inf = 1 / (1 + exp((prevV_dend + 80) / 4));
    double tau_q = 1 / (exp(-0.086 * prevV_dend - 14.6) + exp(0.070 * prevV_dend - 1.87));
    double dq_dt = (q_inf - s[HCUR_Q][i]) / tau_q;
    double q = P->DELTA * dq_dt + s[HCUR_Q][i];

    /* DendCaCurr ST:449-468 */
    double alpha_r = 1.7 / (1 + exp(-(prevV_dend - 5) / 13.9));
    /* removable singularity (0/0 at V = -8.5 exactly): its limit 0.02 * 5 */
    const double xr = prevV_dend + 8.5;
    double beta_r = xr == 0 ? 0.1 : 0.02 * xr / (exp(xr / 5) - 1);
    double r_inf = alpha_r / (alpha_r + beta_r);
    double tau_r = 5 / (alpha_r + beta_r);
    double dr_dt = (r_inf - s[CAL_R][i]) / tau_r;
    double r = P->DELTA * dr_dt + s[CAL_R][i];

    /* DendKCurr ST:470-489 (reads previous Ca2Plus) */
    double alpha_s = min64((0.00002 * s[CA2P][i]), 0.01);
    double beta_s = 0.015;
    double s_inf = alpha_s / (alpha_s + beta_s);
    double tau_s = 1 / (alpha_s + beta_s);
    double ds_dt = (s_inf - s[POT_S][i]) / tau_s;
    double sK = P->DELTA * ds_dt + s[POT_S][i];

    /* DendCal ST:492-507 (previous I_CaH and Ca2Plus) */
    double dCa_dt = -3 * s[I_CAH][i] - 0.075 * s[CA2P][i];
    double Ca = P->DELTA * dCa_dt + s[CA2P][i];

    /* DendCurrVolt ST:509-542 (new q, r, s; previous voltages) */
    double I_sd = (P->G_INT / (1 - P->P1)) * (prevV_dend - prevV_soma);
    double I_CaH = P->G_CAH * r * r * (prevV_dend - P->V_CA);
    double I_K_Ca = P->G_K_CA * sK * (prevV_dend - P->V_K);
    double I_ld = P->G_LD * (prevV_dend - P->V_L);
    double I_h = P->G_H * q * (prevV_dend - P->V_H);
    double dVd_dt = (-(I_CaH + I_sd + I_ld + I_K_Ca + I_c + I_h) + I_app) / P->C_M;

    /* SomaCalcium ST:618-640 */
    double k_inf = (1 / (1 + exp(-1 * (prevV_soma + 61) / 4.2)));
    double l_inf = (1 / (1 + exp((prevV_soma + 85.5) / 8.5)));
    double tau_k = 1;
    double tau_l = ((20 * exp((prevV_soma + 160) / 30) / (1 + exp((prevV_soma + 84) / 7.3))) + 35);
    double dk_dt = (k_inf - s[CAL_K][i]) / tau_k;
    double dl_dt = (l_inf - s[CAL_L][i]) / tau_l;
    double k = P->DELTA * dk_dt + s[CAL_K][i];
    double l = P->DELTA * dl_dt + s[CAL_L][i];

    /* SomaSodium ST:642-663 */
    double m_inf = 1 / (1 + (exp((-30 - prevV_soma) / 5.5)));
    double h_inf = 1 / (1 + (exp((-70 - prevV_soma) / -5.8)));
    double tau_h = 3 * exp((-40 - prevV_soma) / 33);
    double dh_dt = (h_inf - s[NA_H][i]) / tau_h;
    double m = m_inf;
    double h = s[NA_H][i] + P->DELTA * dh_dt;

    /* SomaPotassium ST:665-688 */
    double n_inf = 1 / (1 + exp((-3 - prevV_soma) / 10));
    double p_inf = 1 / (1 + exp((-51 - prevV_soma) / -12));
    double tau_n = 5 + (47 * exp(-(-50 - prevV_soma) / 900));
    double tau_p = tau_n;
    double dn_dt = (n_inf - s[POT_N][i]) / tau_n;
    double dp_dt = (p_inf - s[POT_P][i]) / tau_p;
    double n = P->DELTA * dn_dt + s[POT_N][i];
    double p = P->DELTA * dp_dt + s[POT_P][i];

    /* SomaPotassiumX ST:690-709 (literal 0.05, not DELTA) */
    /* removable singularity (0/0 at V = -25 exactly): its limit 0.13 * 10 */
    const double xs = prevV_soma + 25;
    double alpha_x_s = xs == 0 ? 1.3 : 0.13 * xs / (1 - exp(-xs / 10));
    double beta_x_s = 1.69 * exp(-0.0125 * (prevV_soma + 35));
    double x_inf_s = alpha_x_s / (alpha_x_s + beta_x_s);
    double tau_x_s = 1 / (alpha_x_s + beta_x_s);
    double dx_dt_s = (x_inf_s - s[POT_XS][i]) / tau_x_s;
    double x_s = 0.05 * dx_dt_s + s[POT_XS][i];

    /* SomaCurrVolt ST:711-749 */
    double g_CaL = s[G_CAL][i];
    double I_ds = (P->G_INT / P->P1) * (prevV_soma - prevV_dend);
    double I_CaL = g_CaL * k * k * k * l * (prevV_soma - P->V_CA);
    double I_Na_s = P->G_NA_S * m * m * m * h * (prevV_soma - P->V_NA);
    double I_ls = P->G_LS * (prevV_soma - P->V_L);
    double I_Kdr_s = P->G_KDR_S * n * n * n * n * (prevV_soma - P->V_K);
    double I_K_s = P->G_K_S * pow(x_s, 4) * (prevV_soma - P->V_K);
    double I_as = (P->G_INT / (1 - P->P2)) * (prevV_soma - prevV_axon);
    double dVs_dt = (-(I_CaL + I_ds + I_as + I_Na_s + I_ls + I_Kdr_s + I_K_s)) / P->C_M;

    /* AxonSodium ST:788-810 */
    double m_inf_a = 1 / (1 + (exp((-30 - prevV_axon) / 5.5)));
    double h_inf_a = 1 / (1 + (exp((-60 - prevV_axon) / -5.8)));
    double tau_h_a = 1.5 * exp((-40 - prevV_axon) / 33);
    double dh_dt_a = (h_inf_a - s[NA_HA][i]) / tau_h_a;
    double m_a = m_inf_a;
    double h_a = s[NA_HA][i] + P->DELTA * dh_dt_a;

    /* AxonPotassium ST:812-831 */
    const double xa = prevV_axon + 25;
    double alpha_x_a = xa == 0 ? 1.3 : 0.13 * xa / (1 - exp(-xa / 10));
    double beta_x_a = 1.69 * exp(-0.0125 * (prevV_axon + 35));
    double x_inf_a = alpha_x_a / (alpha_x_a + beta_x_a);
    double tau_x_a = 1 / (alpha_x_a + beta_x_a);
    double dx_dt_a = (x_inf_a - s[POT_XA][i]) / tau_x_a;
    double x_a = 0.05 * dx_dt_a + s[POT_XA][i];

    /* AxonCurrVolt ST:833-858 */
    double I_Na_a = P->G_NA_A * m_a * m_a * m_a * h_a * (prevV_axon - P->V_NA);
    double I_la = P->G_LA * (prevV_axon - P->V_L);
    double I_sa = (P->G_INT / P->P2) * (prevV_axon - prevV_soma);
    double I_K_a = P->G_K_A * pow(x_a, 4) * (prevV_axon - P->V_K);
    double dVa_dt = (-(I_K_a + I_sa + I_la + I_Na_a)) / P->C_M;

    s[HCUR_Q][i] = q; s[CAL_R][i] = r; s[POT_S][i] = sK; s[CA2P][i] = Ca;
    s[I_CAH][i] = I_CaH;
    s[CAL_K][i] = k; s[CAL_L][i] = l; s[NA_M][i] = m; s[NA_H][i] = h;
    s[POT_N][i] = n; s[POT_P][i] = p; s[POT_XS][i] = x_s;
    s[NA_MA][i] = m_a; s[NA_HA][i] = h_a; s[POT_XA][i] = x_a;
    s[V_DEND][i] = P->DELTA * dVd_dt + prevV_dend;
    s[V_SOMA][i] = P->DELTA * dVs_dt + prevV_soma;
    s[V_AXON][i] = P->DELTA * dVa_dt + prevV_axon;
}

/* ------------------------------------------------------------------------ */
/* FP32: src/openmp/infoli.c                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    float DELTA, C_M, G_NA_S, G_KDR_S, G_K_S, G_LS, G_K_CA, G_CAH, G_LD, G_H,
          G_NA_A, G_K_A, G_LA, P1, P2, G_INT, V_NA, V_K, V_CA, V_H, V_L;
} p32;

static void load_p32(const float *c, p32 *p)
{
    p->DELTA = c[0]; p->C_M = c[2]; p->G_NA_S = c[3]; p->G_KDR_S = c[4];
    p->G_K_S = c[5]; p->G_LS = c[6]; p->G_K_CA = c[7]; p->G_CAH = c[8];
    p->G_LD = c[9]; p->G_H = c[10]; p->G_NA_A = c[11]; p->G_K_A = c[13];
    p->G_LA = c[14]; p->P1 = c[15]; p->P2 = c[16]; p->G_INT = c[17];
    p->V_NA = c[18]; p->V_K = c[19]; p->V_CA = c[20]; p->V_H = c[21];
    p->V_L = c[22];
}

static inline float min32(float a, float b) { return (a < b) ? a : b; } /* utils.c:340-343 */

/* gap-junction loop body, OMP:507-519 */
static float gj32(const float *vd, int64_t i, const int64_t *rp,
                  const int32_t *col, const float *w)
{
    float I_c_storage = 0;
    for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
        float voltage = vd[col[e]];
        float V = vd[i] - voltage;
        float f = 0.8f * expf(-1 * powf(V, 2) / 100) + 0.2f;
        I_c_storage += w[e] * f * V;
    }
    return I_c_storage;
}

/* the same terms, compensated: TwoSum(sum, term) per entry */
static float gj32k(const float *vd, int64_t i, const int64_t *rp,
                   const int32_t *col, const float *w)
{
    float sum = 0, comp = 0;
    for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
        float voltage = vd[col[e]];
        float V = vd[i] - voltage;
        float f = 0.8f * expf(-1 * powf(V, 2) / 100) + 0.2f;
        float term = w[e] * f * V;
        float t = sum + term;
        float bp = t - sum;
        float err = (sum - (t - bp)) + (term - bp);
        comp += err;
        sum = t;
    }
    return sum + comp;
}

/* Cells with at least this many neighbours take FP64 terms and an FP64 sum
 * in ORACLE_FP32K (0 = none): the product's degree budget, restated as
 * ST's FP64 term (ST:548-553) applied to the FP32 voltages and the FP32
 * conductances, summed in list order and rounded to float. */
static int64_t g_fp64_degree = 0;

void oracle_set_fp64_degree(int64_t d) { g_fp64_degree = d < 0 ? 0 : d; }

static float gj32w(const float *vd, int64_t i, const int64_t *rp,
                   const int32_t *col, const float *w)
{
    double sum = 0;
    for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
        double V = (double)vd[i] - (double)vd[col[e]];
        sum += (double)w[e] * (0.8 * exp(-1 * (V * V) / 100) + 0.2) * V;
    }
    return (float)sum;
}

/* ORACLE_FP32K neighbour current: compensated FP32, or FP64 at or above the
 * degree budget */
static float gj32k_cell(const float *vd, int64_t i, const int64_t *rp,
                        const int32_t *col, const float *w)
{
    if (g_fp64_degree > 0 && rp[i + 1] - rp[i] >= g_fp64_degree)
        return gj32w(vd, i, rp, col, w);
    return gj32k(vd, i, rp, col, w);
}

/* compartment loop body, OMP:556-676, one cell, in place */
static void cell32(const p32 *P, float **s, int64_t i, float I_c, float iApp)
{
    float *V_dend = s[V_DEND], *V_soma = s[V_SOMA], *V_axon = s[V_AXON];
    float q_inf, tau_q, dq_dt, alpha_r, beta_r, r_inf, tau_r, dr_dt, xr, xs, xa;
    float alpha_s, beta_s, s_inf, tau_s, ds_dt, dCa_dt;
    float I_sd, I_CaH_temp, I_K_Ca, I_ld, I_h, dVd_dt;
    float k_inf, l_inf, tau_k, tau_l, dk_dt, dl_dt, m_inf, h_inf, tau_h, dh_dt;
    float n_inf, p_inf, tau_n, tau_p, dn_dt, dp_dt;
    float alpha_x_s, beta_x_s, x_inf_s, tau_x_s, dx_dt_s;
    float I_ds, I_CaL, I_Na_s, I_ls, I_Kdr_s, I_K_s, I_as, dVs_dt;
    float m_inf_a, h_inf_a, tau_h_a, dh_dt_a, alpha_x_a, beta_x_a, x_inf_a,
          tau_x_a, dx_dt_a, I_Na_a, I_la, I_sa, I_K_a, dVa_dt;

    /* Dend H current OMP:561-564 */
    q_inf = 1 / (1 + expf((V_dend[i] + 80) / 4));
    tau_q = 1 / (expf(-0.086f * V_dend[i] - 14.6f) + expf(0.070f * V_dend[i] - 1.87f));
    dq_dt = (q_inf - s[HCUR_Q][i]) / tau_q;
    s[HCUR_Q][i] = P->DELTA * dq_dt + s[HCUR_Q][i];

    /* Dend Ca current OMP:567-572 */
    alpha_r = 1.7f / (1 + expf(-(V_dend[i] - 5) / 13.9f));
    xr = V_dend[i] + 8.5f;   /* 0/0 at -8.5 exactly: the limit 0.02 * 5 */
    beta_r = xr == 0 ? 0.1f : 0.02f * xr / (expf(xr / 5) - 1);
    r_inf = alpha_r / (alpha_r + beta_r);
    tau_r = 5 / (alpha_r + beta_r);
    dr_dt = (r_inf - s[CAL_R][i]) / tau_r;
    s[CAL_R][i] = P->DELTA * dr_dt + s[CAL_R][i];

    /* Dend K current OMP:575-580 */
    alpha_s = min32((0.00002f * s[CA2P][i]), 0.01f);
    beta_s = 0.015f;
    s_inf = alpha_s / (alpha_s + beta_s);
    tau_s = 1 / (alpha_s + beta_s);
    ds_dt = (s_inf - s[POT_S][i]) / tau_s;
    s[POT_S][i] = P->DELTA * ds_dt + s[POT_S][i];

    /* Dend Cal OMP:583-584 */
    dCa_dt = -3 * s[I_CAH][i] - 0.075f * s[CA2P][i];
    s[CA2P][i] = P->DELTA * dCa_dt + s[CA2P][i];

    /* Dendritic voltage and currents OMP:587-593 */
    I_sd = (P->G_INT / (1 - P->P1)) * (V_dend[i] - V_soma[i]);
    I_CaH_temp = P->G_CAH * powf(s[CAL_R][i], 2) * (V_dend[i] - P->V_CA);
    I_K_Ca = P->G_K_CA * s[POT_S][i] * (V_dend[i] - P->V_K);
    I_ld = P->G_LD * (V_dend[i] - P->V_L);
    I_h = P->G_H * s[HCUR_Q][i] * (V_dend[i] - P->V_H);
    dVd_dt = (-(I_CaH_temp + I_sd + I_ld + I_K_Ca + I_c + I_h) + iApp) / P->C_M;
    s[I_CAH][i] = I_CaH_temp;

    /* Soma calcium OMP:599-606 */
    k_inf = (1 / (1 + expf(-1 * (V_soma[i] + 61) / 4.2f)));
    l_inf = (1 / (1 + expf((V_soma[i] + 85.5f) / 8.5f)));
    tau_k = 1;
    tau_l = ((20 * expf((V_soma[i] + 160) / 30) / (1 + expf((V_soma[i] + 84) / 7.3f))) + 35);
    dk_dt = (k_inf - s[CAL_K][i]) / tau_k;
    dl_dt = (l_inf - s[CAL_L][i]) / tau_l;
    s[CAL_K][i] = P->DELTA * dk_dt + s[CAL_K][i];
    s[CAL_L][i] = P->DELTA * dl_dt + s[CAL_L][i];

    /* Soma sodium OMP:609-614 */
    m_inf = 1 / (1 + (expf((-30 - V_soma[i]) / 5.5f)));
    h_inf = 1 / (1 + (expf((-70 - V_soma[i]) / -5.8f)));
    tau_h = 3 * expf((-40 - V_soma[i]) / 33);
    dh_dt = (h_inf - s[NA_H][i]) / tau_h;
    s[NA_M][i] = m_inf;
    s[NA_H][i] = s[NA_H][i] + P->DELTA * dh_dt;

    /* Soma potassium OMP:617-624 */
    n_inf = 1 / (1 + expf((-3 - V_soma[i]) / 10));
    p_inf = 1 / (1 + expf((-51 - V_soma[i]) / -12));
    tau_n = 5 + (47 * expf(-(-50 - V_soma[i]) / 900));
    tau_p = tau_n;
    dn_dt = (n_inf - s[POT_N][i]) / tau_n;
    dp_dt = (p_inf - s[POT_P][i]) / tau_p;
    s[POT_N][i] = P->DELTA * dn_dt + s[POT_N][i];
    s[POT_P][i] = P->DELTA * dp_dt + s[POT_P][i];

    /* Soma potassium X OMP:627-632 (literal 0.05f) */
    xs = V_soma[i] + 25;     /* 0/0 at -25 exactly: the limit 0.13 * 10 */
    alpha_x_s = xs == 0 ? 1.3f : 0.13f * xs / (1 - expf(-xs / 10));
    beta_x_s = 1.69f * expf(-0.0125f * (V_soma[i] + 35));
    x_inf_s = alpha_x_s / (alpha_x_s + beta_x_s);
    tau_x_s = 1 / (alpha_x_s + beta_x_s);
    dx_dt_s = (x_inf_s - s[POT_XS][i]) / tau_x_s;
    s[POT_XS][i] = 0.05f * dx_dt_s + s[POT_XS][i];

    /* Somatic voltage and currents OMP:635-642 */
    I_ds = (P->G_INT / P->P1) * (V_soma[i] - V_dend[i]);
    I_CaL = s[G_CAL][i] * powf(s[CAL_K][i], 3) * s[CAL_L][i] * (V_soma[i] - P->V_CA);
    I_Na_s = P->G_NA_S * powf(s[NA_M][i], 3) * s[NA_H][i] * (V_soma[i] - P->V_NA);
    I_ls = P->G_LS * (V_soma[i] - P->V_L);
    I_Kdr_s = P->G_KDR_S * powf(s[POT_N][i], 4) * (V_soma[i] - P->V_K);
    I_K_s = P->G_K_S * powf(s[POT_XS][i], 4) * (V_soma[i] - P->V_K);
    I_as = (P->G_INT / (1 - P->P2)) * (V_soma[i] - V_axon[i]);
    dVs_dt = (-(I_CaL + I_ds + I_as + I_Na_s + I_ls + I_Kdr_s + I_K_s)) / P->C_M;

    /* Axon sodium OMP:648-653 */
    m_inf_a = 1 / (1 + (expf((-30 - V_axon[i]) / 5.5f)));
    h_inf_a = 1 / (1 + (expf((-60 - V_axon[i]) / -5.8f)));
    tau_h_a = 1.5f * expf((-40 - V_axon[i]) / 33);
    dh_dt_a = (h_inf_a - s[NA_HA][i]) / tau_h_a;
    s[NA_MA][i] = m_inf_a;
    s[NA_HA][i] = s[NA_HA][i] + P->DELTA * dh_dt_a;

    /* Axon potassium OMP:656-661 */
    xa = V_axon[i] + 25;
    alpha_x_a = xa == 0 ? 1.3f : 0.13f * xa / (1 - expf(-xa / 10));
    beta_x_a = 1.69f * expf(-0.0125f * (V_axon[i] + 35));
    x_inf_a = alpha_x_a / (alpha_x_a + beta_x_a);
    tau_x_a = 1 / (alpha_x_a + beta_x_a);
    dx_dt_a = (x_inf_a - s[POT_XA][i]) / tau_x_a;
    s[POT_XA][i] = 0.05f * dx_dt_a + s[POT_XA][i];

    /* Axonal voltage and currents OMP:664-668 */
    I_Na_a = P->G_NA_A * powf(s[NA_MA][i], 3) * s[NA_HA][i] * (V_axon[i] - P->V_NA);
    I_la = P->G_LA * (V_axon[i] - P->V_L);
    I_sa = (P->G_INT / P->P2) * (V_axon[i] - V_soma[i]);
    I_K_a = P->G_K_A * powf(s[POT_XA][i], 4) * (V_axon[i] - P->V_K);
    dVa_dt = (-(I_K_a + I_sa + I_la + I_Na_a)) / P->C_M;

    /* new voltages OMP:673-675 */
    V_dend[i] = P->DELTA * dVd_dt + V_dend[i];
    V_soma[i] = P->DELTA * dVs_dt + V_soma[i];
    V_axon[i] = P->DELTA * dVa_dt + V_axon[i];
}

/* ------------------------------------------------------------------------ */
/* driver loop (OMP:480-698 structure: GJ loop, then compartment loop)      */
/* ------------------------------------------------------------------------ */

/* Iapp of cell i at step t: table row if t is in the table, else
 * base + pulse (the reference: 6 on steps [20000, 20499), OMP:482-485). */
static inline double iapp_at(int64_t i, int64_t t, int64_t n,
                             const double *base, const uint8_t *mask,
                             double amp, int64_t t0, int64_t t1,
                             int64_t tab_t0, int64_t tab_n, const double *tab,
                             int *from_table)
{
    if (tab && t >= tab_t0 && t < tab_t0 + tab_n) {
        *from_table = 1;
        return tab[(t - tab_t0) * n + i];
    }
    *from_table = 0;
    return 0; /* caller composes base + pulse in its own precision */
}

int oracle_run(int prec, const double *conf, const float *conff, int64_t n,
               const int64_t *rp, const int32_t *col, const double *w,
               const double *iapp_base, const uint8_t *mask, double amp,
               int64_t t0, int64_t t1, int64_t tab_t0, int64_t tab_n,
               const double *tab, double *state, int64_t start_step,
               int64_t nsteps, int64_t stride, double *trace, int nthreads,
               double *elapsed_s)
{
    if (n <= 0 || !state || nsteps < 0 || (trace && stride <= 0)) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t nnz = rp ? rp[n] : 0;
    int64_t rec = 0;
    double tic, toc;

    if (prec == ORACLE_FP64) {
        p64 P; load_p64(conf, &P);
        double *s[NV];
        for (int v = 0; v < NV; v++) s[v] = state + (int64_t)v * n;
        double *ic = (double *)malloc(sizeof(double) * n);
        if (!ic) return -3;
        tic = omp_get_wtime();
        for (int64_t k = 0; k < nsteps; k++) {
            int64_t t = start_step + k;
            int in_pulse_t = (t >= t0 && t < t1);
            #pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t i = 0; i < n; i++)
                ic[i] = (nnz > 0) ? gj64(s[V_DEND], i, rp, col, w) : 0.0;
            #pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t i = 0; i < n; i++) {
                int ft;
                double ia = iapp_at(i, t, n, iapp_base, mask, amp, t0, t1,
                                    tab_t0, tab_n, tab, &ft);
                if (!ft) {
                    double b = iapp_base ? iapp_base[i] : 0.0;
                    ia = (in_pulse_t && (!mask || mask[i])) ? b + amp : b;
                }
                cell64(&P, s, i, ic[i], ia);
            }
            if (trace && t % stride == 0) {
                memcpy(trace + rec * n, s[V_AXON], sizeof(double) * n);
                rec++;
            }
        }
        toc = omp_get_wtime();
        free(ic);
    } else {
        p32 P; load_p32(conff, &P);
        float *buf = (float *)malloc(sizeof(float) * (size_t)n * (NV + 2));
        float *wf = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
        if (!buf || !wf) { free(buf); free(wf); return -3; }
        float *s[NV];
        for (int v = 0; v < NV; v++) {
            s[v] = buf + (int64_t)v * n;
            for (int64_t i = 0; i < n; i++) s[v][i] = (float)state[(int64_t)v * n + i];
        }
        float *ic = buf + (int64_t)NV * n;
        float *basef = buf + (int64_t)(NV + 1) * n;
        for (int64_t i = 0; i < n; i++) basef[i] = iapp_base ? (float)iapp_base[i] : 0.0f;
        for (int64_t e = 0; e < nnz; e++) wf[e] = (float)w[e];
        float ampf = (float)amp;
        tic = omp_get_wtime();
        for (int64_t k = 0; k < nsteps; k++) {
            int64_t t = start_step + k;
            int in_pulse_t = (t >= t0 && t < t1);
            if (prec == ORACLE_FP32K) {
                #pragma omp parallel for schedule(static) num_threads(nthreads)
                for (int64_t i = 0; i < n; i++)
                    ic[i] = (nnz > 0) ? gj32k_cell(s[V_DEND], i, rp, col, wf) : 0.0f;
            } else {
                #pragma omp parallel for schedule(static) num_threads(nthreads)
                for (int64_t i = 0; i < n; i++)
                    ic[i] = (nnz > 0) ? gj32(s[V_DEND], i, rp, col, wf) : 0.0f;
            }
            #pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t i = 0; i < n; i++) {
                int ft;
                double iad = iapp_at(i, t, n, iapp_base, mask, amp, t0, t1,
                                     tab_t0, tab_n, tab, &ft);
                float ia;
                if (ft) ia = (float)iad;
                else ia = (in_pulse_t && (!mask || mask[i])) ? basef[i] + ampf : basef[i];
                cell32(&P, s, i, ic[i], ia);
            }
            if (trace && t % stride == 0) {
                for (int64_t i = 0; i < n; i++) trace[rec * n + i] = (double)s[V_AXON][i];
                rec++;
            }
        }
        toc = omp_get_wtime();
        for (int v = 0; v < NV; v++)
            for (int64_t i = 0; i < n; i++) state[(int64_t)v * n + i] = (double)s[v][i];
        free(buf); free(wf);
    }
    if (elapsed_s) *elapsed_s = toc - tic;
    return 0;
}

/* I_c of every cell from V_dend, in the given arithmetic. */
int oracle_gap_currents(int prec, int64_t n, const int64_t *rp,
                        const int32_t *col, const double *w,
                        const double *vdend, double *ic)
{
    if (n <= 0 || !rp || !vdend || !ic) return -1;
    int64_t nnz = rp[n];
    if (prec == ORACLE_FP64) {
        for (int64_t i = 0; i < n; i++) ic[i] = gj64(vdend, i, rp, col, w);
        return 0;
    }
    float *vf = (float *)malloc(sizeof(float) * n);
    float *wf = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    if (!vf || !wf) { free(vf); free(wf); return -3; }
    for (int64_t i = 0; i < n; i++) vf[i] = (float)vdend[i];
    for (int64_t e = 0; e < nnz; e++) wf[e] = (float)w[e];
    for (int64_t i = 0; i < n; i++)
        ic[i] = (prec == ORACLE_FP32K) ? gj32k_cell(vf, i, rp, col, wf)
                                       : gj32(vf, i, rp, col, wf);
    free(vf); free(wf);
    return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
