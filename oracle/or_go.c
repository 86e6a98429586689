/*
 * Oracle: Golgi cell model on the RKC path (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Restates:
 *   constants                 go.cuh:18-53
 *   initialize_host_go        go.cu:3-88     (Go_131.txt: 20 columns, densities x area, Cm = Cm' x area)
 *   set_host_go_solve         go_solve.cu:4-215 (cable conductances, grid gap junctions on
 *                                               compartment 0, CSR with sorted GJ inserts, RKC sign)
 *   rate functions            go_ion.cu:3-86
 *   go_KAHP_update_2order     go_ion.cu:88-213 (6-state KAHP, Crank-Nicolson, 6 x 6 partial-pivot solve)
 *   go_update_ion_RKC_exp_imp go_ion.cu:379-437 (two Ca pools by RK4, Nernst reversal of the
 *                                               second pool at 23 C, CN gates at the mid-step V)
 *   go_initialize_ion         go_ion.cu:590-634
 *   go_update_matrix          solve_rkc.cu:288-314
 *   go_update_ion_and_matrix  solve_rkc.cu:1163-1182 (ion update, KAHP, then b)
 *   go_set_current            go.cu:258-268
 *   go_update_ion_exp_imp     go_ion.cu:319-377 (CN path)
 *   CN conductance sums       solve_cnm.cu:259-362
 *
 * Reference quirks kept: the Ca currents of the update divide the stored
 * conductance (density x area) by the area again; the second pool's current
 * uses the reversal from before this update; the KAHP scheme's explicit half
 * uses the Ca of the previous ion update (ca_old) and its implicit half the
 * new one; the initial reversal of the second pool is V_Ca (120 mV), not its
 * Nernst value; rows of the compartment table are connected by line number
 * (go_solve.cu:37) while geometry is indexed by the file's first column.
 * Without granule-cell synapses the reference's b += v_new term
 * (solve_rkc.cu:280-287, v_new reset to zero at each macro step) adds zero
 * and is left out.
 *
 * Pinning (tests/test_oracle_go.py): rates against their alpha/(alpha+beta)
 * identities and closed-form limits, the HCN2 ratio's clamps, CN gates
 * against the (1,1) Pade form, both Ca RK4 updates against the exact
 * exponential, the reversal against the Nernst constant, the KAHP step
 * against an independent dense solve of the same Crank-Nicolson system and
 * its conservation of total occupancy, the cable matrix against the series
 * half-segment conductances from Go_131.txt geometry, and the gap-junction
 * entries (count, symmetry, value) against the grid.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GO_COMP 131
#define GO_COMP_DEND 10
#define V_INIT_GO (-60.0)
#define V_LEAK_GO (-55.0)
#define V_Na_GO (87.39)
#define V_K_GO (-84.69)
#define V_Ca_GO (120.0)
#define V_H_GO (-20.0)
#define SHELL1_D_GO (0.2e-4)
#define B_Ca1_GO (1.3)
#define Ca1_0_GO (5e-5)
#define Ca1OUT_GO (2.0)
#define F_GO (9.6485e4)
#define G_GJ_GO (2.5e-6)

/* ion order (go.cuh:46-50) */
enum { m_NaT, h_NaT, r_NaR, s_NaR, p_NaP, n_KV, a_KA, b_KA, c_KC, sl_Kslow, ch_CaHVA, ci_CaHVA, cl_CaLVA,
       cm_CaLVA, hf_HCN1, hf_HCN2, hs_HCN1, hs_HCN2, o1_KAHP, o2_KAHP, c1_KAHP, c2_KAHP, c3_KAHP, c4_KAHP,
       N_ION_GO };
/* conductance order (go.cuh:52-54) */
enum { g_leak, g_NaT, g_NaR, g_NaP, g_KV, g_KA, g_KC, g_Kslow, g_CaHVA, g_CaLVA, g_HCN1, g_HCN2, g_KAHP,
       N_COND_GO };

/* ---------------- rate functions, go_ion.cu:3-86 ---------------- */
static double a_m_NaT(double v) { return 0.417 * (v + 25.0) / (1 - exp(-(v + 25.0) / 10.0)); }
static double b_m_NaT(double v) { return 16.68 * exp(-0.055 * (v + 50.0)); }
static double a_h_NaT(double v) { return 0.292 * exp(-0.3 * (v + 50.0)); }
static double b_h_NaT(double v) { return 4.17 / (1 + exp(-(v + 17.0) / 5.0)); }
static double a_r_NaR(double v) { return ((1.11 - 68.5 * (v - 4.48) / (exp(-(v - 4.48) / 6.8) - 1.0)) * 1.0e-4); }
static double b_r_NaR(double v)
{
  double x = (v + 44.0) / 0.11;
  if (x > 200.0) x = 200.0;
  return ((66.0 + 21.7 * (v + 44) / (exp(x) - 1.0)) * 1.0e-3);
}
static double a_s_NaR(double v) { return (0.443 * exp(-(v + 80.0) / 62.5)); }
static double b_s_NaR(double v) { return (0.014 * exp((v + 83.3) / 16.1)); }
static double a_p_NaP(double v) { return (0.421 * (v + 40.0) / (1.0 - exp(-(v + 40.0) / 5.0))); }
static double b_p_NaP(double v) { return (-0.287 * (v + 40.0) / (1.0 - exp((v + 40.0) / 5.0))); }
static double a_ch_CaHVA(double v) { return (0.0687 * exp(0.063 * (v + 29.0))); }
static double b_ch_CaHVA(double v) { return (0.115 * exp(-0.039 * (v + 18.66))); }
static double a_ci_CaHVA(double v) { return (1.8e-3 * exp(-(v + 48.0) / 18.0)); }
static double b_ci_CaHVA(double v) { return (1.8e-3 * exp((v + 48.0) / 83.0)); }
static double a_n_KV(double v) { return 0.062 * (v + 26.0) / (1.0 - exp(-(v + 26.0) / 10.0)); }
static double b_n_KV(double v) { return 0.78 * exp(-(v + 36.0) / 80.0); }
static double a_a_KA(double v) { return (0.62 / (1.0 + exp(-(v + 9.17) / 23.32))); }
static double b_a_KA(double v) { return (0.126 / (exp((v + 18.28) / 19.47))); }
static double a_b_KA(double v) { return (0.028 / (1.0 + exp((v + 111.0) / 12.84))); }
static double b_b_KA(double v) { return (0.026 / (1.0 + exp(-(v + 49.95) / 8.9))); }
static double a_c_KC(double v, double ca) { return (3.2 / (1.0 + 0.0015 * exp(-(v) / 11.7) / ca)); }
static double b_c_KC(double v, double ca) { return (0.46 / (1.0 + ca / (1.5e-4 * exp(-(v) / 11.7)))); }
static double a_sl_Kslow(double v) { return (0.0037 * exp((v + 30.0) / 40.0)); }
static double b_sl_Kslow(double v) { return (0.0037 * exp(-(v + 30.0) / 20.0)); }
static double r_HCN1(double v) { return (0.0021 * (v) + 0.97); }
static double r_HCN2(double v)
{
  if (v >= -64.70) return 0.0;
  else if (v <= -108.70) return 1.0;
  else return (-0.0227 * (v) - 1.47);
}

static double inf_m_NaT(double v) { return a_m_NaT(v) / (a_m_NaT(v) + b_m_NaT(v)); }
static double tau_m_NaT(double v) { return 1.0 / (a_m_NaT(v) + b_m_NaT(v)); }
static double inf_h_NaT(double v) { return a_h_NaT(v) / (a_h_NaT(v) + b_h_NaT(v)); }
static double tau_h_NaT(double v) { return 1.0 / (a_h_NaT(v) + b_h_NaT(v)); }
static double inf_r_NaR(double v) { return a_r_NaR(v) / (a_r_NaR(v) + b_r_NaR(v)); }
static double tau_r_NaR(double v) { return 1.0 / (a_r_NaR(v) + b_r_NaR(v)); }
static double inf_s_NaR(double v) { return a_s_NaR(v) / (a_s_NaR(v) + b_s_NaR(v)); }
static double tau_s_NaR(double v) { return 1.0 / (a_s_NaR(v) + b_s_NaR(v)); }
static double inf_p_NaP(double v) { return (1 / (1.0 + exp(-(v + 43.0) / 5.0))); }
static double tau_p_NaP(double v) { return (5.0 / (a_p_NaP(v) + b_p_NaP(v))); }
static double inf_ch_CaHVA(double v) { return (a_ch_CaHVA(v) / (a_ch_CaHVA(v) + b_ch_CaHVA(v))); }
static double tau_ch_CaHVA(double v) { return (1.0 / (a_ch_CaHVA(v) + b_ch_CaHVA(v))); }
static double inf_ci_CaHVA(double v) { return (a_ci_CaHVA(v) / (a_ci_CaHVA(v) + b_ci_CaHVA(v))); }
static double tau_ci_CaHVA(double v) { return (1.0 / (a_ci_CaHVA(v) + b_ci_CaHVA(v))); }
static double inf_cl_CaLVA(double v) { return (1.0 / (1.0 + exp(-(v + 52.0) / 7.4))); }
static double tau_cl_CaLVA(double v) { return ((3.0 + 1.0 / (exp((v + 27.0) / 10.0) + exp(-(v + 102.0) / 15.0))) / 0.85); }
static double inf_cm_CaLVA(double v) { return (1.0 / (1.0 + exp((v + 80.0) / 5.0))); }
static double tau_cm_CaLVA(double v) { return ((85.0 + 1.0 / (exp((v + 48.0) / 4.0) + exp(-(v + 407.0) / 50.0))) / 0.9); }
static double inf_n_KV(double v) { return (a_n_KV(v) / (a_n_KV(v) + b_n_KV(v))); }
static double tau_n_KV(double v) { return (1.0 / (a_n_KV(v) + b_n_KV(v))); }
static double inf_a_KA(double v) { return (1.0 / (1.0 + exp(-(v + 38.0) / 17.0))); }
static double tau_a_KA(double v) { return (1.0 / (a_a_KA(v) + b_a_KA(v))); }
static double inf_b_KA(double v) { return (1.0 / (1.0 + exp((v + 78.8) / 8.4))); }
static double tau_b_KA(double v) { return (1.0 / (a_b_KA(v) + b_b_KA(v))); }
static double inf_c_KC(double v, double ca) { return (a_c_KC(v, ca) / (a_c_KC(v, ca) + b_c_KC(v, ca))); }
static double tau_c_KC(double v, double ca) { return (1.0 / (a_c_KC(v, ca) + b_c_KC(v, ca))); }
static double inf_sl_Kslow(double v) { return (1.0 / (1.0 + exp(-(v + 35.0) / 6.0))); }
static double tau_sl_Kslow(double v) { return (1.0 / (a_sl_Kslow(v) + b_sl_Kslow(v))); }
static double inf_hf_HCN1(double v) { return (r_HCN1(v) * (1.0 / (1.0 + exp((v + 72.5) * 0.11)))); }
static double inf_hs_HCN1(double v) { return ((1.0 - r_HCN1(v)) * (1.0 / (1.0 + exp((v + 72.5) * 0.11)))); }
static double tau_hf_HCN1(double v) { return (exp((0.0137 * v + 3.37) * 2.3)); }
static double tau_hs_HCN1(double v) { return (exp((0.0145 * v + 4.06) * 2.3)); }
static double inf_hf_HCN2(double v) { return (r_HCN2(v) * (1.0 / (1.0 + exp((v + 81.9) * 0.16)))); }
static double inf_hs_HCN2(double v) { return ((1.0 - r_HCN2(v)) * (1.0 / (1.0 + exp((v + 81.9) * 0.16)))); }
static double tau_hf_HCN2(double v) { return (exp((0.027 * v + 5.6) * 2.3)); }
static double tau_hs_HCN2(double v) { return (exp((0.015 * v + 5.3) * 2.3)); }

/* exported for the closed-form tests: which = 2 * gate (+1 for tau), gates in
   the ion order up to hs_HCN2; ca is used by c_KC.  which = 1000 + k returns
   the k-th alpha/beta (k = 2 * gate + (beta)) for the gates that have them. */
double or_go_rate(int which, double v, double ca)
{
  if (which >= 1000) {
    const int k = which - 1000, g = k / 2, b = k % 2;
    switch (g) {
      case m_NaT: return b ? b_m_NaT(v) : a_m_NaT(v);
      case h_NaT: return b ? b_h_NaT(v) : a_h_NaT(v);
      case r_NaR: return b ? b_r_NaR(v) : a_r_NaR(v);
      case s_NaR: return b ? b_s_NaR(v) : a_s_NaR(v);
      case p_NaP: return b ? b_p_NaP(v) : a_p_NaP(v);
      case n_KV: return b ? b_n_KV(v) : a_n_KV(v);
      case a_KA: return b ? b_a_KA(v) : a_a_KA(v);
      case b_KA: return b ? b_b_KA(v) : a_b_KA(v);
      case c_KC: return b ? b_c_KC(v, ca) : a_c_KC(v, ca);
      case sl_Kslow: return b ? b_sl_Kslow(v) : a_sl_Kslow(v);
      case ch_CaHVA: return b ? b_ch_CaHVA(v) : a_ch_CaHVA(v);
      case ci_CaHVA: return b ? b_ci_CaHVA(v) : a_ci_CaHVA(v);
      case hf_HCN1: return b ? 1.0 - r_HCN1(v) : r_HCN1(v);
      case hf_HCN2: return b ? 1.0 - r_HCN2(v) : r_HCN2(v);
    }
    return NAN;
  }
  const int g = which / 2, t = which % 2;
  switch (g) {
    case m_NaT: return t ? tau_m_NaT(v) : inf_m_NaT(v);
    case h_NaT: return t ? tau_h_NaT(v) : inf_h_NaT(v);
    case r_NaR: return t ? tau_r_NaR(v) : inf_r_NaR(v);
    case s_NaR: return t ? tau_s_NaR(v) : inf_s_NaR(v);
    case p_NaP: return t ? tau_p_NaP(v) : inf_p_NaP(v);
    case n_KV: return t ? tau_n_KV(v) : inf_n_KV(v);
    case a_KA: return t ? tau_a_KA(v) : inf_a_KA(v);
    case b_KA: return t ? tau_b_KA(v) : inf_b_KA(v);
    case c_KC: return t ? tau_c_KC(v, ca) : inf_c_KC(v, ca);
    case sl_Kslow: return t ? tau_sl_Kslow(v) : inf_sl_Kslow(v);
    case ch_CaHVA: return t ? tau_ch_CaHVA(v) : inf_ch_CaHVA(v);
    case ci_CaHVA: return t ? tau_ci_CaHVA(v) : inf_ci_CaHVA(v);
    case cl_CaLVA: return t ? tau_cl_CaLVA(v) : inf_cl_CaLVA(v);
    case cm_CaLVA: return t ? tau_cm_CaLVA(v) : inf_cm_CaLVA(v);
    case hf_HCN1: return t ? tau_hf_HCN1(v) : inf_hf_HCN1(v);
    case hf_HCN2: return t ? tau_hf_HCN2(v) : inf_hf_HCN2(v);
    case hs_HCN1: return t ? tau_hs_HCN1(v) : inf_hs_HCN1(v);
    case hs_HCN2: return t ? tau_hs_HCN2(v) : inf_hs_HCN2(v);
  }
  return NAN;
}

/* Gaussian elimination with partial pivoting on the 6 x 6 system A x = rhs,
   as go_ion.cu:146-198 and :252-304 write it (strictly larger |pivot| wins,
   rows swapped from the pivot column on); returns 1 and leaves x alone on a
   pivot below 1e-12 (the reference returns there). */
static int solve6(double A[6][6], double *x, const double *rhs_in)
{
  double rhs[6];
  memcpy(rhs, rhs_in, sizeof(rhs));
  for (int i = 0; i < 5; i++) {
    int piv = i;
    double pmax = fabs(A[i][i]);
    for (int j = i + 1; j < 6; j++)
      if (fabs(A[j][i]) > pmax) { piv = j; pmax = fabs(A[j][i]); }
    if (fabs(pmax) < 1.0e-12) return 1;
    if (piv != i) {
      for (int j = i; j < 6; j++) { double t = A[i][j]; A[i][j] = A[piv][j]; A[piv][j] = t; }
      double t = rhs[i]; rhs[i] = rhs[piv]; rhs[piv] = t;
    }
    for (int j = i + 1; j < 6; j++) {
      double w = A[j][i] / A[i][i];
      A[j][i] = 0.0;
      for (int k = i + 1; k < 6; k++) A[j][k] = A[j][k] - A[i][k] * w;
      rhs[j] = rhs[j] - rhs[i] * w;
    }
  }
  for (int i = 5; i >= 0; i--) {
    for (int j = i + 1; j < 6; j++) rhs[i] = rhs[i] - A[i][j] * rhs[j];
    rhs[i] = rhs[i] / A[i][i];
  }
  memcpy(x, rhs, sizeof(rhs));
  return 0;
}

/* One KAHP step, go_ion.cu:88-213: Crank-Nicolson on the 6-state scheme
   x = (c1, c2, c3, c4, o1, o2) with Ca-dependent forward rates,
   (A(ca) - 2/DT) x' = -(A(ca_old) + 2/DT) x, solved by Gaussian elimination
   with partial pivoting.  Returns 1 (state unchanged) on a pivot below 1e-12,
   as the reference returns early there. */
static int kahp_step(double ca_old, double ca, double DT, double *st /* c1 c2 c3 c4 o1 o2 */)
{
  const double i1 = 80e-3, i2 = 80e-3, i3 = 200e-3, a1 = 1.0, a2 = 100e-3, b1 = 160e-3, b2 = 1.2;
  double l_ca = ca_old;
  const double l_dt = 2.0 / DT;
  double d1 = 200 * l_ca / 3.0, d2 = 160 * l_ca / 3.0, d3 = 80 * l_ca / 3.0;
  double An[6][6] = {
    {-d1 + l_dt, i1, 0.0, 0.0, 0.0, 0.0},
    {d1, -d2 - i1 + l_dt, i2, 0.0, 0.0, 0.0},
    {0.0, d2, -(d3 + i2 + b1) + l_dt, i3, a1, 0.0},
    {0.0, 0.0, d3, -(i3 + b2) + l_dt, 0.0, a2},
    {0.0, 0.0, b1, 0.0, -a1 + l_dt, 0.0},
    {0.0, 0.0, 0.0, b2, 0.0, -a2 + l_dt}};
  double x[6], rhs[6] = {0, 0, 0, 0, 0, 0};
  memcpy(x, st, sizeof(x));
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) rhs[i] -= x[j] * An[i][j];

  l_ca = ca;
  d1 = 200 * l_ca / 3.0; d2 = 160 * l_ca / 3.0; d3 = 80 * l_ca / 3.0;
  double A[6][6] = {
    {-d1 - l_dt, i1, 0.0, 0.0, 0.0, 0.0},
    {d1, -d2 - i1 - l_dt, i2, 0.0, 0.0, 0.0},
    {0.0, d2, -(d3 + i2 + b1) - l_dt, i3, a1, 0.0},
    {0.0, 0.0, d3, -(i3 + b2) - l_dt, 0.0, a2},
    {0.0, 0.0, b1, 0.0, -a1 - l_dt, 0.0},
    {0.0, 0.0, 0.0, b2, 0.0, -a2 - l_dt}};
  return solve6(A, st, rhs);
}

/* One implicit-Euler KAHP step at ca (go_KAHP_update, go_ion.cu:214-317, the
   BE path): (I - DT A(ca)) x' = x, the matrix formed as the reference forms
   it (every entry times -DT, then 1 added to the diagonal). */
static int kahp_be(double ca, double DT, double *st)
{
  const double i1 = 80e-3, i2 = 80e-3, i3 = 200e-3, a1 = 1.0, a2 = 100e-3, b1 = 160e-3, b2 = 1.2;
  const double d1 = 200 * ca / 3.0, d2 = 160 * ca / 3.0, d3 = 80 * ca / 3.0;
  double A[6][6] = {
    {-d1, i1, 0.0, 0.0, 0.0, 0.0},
    {d1, -d2 - i1, i2, 0.0, 0.0, 0.0},
    {0.0, d2, -(d3 + i2 + b1), i3, a1, 0.0},
    {0.0, 0.0, d3, -(i3 + b2), 0.0, a2},
    {0.0, 0.0, b1, 0.0, -a1, 0.0},
    {0.0, 0.0, 0.0, b2, 0.0, -a2}};
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) A[i][j] *= -DT;
  for (int i = 0; i < 6; i++) A[i][i] += 1.0;
  return solve6(A, st, st);
}

/* test hook: one implicit-Euler KAHP step, st6 in and out */
int or_go_kahp_be(double ca, double DT, double *st6) { return kahp_be(ca, DT, st6); }

/* test hook: st6 = (c1, c2, c3, c4, o1, o2) in and out */
int or_go_kahp(double ca_old, double ca, double DT, double *st6) { return kahp_step(ca_old, ca, DT, st6); }

static double *dal(int n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

typedef struct { int row, col; double val; } ent_t;
static int ent_cmp(const void *a, const void *b)
{
  const ent_t *x = (const ent_t *)a, *y = (const ent_t *)b;
  if (x->row != y->row) return x->row < y->row ? -1 : 1;
  return (x->col > y->col) - (x->col < y->col);
}

or_pop *or_go_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand)
{
  FILE *f = fopen(param_file, "r");
  if (!f) return NULL;
  double rad[GO_COMP], len[GO_COMP], Ra[GO_COMP], cmd[GO_COMP], dens[GO_COMP][N_COND_GO];
  int l_connect[GO_COMP], compart[GO_COMP];
  for (int i = 0; i < GO_COMP; i++) {
    int i1, i2, i3;
    double d1, d2, d3, d4, x[N_COND_GO];
    if (fscanf(f, "%d %d %lf %lf %lf %lf %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &i1, &i2, &d1,
               &d2, &d3, &d4, &i3, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6], &x[7], &x[8], &x[9], &x[10],
               &x[11], &x[12]) != 20 || i1 < 0 || i1 >= GO_COMP) {
      fclose(f);
      return NULL;
    }
    rad[i1] = 0.5 * d1 * 1e-4;
    len[i1] = d2 * 1e-4;
    Ra[i1] = d3 * 1e-3;
    cmd[i1] = d4;
    compart[i1] = i3;
    for (int q = 0; q < N_COND_GO; q++) dens[i1][q] = x[q];
    l_connect[i] = i2;   /* by line, go_solve.cu:37 */
  }
  fclose(f);

  const int n = nx * ny, nc = n * GO_COMP;
  or_pop *p = (or_pop *)calloc(1, sizeof(or_pop));
  p->type = OR_GO; p->n = n; p->nx = nx; p->ny = ny; p->comp = GO_COMP; p->nc = nc;
  p->ncond = N_COND_GO; p->nion = N_ION_GO;
  p->v = dal(nc); p->ca = dal(nc); p->cm = dal(nc); p->area = dal(nc);
  p->iext = dal(nc); p->shell = dal(nc); p->rev = dal(nc);
  p->ca2 = dal(nc); p->ca_old = dal(nc);
  p->compart = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < N_COND_GO; i++) p->cond[i] = dal(nc);
  for (int i = 0; i < N_ION_GO; i++) p->ion[i] = dal(nc);
  p->b = dal(nc); p->dammy = dal(nc);
  for (int i = 0; i < OR_NVEC; i++) p->vec[i] = dal(nc);
  p->hmax = 0.125;
  p->hmin = pow(2.0, -10.0);   /* H_MIN_GO, solve_rkc.cuh:25 */

  /* go.cu:42-87: area of a cylinder, Cm and conductances per area x area */
  for (int c = 0; c < n; c++)
    for (int k = 0; k < GO_COMP; k++) {
      const int i = c * GO_COMP + k;
      const double a = 2.0 * M_PI * rad[k] * len[k];
      p->area[i] = a;
      p->cm[i] = cmd[k] * a;
      p->compart[i] = compart[k];
      for (int q = 0; q < N_COND_GO; q++) p->cond[q][i] = dens[k][q] * a;
    }

  /* go_solve.cu:40-46: series conductance of the two half segments */
  static double mat[GO_COMP * GO_COMP];
  memset(mat, 0, sizeof(mat));
  for (int i = 0; i < GO_COMP; i++) {
    int d = l_connect[i];
    if (d >= 0) {
      mat[d + GO_COMP * i] = (2.0 / ((Ra[i] * len[i]) / (rad[i] * rad[i] * M_PI) + (Ra[d] * len[d]) / (rad[d] * rad[d] * M_PI)));
      mat[i + GO_COMP * d] = mat[d + GO_COMP * i];
    }
  }
  /* grid gap junctions between compartments 0, go_solve.cu:52-78 */
  int n_gj = (n > 1) ? 2 * ((nx - 1) * ny + (ny - 1) * nx) : 0;
  int *gpre = (int *)calloc(n_gj > 0 ? n_gj : 1, sizeof(int)), *gpost = (int *)calloc(n_gj > 0 ? n_gj : 1, sizeof(int));
  int c_gj = 0;
  if (n_gj > 0) {
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx - 1; x++) {
        gpre[c_gj * 2] = (y * nx + x) * GO_COMP;
        gpre[c_gj * 2 + 1] = (y * nx + x + 1) * GO_COMP;
        c_gj++;
      }
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny - 1; y++) {
        gpre[c_gj * 2] = (y * nx + x) * GO_COMP;
        gpre[c_gj * 2 + 1] = ((y + 1) * nx + x) * GO_COMP;
        c_gj++;
      }
    for (int i = 0; i < n_gj; i += 2) { gpost[i + 1] = gpre[i]; gpost[i] = gpre[i + 1]; }
  }
  /* CSR (go_solve.cu:108-213): per row the in-cell entries and the diagonal
     in column order, gap-junction entries inserted at their sorted place;
     the RKC sign makes off-diagonals +g and the diagonal -(sum) */
  int nnz1 = GO_COMP;
  for (int i = 0; i < GO_COMP * GO_COMP; i++) nnz1 += (mat[i] > 0.0);
  const int nnz = nnz1 * n + n_gj;
  ent_t *E = (ent_t *)malloc(sizeof(ent_t) * (nnz > 0 ? nnz : 1));
  double dg1[GO_COMP];
  for (int r = 0; r < GO_COMP; r++) {
    double s = 0.0;
    for (int c = 0; c < GO_COMP; c++)
      if (c != r) s += mat[c + GO_COMP * r];
    dg1[r] = s;
  }
  int k = 0;
  int *dpos = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  for (int cell = 0; cell < n; cell++)
    for (int r = 0; r < GO_COMP; r++)
      for (int c = 0; c < GO_COMP; c++) {
        const double m = mat[c + GO_COMP * r];
        if (m > 0.0) { E[k].row = cell * GO_COMP + r; E[k].col = cell * GO_COMP + c; E[k].val = m; k++; }
        else if (r == c) {
          E[k].row = cell * GO_COMP + r; E[k].col = cell * GO_COMP + c; E[k].val = -dg1[r];
          dpos[cell * GO_COMP + r] = k++;
        }
      }
  /* the diagonal of a GJ row gains G_GJ per junction (sign-flipped: -G) */
  for (int i = 0; i < n_gj; i++) {
    E[k].row = gpre[i]; E[k].col = gpost[i]; E[k].val = G_GJ_GO; k++;
    E[dpos[gpre[i]]].val -= G_GJ_GO;
  }
  free(dpos);
  qsort(E, nnz, sizeof(ent_t), ent_cmp);
  p->nnz = nnz;
  p->row = (int *)calloc(nc + 1, sizeof(int));
  p->col = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
  p->val = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
  for (int j = 0; j < nnz; j++) { p->col[j] = E[j].col; p->val[j] = E[j].val; p->row[E[j].row + 1]++; }
  for (int i = 0; i < nc; i++) p->row[i + 1] += p->row[i];
  free(E); free(gpre); free(gpost);

  /* initial state, go_ion.cu:590-634 */
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % GO_COMP == 0 && !v0_per_cell)
      init_v_rand = use_libc_rand ? ((double)rand() / RAND_MAX) - 0.5 : 0.0;
    const double vv = v0_per_cell ? v0_per_cell[i / GO_COMP] : V_INIT_GO + 2.0 * init_v_rand;
    p->v[i] = vv;
    p->ca[i] = Ca1_0_GO;
    p->ca2[i] = Ca1_0_GO;
    p->rev[i] = V_Ca_GO;
    p->ca_old[i] = Ca1_0_GO;
    p->iext[i] = 0.0;
    const double cav = p->ca[i];
    double **g = p->ion;
    g[m_NaT][i] = inf_m_NaT(vv);
    g[h_NaT][i] = inf_h_NaT(vv);
    g[r_NaR][i] = inf_r_NaR(vv);
    g[s_NaR][i] = inf_s_NaR(vv);
    g[p_NaP][i] = inf_p_NaP(vv);
    g[n_KV][i] = inf_n_KV(vv);
    g[a_KA][i] = inf_a_KA(vv);
    g[b_KA][i] = inf_b_KA(vv);
    g[c_KC][i] = inf_c_KC(vv, cav);
    g[sl_Kslow][i] = inf_sl_Kslow(vv);
    g[ch_CaHVA][i] = inf_ch_CaHVA(vv);
    g[ci_CaHVA][i] = inf_ci_CaHVA(vv);
    g[cl_CaLVA][i] = inf_cl_CaLVA(vv);
    g[cm_CaLVA][i] = inf_cm_CaLVA(vv);
    g[hf_HCN1][i] = inf_hf_HCN1(vv);
    g[hf_HCN2][i] = inf_hf_HCN2(vv);
    g[hs_HCN1][i] = inf_hs_HCN1(vv);
    g[hs_HCN2][i] = inf_hs_HCN2(vv);
    g[o1_KAHP][i] = 0.0; g[o2_KAHP][i] = 0.0;
    g[c2_KAHP][i] = 0.0; g[c3_KAHP][i] = 0.0; g[c4_KAHP][i] = 0.0;
    g[c1_KAHP][i] = 1.0;
  }
  /* go_set_current (go.cu:258-268): 0.2 nA into compartment 30 over [800, 1100) ms */
  p->stim_comp = GO_COMP_DEND * 3;
  p->stim_t0 = 800.0;
  p->stim_t1 = 1100.0;
  p->stim_amp = dal(n);
  for (int c = 0; c < n; c++) p->stim_amp[c] = 0.2e-3;
  p->others[OR_ATOL] = atol;
  p->others[OR_RTOL] = rtol;
  if (or_rkc_init(p) != 0) { or_pop_free(p); return NULL; }
  return p;
}

/* go_update_matrix, solve_rkc.cu:288-314 (terms in the reference's order) */
void or_go_update_b(or_pop *p, const double *vv)
{
  double **c = p->cond, **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const double l_v = vv[id];
    p->b[id] = (-p->iext[id]
                + c[g_leak][id] * (l_v - V_LEAK_GO)
                + c[g_NaT][id] * (l_v - V_Na_GO) * g[m_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[h_NaT][id]
                + c[g_NaR][id] * (l_v - V_Na_GO) * g[r_NaR][id] * g[s_NaR][id]
                + c[g_NaP][id] * (l_v - V_Na_GO) * g[p_NaP][id]
                + c[g_CaHVA][id] * (l_v - V_Ca_GO) * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id]
                + c[g_CaLVA][id] * (l_v - p->rev[id]) * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id]
                + c[g_KAHP][id] * (l_v - V_K_GO) * (g[o1_KAHP][id] + g[o2_KAHP][id])
                + c[g_KV][id] * (l_v - V_K_GO) * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
                + c[g_KA][id] * (l_v - V_K_GO) * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
                + c[g_KC][id] * (l_v - V_K_GO) * g[c_KC][id]
                + c[g_Kslow][id] * (l_v - V_K_GO) * g[sl_Kslow][id]
                + c[g_HCN1][id] * (l_v - V_H_GO) * (g[hf_HCN1][id] + g[hs_HCN1][id])
                + c[g_HCN2][id] * (l_v - V_H_GO) * (g[hf_HCN2][id] + g[hs_HCN2][id]));
  }
}

#define CNU(gt, INF, TAU) \
  g[gt][id] = (2.0 * DT * (INF) + (2.0 * (TAU) - DT) * g[gt][id]) / (2.0 * (TAU) + DT)

/* both Ca pools, the reversal and the gates of one compartment: Ca currents
   at v_ca, gates at v_val = (p->v + v_other) / 2 (go_update_ion_RKC_exp_imp,
   go_ion.cu:379-437, with v_ca = elem[v] and v_other = vnew;
   go_update_ion_exp_imp, :319-377, with v_ca = v_other = v_old) */
static void go_ion_at(or_pop *p, int id, double v_ca, double v_other, double DT)
{
  double **g = p->ion, **c = p->cond;
  const double v_val = (p->v[id] + v_other) / 2.0;
  double Ca_val = p->ca[id];
  const double Ca2_val = p->ca2[id];
  const double I_Ca1 = 1e-3 * c[g_CaHVA][id] / p->area[id] * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id] *
                       (v_ca - V_Ca_GO);
  const double I_Ca2 = 1e-3 * c[g_CaLVA][id] / p->area[id] * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id] *
                       (v_ca - (p->rev[id]));
  double k1 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val - Ca1_0_GO));
  double k2 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k1 / 2.0 - Ca1_0_GO));
  double k3 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k2 / 2.0 - Ca1_0_GO));
  double k4 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k3 - Ca1_0_GO));
  p->ca[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  k1 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val - Ca1_0_GO));
  k2 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k1 / 2.0 - Ca1_0_GO));
  k3 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k2 / 2.0 - Ca1_0_GO));
  k4 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k3 - Ca1_0_GO));
  p->ca2[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  p->rev[id] = (1e3) * (8.313424 * (23.0 + 273.15)) / (2 * F_GO) * log(Ca1OUT_GO / p->ca2[id]);
  Ca_val = (Ca_val + p->ca[id]) / 2.0;
  CNU(m_NaT, inf_m_NaT(v_val), tau_m_NaT(v_val));
  CNU(h_NaT, inf_h_NaT(v_val), tau_h_NaT(v_val));
  CNU(r_NaR, inf_r_NaR(v_val), tau_r_NaR(v_val));
  CNU(s_NaR, inf_s_NaR(v_val), tau_s_NaR(v_val));
  CNU(p_NaP, inf_p_NaP(v_val), tau_p_NaP(v_val));
  CNU(n_KV, inf_n_KV(v_val), tau_n_KV(v_val));
  CNU(a_KA, inf_a_KA(v_val), tau_a_KA(v_val));
  CNU(b_KA, inf_b_KA(v_val), tau_b_KA(v_val));
  CNU(sl_Kslow, inf_sl_Kslow(v_val), tau_sl_Kslow(v_val));
  CNU(c_KC, inf_c_KC(v_val, Ca_val), tau_c_KC(v_val, Ca_val));
  CNU(ch_CaHVA, inf_ch_CaHVA(v_val), tau_ch_CaHVA(v_val));
  CNU(ci_CaHVA, inf_ci_CaHVA(v_val), tau_ci_CaHVA(v_val));
  CNU(cl_CaLVA, inf_cl_CaLVA(v_val), tau_cl_CaLVA(v_val));
  CNU(cm_CaLVA, inf_cm_CaLVA(v_val), tau_cm_CaLVA(v_val));
  CNU(hf_HCN1, inf_hf_HCN1(v_val), tau_hf_HCN1(v_val));
  CNU(hf_HCN2, inf_hf_HCN2(v_val), tau_hf_HCN2(v_val));
  CNU(hs_HCN1, inf_hs_HCN1(v_val), tau_hs_HCN1(v_val));
  CNU(hs_HCN2, inf_hs_HCN2(v_val), tau_hs_HCN2(v_val));
}

/* go_KAHP_update_2order over every compartment (go_ion.cu:88-213) */
static void go_kahp_all(or_pop *p, double DT)
{
  double **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const double l_ca_old = p->ca_old[id];
    p->ca_old[id] = p->ca[id];
    double st[6] = {g[c1_KAHP][id], g[c2_KAHP][id], g[c3_KAHP][id], g[c4_KAHP][id], g[o1_KAHP][id], g[o2_KAHP][id]};
    if (kahp_step(l_ca_old, p->ca[id], DT, st) != 0) continue;
    g[c1_KAHP][id] = st[0]; g[c2_KAHP][id] = st[1]; g[c3_KAHP][id] = st[2];
    g[c4_KAHP][id] = st[3]; g[o1_KAHP][id] = st[4]; g[o2_KAHP][id] = st[5];
  }
}

/* go_update_ion_RKC_exp_imp (go_ion.cu:379-437) then go_KAHP_update_2order
   (go_ion.cu:88-213), as go_update_ion_and_matrix runs them (solve_rkc.cu:1174-1177) */
void or_go_update_ion(or_pop *p, const double *vnew, double DT)
{
  for (int id = 0; id < p->nc; id++) {
    go_ion_at(p, id, p->v[id], vnew[id], DT);
    p->v[id] = vnew[id];
  }
  go_kahp_all(p, DT);
}

/* go_update_ion_exp_imp (go_ion.cu:319-377) then the KAHP step, the CN path
   (go_solve_by_cnm, solve_cnm.cu:371-375) */
void or_go_cn_ion(or_pop *p, double DT)
{
  for (int id = 0; id < p->nc; id++) {
    go_ion_at(p, id, p->cn_vold[id], p->cn_vold[id], DT);
    p->cn_vold[id] = p->v[id];
  }
  go_kahp_all(p, DT);
}

/* conductance and reversal sums of the CN path before halving
   (go_cnm_vec_initialize / go_update_matrix, solve_cnm.cu:269-300, 320-351) */
void or_go_cn_ge(const or_pop *p, int id, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = (c[g_leak][id]
           + c[g_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[h_NaT][id]
           + c[g_NaR][id] * g[r_NaR][id] * g[s_NaR][id]
           + c[g_NaP][id] * g[p_NaP][id]
           + c[g_CaHVA][id] * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id]
           + c[g_CaLVA][id] * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id]
           + c[g_KAHP][id] * (g[o1_KAHP][id] + g[o2_KAHP][id])
           + c[g_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KA][id] * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KC][id] * g[c_KC][id]
           + c[g_Kslow][id] * g[sl_Kslow][id]
           + c[g_HCN1][id] * (g[hf_HCN1][id] + g[hs_HCN1][id])
           + c[g_HCN2][id] * (g[hf_HCN2][id] + g[hs_HCN2][id]));
  *esum = (c[g_leak][id] * V_LEAK_GO + p->iext[id]
           + c[g_NaT][id] * V_Na_GO * g[m_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[h_NaT][id]
           + c[g_NaR][id] * V_Na_GO * g[r_NaR][id] * g[s_NaR][id]
           + c[g_NaP][id] * V_Na_GO * g[p_NaP][id]
           + c[g_CaHVA][id] * V_Ca_GO * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id]
           + c[g_CaLVA][id] * (p->rev[id]) * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id]
           + c[g_KAHP][id] * V_K_GO * (g[o1_KAHP][id] + g[o2_KAHP][id])
           + c[g_KV][id] * V_K_GO * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KA][id] * V_K_GO * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KC][id] * V_K_GO * g[c_KC][id]
           + c[g_Kslow][id] * V_K_GO * g[sl_Kslow][id]
           + c[g_HCN1][id] * V_H_GO * (g[hf_HCN1][id] + g[hs_HCN1][id])
           + c[g_HCN2][id] * V_H_GO * (g[hf_HCN2][id] + g[hs_HCN2][id]));
}

/* ---------------- the backward-Euler (BE) path ---------------- */
#define EEU(gt, INF, TAU) \
  do { const double inf_ = (INF); g[gt][id] = inf_ + (g[gt][id] - inf_) * exp(-DT / (TAU)); } while (0)

/* go_KAHP_update at the present Ca (go_ion.cu:214-317), then go_update_ion
   (go_ion.cu:438-511): Ca currents with the present gates and reversal,
   every gate by exponential Euler at V (c_KC at the present Ca), both Ca
   pools by RK4 and the reversal from the new second pool; the order of
   go_solve_by_bem (solve_bem.cu:350-354) */
void or_go_be_ion(or_pop *p, double DT)
{
  double **g = p->ion, **c = p->cond;
  for (int id = 0; id < p->nc; id++) {
    double st[6] = {g[c1_KAHP][id], g[c2_KAHP][id], g[c3_KAHP][id], g[c4_KAHP][id], g[o1_KAHP][id], g[o2_KAHP][id]};
    if (kahp_be(p->ca[id], DT, st) != 0) continue;
    g[c1_KAHP][id] = st[0]; g[c2_KAHP][id] = st[1]; g[c3_KAHP][id] = st[2];
    g[c4_KAHP][id] = st[3]; g[o1_KAHP][id] = st[4]; g[o2_KAHP][id] = st[5];
  }
  for (int id = 0; id < p->nc; id++) {
    const double v_val = p->v[id], Ca_val = p->ca[id], Ca2_val = p->ca2[id];
    const double I_Ca1 = 1e-3 * c[g_CaHVA][id] / p->area[id] * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id] *
                         (p->v[id] - V_Ca_GO);
    const double I_Ca2 = 1e-3 * c[g_CaLVA][id] / p->area[id] * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id] *
                         (p->v[id] - (p->rev[id]));
    EEU(m_NaT, inf_m_NaT(v_val), tau_m_NaT(v_val));
    EEU(h_NaT, inf_h_NaT(v_val), tau_h_NaT(v_val));
    EEU(r_NaR, inf_r_NaR(v_val), tau_r_NaR(v_val));
    EEU(s_NaR, inf_s_NaR(v_val), tau_s_NaR(v_val));
    EEU(p_NaP, inf_p_NaP(v_val), tau_p_NaP(v_val));
    EEU(n_KV, inf_n_KV(v_val), tau_n_KV(v_val));
    EEU(a_KA, inf_a_KA(v_val), tau_a_KA(v_val));
    EEU(b_KA, inf_b_KA(v_val), tau_b_KA(v_val));
    EEU(c_KC, inf_c_KC(v_val, Ca_val), tau_c_KC(v_val, Ca_val));
    EEU(sl_Kslow, inf_sl_Kslow(v_val), tau_sl_Kslow(v_val));
    EEU(ch_CaHVA, inf_ch_CaHVA(v_val), tau_ch_CaHVA(v_val));
    EEU(ci_CaHVA, inf_ci_CaHVA(v_val), tau_ci_CaHVA(v_val));
    EEU(cl_CaLVA, inf_cl_CaLVA(v_val), tau_cl_CaLVA(v_val));
    EEU(cm_CaLVA, inf_cm_CaLVA(v_val), tau_cm_CaLVA(v_val));
    EEU(hf_HCN1, inf_hf_HCN1(v_val), tau_hf_HCN1(v_val));
    EEU(hf_HCN2, inf_hf_HCN2(v_val), tau_hf_HCN2(v_val));
    EEU(hs_HCN1, inf_hs_HCN1(v_val), tau_hs_HCN1(v_val));
    EEU(hs_HCN2, inf_hs_HCN2(v_val), tau_hs_HCN2(v_val));
    double k1 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val - Ca1_0_GO));
    double k2 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k1 / 2.0 - Ca1_0_GO));
    double k3 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k2 / 2.0 - Ca1_0_GO));
    double k4 = DT * (-I_Ca1 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca_val + k3 - Ca1_0_GO));
    p->ca[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
    k1 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val - Ca1_0_GO));
    k2 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k1 / 2.0 - Ca1_0_GO));
    k3 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k2 / 2.0 - Ca1_0_GO));
    k4 = DT * (-I_Ca2 / (2.0 * F_GO * SHELL1_D_GO) - B_Ca1_GO * (Ca2_val + k3 - Ca1_0_GO));
    p->ca2[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
    p->rev[id] = (1e3) * (8.313424 * (23.0 + 273.15)) / (2 * F_GO) * log(Ca1OUT_GO / p->ca2[id]);
  }
}

/* the BE system's diagonal addend and right-hand side, Cm/DT first, in the
   reference's term order (go_update_matrix, solve_bem.cu:323-347) */
void or_go_be_ge(const or_pop *p, int id, double DT, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = ((p->cm[id] / DT)
           + c[g_leak][id]
           + c[g_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[h_NaT][id]
           + c[g_NaR][id] * g[r_NaR][id] * g[s_NaR][id]
           + c[g_NaP][id] * g[p_NaP][id]
           + c[g_CaHVA][id] * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id]
           + c[g_CaLVA][id] * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id]
           + c[g_KAHP][id] * (g[o1_KAHP][id] + g[o2_KAHP][id])
           + c[g_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KA][id] * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KC][id] * g[c_KC][id]
           + c[g_Kslow][id] * g[sl_Kslow][id]
           + c[g_HCN1][id] * (g[hf_HCN1][id] + g[hs_HCN1][id])
           + c[g_HCN2][id] * (g[hf_HCN2][id] + g[hs_HCN2][id]));
  *esum = ((p->cm[id] / DT) * p->v[id]
           + c[g_leak][id] * V_LEAK_GO + p->iext[id]
           + c[g_NaT][id] * V_Na_GO * g[m_NaT][id] * g[m_NaT][id] * g[m_NaT][id] * g[h_NaT][id]
           + c[g_NaR][id] * V_Na_GO * g[r_NaR][id] * g[s_NaR][id]
           + c[g_NaP][id] * V_Na_GO * g[p_NaP][id]
           + c[g_CaHVA][id] * V_Ca_GO * g[ch_CaHVA][id] * g[ch_CaHVA][id] * g[ci_CaHVA][id]
           + c[g_CaLVA][id] * (p->rev[id]) * g[cl_CaLVA][id] * g[cl_CaLVA][id] * g[cm_CaLVA][id]
           + c[g_KAHP][id] * V_K_GO * (g[o1_KAHP][id] + g[o2_KAHP][id])
           + c[g_KV][id] * V_K_GO * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KA][id] * V_K_GO * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KC][id] * V_K_GO * g[c_KC][id]
           + c[g_Kslow][id] * V_K_GO * g[sl_Kslow][id]
           + c[g_HCN1][id] * V_H_GO * (g[hf_HCN1][id] + g[hs_HCN1][id])
           + c[g_HCN2][id] * V_H_GO * (g[hf_HCN2][id] + g[hs_HCN2][id]));
}
