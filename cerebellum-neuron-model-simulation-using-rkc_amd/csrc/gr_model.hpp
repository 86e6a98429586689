// Granule cell model on the RKC path: 578 compartments (4 dendrites, soma,
// hillock, axon initial segment, ascending axon, parallel fibre) with a
// 12-state Markov Na channel on the hillock, AIS, axon and parallel fibre,
// Hodgkin-Huxley K channels and three leaks.  On the RKC path Ca and the Ca
// gates are not advanced (the reference's gr_update_ion_RKC_exp_imp leaves
// them at their initial values).
//
// Constants and formulas follow the reference cell model:
//   constants                 gr.cuh:11-40
//   rate functions            gr_ion.cu:3-43
//   Na scheme (CN, 12 x 12)   gr_ion.cu:45-207
//   ion update (RKC)          gr_ion.cu:461-480
//   frozen current b          solve_rkc.cu:226-249
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "exp_acc.hpp"

namespace cbrkc {
namespace gr {

constexpr int COMP = 578;     // gr.cuh:11
constexpr int REC = 16;       // soma: stimulated and written to gr_v (gr.cu:250, :264)
constexpr int RASTER = 77;    // raster compartment (gr.cu:270)
constexpr int HILL = 5;       // first compartment type with the Na scheme (gr.cuh:28)
constexpr int NTYPE = 9;
constexpr double V_INIT = -80.0, V_LEAK1 = -16.5, V_LEAK2 = -65.0, V_LEAK3 = -74.5, V_NA = 87.39, V_K = -84.69,
                 V_CA = 129.33;
constexpr double CA1_0 = 1e-4;
constexpr double STIM = 0.01e-3, STIM_T0 = 50.0, STIM_T1 = 100.0;   // gr.cu:252

enum Gate { n_KV, a_KA, b_KA, c_KCa, s_KM, ir_KIR, ch_Ca, ci_Ca, o_Na, c1_Na, c2_Na, c3_Na, c4_Na, c5_Na, i1_Na,
            i2_Na, i3_Na, i4_Na, i5_Na, i6_Na, NGATE };
enum Cond { g_leak1, g_leak2, g_leak3, g_Na, g_KV, g_KA, g_KCa, g_KM, g_KIR, g_Ca, NCOND };

#define HD __host__ __device__ inline
HD double a_n_KV(double v) {
  return (fabs(v + 25.0) > 1.0e-6) ? -0.017 * ((v + 25.0) / (mexp(-(v + 25.0) / 10.0) - 1.0))
                                   : -0.17 * (1.0 + (v + 25.0) / 20.0);
}
HD double b_n_KV(double v) { return 0.2223 * mexp(-(v + 35.0) / 80.0); }
HD double a_a_KA(double v) { return 0.771 / (1.0 + mexp(-(v + 9.17203) / 23.32708)); }
HD double b_a_KA(double v) { return 0.1567 / mexp((v + 18.27914) / 19.47175); }
HD double a_b_KA(double v) { return 0.0348 / (1.0 + mexp((v + 111.33208) / 12.8433)); }
HD double b_b_KA(double v) { return 0.0327 / (1.0 + mexp(-(v + 49.9537) / 8.90123)); }
HD double a_ir_KIR(double v) { return 0.2302 * mexp(-(v + 83.94) / 24.3902); }
HD double b_ir_KIR(double v) { return 0.2943 * mexp((v + 83.94) / 35.714); }
HD double a_s_KM(double v) { return 0.0046 * mexp((v + 30.0) / 40.0); }
HD double b_s_KM(double v) { return 0.0046 * mexp(-(v + 30.0) / 20.0); }
HD double a_c_KCa(double v, double ca) { return 1.4433 / (1.0 + (1.5e-3 * mexp(-v / 11.765)) / ca); }
HD double b_c_KCa(double v, double ca) { return 0.8660 / (1.0 + ca / (1.5e-4 * mexp(-v / 11.765))); }
HD double a_ch_Ca(double v) { return 0.0856 * mexp((v + 29.6) / 15.873); }
HD double b_ch_Ca(double v) { return 0.1437 * mexp(-(v + 18.66) / 25.641); }
HD double a_ci_Ca(double v) { return 0.0023 * mexp(-(v + 48.0) / 18.183); }
HD double b_ci_Ca(double v) { return 0.0025 * mexp((v + 48.0) / 83.33); }

HD double inf_n_KV(double v) { return (a_n_KV(v) / (a_n_KV(v) + b_n_KV(v))); }
HD double tau_n_KV(double v) { return (1.0 / (a_n_KV(v) + b_n_KV(v))); }
HD double inf_a_KA(double v) { return 1.0 / (1.0 + mexp(-(v + 38.0) / 17.0)); }
HD double tau_a_KA(double v) { return (1.0 / (a_a_KA(v) + b_a_KA(v))); }
HD double inf_b_KA(double v) { return 1.0 / (1.0 + mexp((v + 78.8) / 8.4)); }
HD double tau_b_KA(double v) { return (1.0 / (a_b_KA(v) + b_b_KA(v))); }
HD double inf_ir_KIR(double v) { return (a_ir_KIR(v) / (a_ir_KIR(v) + b_ir_KIR(v))); }
HD double tau_ir_KIR(double v) { return (1.0 / (a_ir_KIR(v) + b_ir_KIR(v))); }
HD double inf_s_KM(double v) { return 1.0 / (1.0 + mexp(-(v + 35.0) / 6.0)); }
HD double tau_s_KM(double v) { return (1.0 / (a_s_KM(v) + b_s_KM(v))); }
HD double inf_c_KCa(double v, double ca) { return a_c_KCa(v, ca) / (a_c_KCa(v, ca) + b_c_KCa(v, ca)); }
HD double tau_c_KCa(double v, double ca) { return (1.0 / (a_c_KCa(v, ca) + b_c_KCa(v, ca))); }
HD double inf_ch_Ca(double v) { return (a_ch_Ca(v) / (a_ch_Ca(v) + b_ch_Ca(v))); }
HD double inf_ci_Ca(double v) { return (a_ci_Ca(v) / (a_ci_Ca(v) + b_ci_Ca(v))); }

HD void cn(double &x, double inf, double tau, double DT) { x = (2.0 * DT * inf + (2.0 * tau - DT) * x) / (2.0 * tau + DT); }
// exponential Euler, the BE path's gate update: x = inf + (x - inf) exp(-DT / tau)
HD void ee(double &x, double inf, double tau, double DT) { x = inf + (x - inf) * mexp(-DT / tau); }

// Na scheme constants of one compartment type (gr_ion.cu:57-69, 88-89), with
// a = (Oon / Con)^(1/4), b = (Ooff / Coff)^(1/4) and their fourth powers
// evaluated once on the host by the C library's pow (the reference evaluates
// the same pow calls per compartment).
struct NaType {
  double n1, n2, n3, n4, Aalpha, Abeta, Vshift, Con, Coff, Oon, Ooff, a, b, a4, b4;
};
struct NaTable {
  NaType t[NTYPE];
};
NaTable na_table();   // gr_host.cpp

// Generator of the Na scheme at (alpha, beta) with s2dt added inside every
// diagonal entry's negation: s2dt = -2/DT gives the explicit half's matrix
// A + 2/DT, s2dt = +2/DT the implicit half's A - 2/DT (gr_ion.cu:93-106,
// 121-134).  States: c1..c5, i1..i6, o.
HD void na_matrix(double M[12][12], const NaType &T, double alpha, double beta, double s2dt) {
  const double n1 = T.n1, n2 = T.n2, n3 = T.n3, n4 = T.n4, Con = T.Con, Coff = T.Coff, Oon = T.Oon, Ooff = T.Ooff;
  const double a = T.a, b = T.b, a4 = T.a4, b4 = T.b4;
#pragma unroll
  for (int i = 0; i < 12; i++)
#pragma unroll
    for (int j = 0; j < 12; j++) M[i][j] = 0.0;
  M[0][0] = -(n1 * alpha + Con + s2dt);
  M[0][1] = n4 * beta;
  M[0][5] = Coff;
  M[1][0] = n1 * alpha;
  M[1][1] = -(n4 * beta + n2 * alpha + Con * a + s2dt);
  M[1][2] = n3 * beta;
  M[1][6] = Coff * b;
  M[2][1] = n2 * alpha;
  M[2][2] = -(n3 * beta + n3 * alpha + Con * a * a + s2dt);
  M[2][3] = n2 * beta;
  M[2][7] = Coff * b * b;
  M[3][2] = n3 * alpha;
  M[3][3] = -(n2 * beta + n4 * alpha + Con * a * a * a + s2dt);
  M[3][4] = n1 * beta;
  M[3][8] = Coff * b * b * b;
  M[4][3] = n4 * alpha;
  M[4][4] = -(n1 * beta + 259.8 + Con * a4 + s2dt);
  M[4][9] = Coff * b4;
  M[4][11] = 69.3;
  M[5][0] = Con;
  M[5][5] = -(Coff + n1 * alpha * a + s2dt);
  M[5][6] = n4 * beta * b;
  M[6][1] = Con * a;
  M[6][5] = n1 * alpha * a;
  M[6][6] = -(Coff * b + n4 * beta * b + n2 * alpha * a + s2dt);
  M[6][7] = n3 * beta * b;
  M[7][2] = Con * a * a;
  M[7][6] = n2 * alpha * a;
  M[7][7] = -(Coff * b * b + n3 * beta * b + n3 * alpha * a + s2dt);
  M[7][8] = n2 * beta * b;
  M[8][3] = Con * a * a * a;
  M[8][7] = n3 * alpha * a;
  M[8][8] = -(Coff * b * b * b + n2 * beta * b + n4 * alpha * a + s2dt);
  M[8][9] = n1 * beta * b;
  M[9][4] = Con * a4;
  M[9][8] = n4 * alpha * a;
  M[9][9] = -(Coff * b4 + n1 * beta * b + 259.8 + s2dt);
  M[9][10] = 69.3;
  M[10][9] = 259.8;
  M[10][10] = -(Ooff + 69.3 + s2dt);
  M[10][11] = Oon;
  M[11][4] = 259.8;
  M[11][10] = Ooff;
  M[11][11] = -(69.3 + Oon + s2dt);
}

// One step of the Na scheme on x = (c1..c5, i1..i6, o), the system solved by
// Gaussian elimination with partial pivoting as gr_ion.cu:150-199 / :276-330
// write it (strictly larger |pivot| wins, rows swapped from the pivot column
// on).  BE = false: the Crank-Nicolson step of gr_Na_update_2order
// (gr_ion.cu:86-202), (A(v_new) - 2/DT) x' = -(A(v_old) + 2/DT) x.
// BE = true: the implicit-Euler step of gr_Na_update (gr_ion.cu:209-357),
// (I - DT A(v_new)) x' = x with the matrix formed as the reference forms it
// (every entry times -DT, then 1 added to the diagonal); v_old is unused.
// Every loop is unrolled so the system stays in registers (the elimination
// sits in the same body as the matrix build for that); a row exchange is a
// conditional swap with constant indices.  Returns false, leaving x alone,
// on a pivot below 1e-12 (the reference returns early there).
template <bool BE>
HD bool na_solve(const NaType &T, double v_old, double v_new, double DT, double *x) {
  double M[12][12], r[12];
  if constexpr (BE) {
    const double alpha = T.Aalpha * mexp((v_new - T.Vshift) / 10.0);
    const double beta = T.Abeta * mexp(-(v_new - T.Vshift) / 10.0);
    na_matrix(M, T, alpha, beta, 0.0);
#pragma unroll
    for (int i = 0; i < 12; i++)
#pragma unroll
      for (int j = 0; j < 12; j++) M[i][j] *= -DT;
#pragma unroll
    for (int i = 0; i < 12; i++) {
      M[i][i] += 1.0;
      r[i] = x[i];
    }
  } else {
    double alpha = T.Aalpha * mexp((v_old - T.Vshift) / 10.0);
    double beta = T.Abeta * mexp(-(v_old - T.Vshift) / 10.0);
    na_matrix(M, T, alpha, beta, -2.0 / DT);
#pragma unroll
    for (int i = 0; i < 12; i++) {
      r[i] = 0.0;
#pragma unroll
      for (int j = 0; j < 12; j++) r[i] -= x[j] * M[i][j];
    }
    alpha = T.Aalpha * mexp((v_new - T.Vshift) / 10.0);
    beta = T.Abeta * mexp(-(v_new - T.Vshift) / 10.0);
    na_matrix(M, T, alpha, beta, 2.0 / DT);
  }
#pragma unroll
  for (int i = 0; i < 11; i++) {
    int piv = i;
    double pmax = fabs(M[i][i]);
#pragma unroll
    for (int j = i + 1; j < 12; j++)
      if (fabs(M[j][i]) > pmax) {
        piv = j;
        pmax = fabs(M[j][i]);
      }
    if (fabs(pmax) < 1.0e-12) return false;
#pragma unroll
    for (int j = i + 1; j < 12; j++)
      if (piv == j) {
#pragma unroll
        for (int k = i; k < 12; k++) {
          const double t = M[i][k];
          M[i][k] = M[j][k];
          M[j][k] = t;
        }
        const double t = r[i];
        r[i] = r[j];
        r[j] = t;
      }
#pragma unroll
    for (int j = i + 1; j < 12; j++) {
      const double w = M[j][i] / M[i][i];
      M[j][i] = 0.0;
#pragma unroll
      for (int k = i + 1; k < 12; k++) M[j][k] = M[j][k] - M[i][k] * w;
      r[j] = r[j] - r[i] * w;
    }
  }
#pragma unroll
  for (int i = 11; i >= 0; i--) {
#pragma unroll
    for (int j = i + 1; j < 12; j++) r[i] = r[i] - M[i][j] * r[j];
    r[i] = r[i] / M[i][i];
  }
#pragma unroll
  for (int i = 0; i < 12; i++) x[i] = r[i];
  return true;
}

HD bool na_step(const NaType &T, double v_old, double v_new, double DT, double *x) {
  return na_solve<false>(T, v_old, v_new, DT, x);
}
HD bool na_be(const NaType &T, double v, double DT, double *x) { return na_solve<true>(T, v, v, DT, x); }

// the Na states in the scheme's order, gathered from / scattered to the gate
// vector (ion order: o_Na first, then c1..c5, i1..i6)
HD void na_gather(const double *g, double *x) {
#pragma unroll
  for (int k = 0; k < 11; k++) x[k] = g[c1_Na + k];
  x[11] = g[o_Na];
}
HD void na_scatter(const double *x, double *g) {
#pragma unroll
  for (int k = 0; k < 11; k++) g[c1_Na + k] = x[k];
  g[o_Na] = x[11];
}

// gr_update_ion_RKC_exp_imp (gr_ion.cu:461-480): CN gates of KV, KA, KIR, KM
// and KCa at the mid-step voltage; Ca and the Ca gates stay
HD void update_gates(double vold, double vnew, double ca, double DT, double *g) {
  const double v_val = (vold + vnew) / 2.0;
  cn(g[n_KV], inf_n_KV(v_val), tau_n_KV(v_val), DT);
  cn(g[a_KA], inf_a_KA(v_val), tau_a_KA(v_val), DT);
  cn(g[b_KA], inf_b_KA(v_val), tau_b_KA(v_val), DT);
  cn(g[ir_KIR], inf_ir_KIR(v_val), tau_ir_KIR(v_val), DT);
  cn(g[s_KM], inf_s_KM(v_val), tau_s_KM(v_val), DT);
  cn(g[c_KCa], inf_c_KCa(v_val, ca), tau_c_KCa(v_val, ca), DT);
}

// gr_update_ion (gr_ion.cu:579-600, the BE path): exponential-Euler gates of
// KV, KA, KIR, KM and KCa at V; Ca and the Ca gates stay
HD void update_gates_be(double v, double ca, double DT, double *g) {
  ee(g[n_KV], inf_n_KV(v), tau_n_KV(v), DT);
  ee(g[a_KA], inf_a_KA(v), tau_a_KA(v), DT);
  ee(g[b_KA], inf_b_KA(v), tau_b_KA(v), DT);
  ee(g[ir_KIR], inf_ir_KIR(v), tau_ir_KIR(v), DT);
  ee(g[s_KM], inf_s_KM(v), tau_s_KM(v), DT);
  ee(g[c_KCa], inf_c_KCa(v, ca), tau_c_KCa(v, ca), DT);
}

// gr_update_matrix (solve_rkc.cu:226-249), terms in the reference's order
HD double current(double v, double iext, const double *c, const double *g) {
  return (+c[g_leak1] * (v - V_LEAK1)
          + c[g_leak2] * (v - V_LEAK2)
          + c[g_leak3] * (v - V_LEAK3)
          - iext
          + c[g_Na] * (v - V_NA) * g[o_Na]
          + c[g_Ca] * (v - V_CA) * g[ch_Ca] * g[ch_Ca] * g[ci_Ca]
          + c[g_KV] * (v - V_K) * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
          + c[g_KIR] * (v - V_K) * g[ir_KIR]
          + c[g_KA] * (v - V_K) * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
          + c[g_KCa] * (v - V_K) * g[c_KCa]
          + c[g_KM] * (v - V_K) * g[s_KM]);
}
// CN path sums before halving (gr_cnm_vec_initialize / gr_update_matrix,
// solve_cnm.cu:138-161, 179-200)
HD void cn_ge(const double *c, const double *g, double iext, double &gs, double &es) {
  gs = (c[g_leak1] + c[g_leak2] + c[g_leak3]
        + c[g_Na] * g[o_Na]
        + c[g_Ca] * g[ch_Ca] * g[ch_Ca] * g[ci_Ca]
        + c[g_KV] * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KIR] * g[ir_KIR]
        + c[g_KA] * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KCa] * g[c_KCa]
        + c[g_KM] * g[s_KM]);
  es = (c[g_leak1] * V_LEAK1 + c[g_leak2] * V_LEAK2 + c[g_leak3] * V_LEAK3 + iext
        + c[g_Na] * V_NA * g[o_Na]
        + c[g_Ca] * V_CA * g[ch_Ca] * g[ch_Ca] * g[ci_Ca]
        + c[g_KV] * V_K * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KIR] * V_K * g[ir_KIR]
        + c[g_KA] * V_K * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KCa] * V_K * g[c_KCa]
        + c[g_KM] * V_K * g[s_KM]);
}
// BE path: the system diagonal addend and right-hand side with the
// capacitive part first (gr_update_matrix, solve_bem.cu:258-291)
HD void be_ge(const double *c, const double *g, double iext, double cmdt, double v, double &gs, double &es) {
  gs = (cmdt
        + c[g_leak1] + c[g_leak2] + c[g_leak3]
        + c[g_Na] * g[o_Na]
        + c[g_Ca] * g[ch_Ca] * g[ch_Ca] * g[ci_Ca]
        + c[g_KV] * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KIR] * g[ir_KIR]
        + c[g_KA] * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KCa] * g[c_KCa]
        + c[g_KM] * g[s_KM]);
  es = (cmdt * v
        + c[g_leak1] * V_LEAK1 + c[g_leak2] * V_LEAK2 + c[g_leak3] * V_LEAK3 + iext
        + c[g_Na] * V_NA * g[o_Na]
        + c[g_Ca] * V_CA * g[ch_Ca] * g[ch_Ca] * g[ci_Ca]
        + c[g_KV] * V_K * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KIR] * V_K * g[ir_KIR]
        + c[g_KA] * V_K * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KCa] * V_K * g[c_KCa]
        + c[g_KM] * V_K * g[s_KM]);
}
#undef HD

}  // namespace gr
}  // namespace cbrkc
