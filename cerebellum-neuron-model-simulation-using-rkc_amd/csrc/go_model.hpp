// Golgi cell model on the RKC path: 131 compartments (three 10-compartment
// dendrites, soma, 100-compartment axon) with 12 channel types, two Ca pools
// and a 6-state Ca-activated K (KAHP) scheme; neighbouring cells of a grid are
// coupled through a linear gap junction between their compartments 0, which is
// part of the cable matrix.
//
// Constants and formulas follow the reference cell model:
//   constants                go.cuh:18-53
//   rate functions           go_ion.cu:3-86
//   KAHP Crank-Nicolson step go_ion.cu:88-213
//   ion update (RKC)         go_ion.cu:379-437
//   frozen current b         solve_rkc.cu:288-314
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "exp_acc.hpp"

namespace cbrkc {
namespace go {

constexpr int COMP = 131;        // go.cuh:19
constexpr int REC = 30;          // GO_COMP_DEND * 3: the soma, stimulated and recorded (go.cu:263, :280)
constexpr double V_INIT = -60.0, V_LEAK = -55.0, V_NA = 87.39, V_K = -84.69, V_CA = 120.0, V_H = -20.0;
constexpr double SHELL1_D = 0.2e-4, B_CA1 = 1.3, CA1_0 = 5e-5, CA1_OUT = 2.0, FARADAY = 9.6485e4;
constexpr double G_GJ = 2.5e-6;  // mS, go.cuh:40
constexpr double STIM = 0.2e-3, STIM_T0 = 800.0, STIM_T1 = 1100.0;   // go.cu:266

enum Gate { m_NaT, h_NaT, r_NaR, s_NaR, p_NaP, n_KV, a_KA, b_KA, c_KC, sl_Kslow, ch_CaHVA, ci_CaHVA, cl_CaLVA,
            cm_CaLVA, hf_HCN1, hf_HCN2, hs_HCN1, hs_HCN2, o1_KAHP, o2_KAHP, c1_KAHP, c2_KAHP, c3_KAHP, c4_KAHP,
            NGATE };
enum Cond { g_leak, g_NaT, g_NaR, g_NaP, g_KV, g_KA, g_KC, g_Kslow, g_CaHVA, g_CaLVA, g_HCN1, g_HCN2, g_KAHP,
            NCOND };

#define HD __host__ __device__ inline
HD double a_m_NaT(double v) { return 0.417 * (v + 25.0) / (1 - mexp(-(v + 25.0) / 10.0)); }
HD double b_m_NaT(double v) { return 16.68 * mexp(-0.055 * (v + 50.0)); }
HD double a_h_NaT(double v) { return 0.292 * mexp(-0.3 * (v + 50.0)); }
HD double b_h_NaT(double v) { return 4.17 / (1 + mexp(-(v + 17.0) / 5.0)); }
HD double a_r_NaR(double v) { return ((1.11 - 68.5 * (v - 4.48) / (mexp(-(v - 4.48) / 6.8) - 1.0)) * 1.0e-4); }
HD double b_r_NaR(double v) {
  double x = (v + 44.0) / 0.11;
  if (x > 200.0) x = 200.0;
  return ((66.0 + 21.7 * (v + 44) / (mexp(x) - 1.0)) * 1.0e-3);
}
HD double a_s_NaR(double v) { return (0.443 * mexp(-(v + 80.0) / 62.5)); }
HD double b_s_NaR(double v) { return (0.014 * mexp((v + 83.3) / 16.1)); }
HD double a_p_NaP(double v) { return (0.421 * (v + 40.0) / (1.0 - mexp(-(v + 40.0) / 5.0))); }
HD double b_p_NaP(double v) { return (-0.287 * (v + 40.0) / (1.0 - mexp((v + 40.0) / 5.0))); }
HD double a_ch_CaHVA(double v) { return (0.0687 * mexp(0.063 * (v + 29.0))); }
HD double b_ch_CaHVA(double v) { return (0.115 * mexp(-0.039 * (v + 18.66))); }
HD double a_ci_CaHVA(double v) { return (1.8e-3 * mexp(-(v + 48.0) / 18.0)); }
HD double b_ci_CaHVA(double v) { return (1.8e-3 * mexp((v + 48.0) / 83.0)); }
HD double a_n_KV(double v) { return 0.062 * (v + 26.0) / (1.0 - mexp(-(v + 26.0) / 10.0)); }
HD double b_n_KV(double v) { return 0.78 * mexp(-(v + 36.0) / 80.0); }
HD double a_a_KA(double v) { return (0.62 / (1.0 + mexp(-(v + 9.17) / 23.32))); }
HD double b_a_KA(double v) { return (0.126 / (mexp((v + 18.28) / 19.47))); }
HD double a_b_KA(double v) { return (0.028 / (1.0 + mexp((v + 111.0) / 12.84))); }
HD double b_b_KA(double v) { return (0.026 / (1.0 + mexp(-(v + 49.95) / 8.9))); }
HD double a_c_KC(double v, double ca) { return (3.2 / (1.0 + 0.0015 * mexp(-(v) / 11.7) / ca)); }
HD double b_c_KC(double v, double ca) { return (0.46 / (1.0 + ca / (1.5e-4 * mexp(-(v) / 11.7)))); }
HD double a_sl_Kslow(double v) { return (0.0037 * mexp((v + 30.0) / 40.0)); }
HD double b_sl_Kslow(double v) { return (0.0037 * mexp(-(v + 30.0) / 20.0)); }
HD double r_HCN1(double v) { return (0.0021 * (v) + 0.97); }
// fraction of the fast HCN2 component, clamped to [0, 1] (go_ion.cu:77-82)
HD double r_HCN2(double v) {
  if (v >= -64.70) return 0.0;
  if (v <= -108.70) return 1.0;
  return (-0.0227 * (v) - 1.47);
}

HD double inf_m_NaT(double v) { return a_m_NaT(v) / (a_m_NaT(v) + b_m_NaT(v)); }
HD double tau_m_NaT(double v) { return 1.0 / (a_m_NaT(v) + b_m_NaT(v)); }
HD double inf_h_NaT(double v) { return a_h_NaT(v) / (a_h_NaT(v) + b_h_NaT(v)); }
HD double tau_h_NaT(double v) { return 1.0 / (a_h_NaT(v) + b_h_NaT(v)); }
HD double inf_r_NaR(double v) { return a_r_NaR(v) / (a_r_NaR(v) + b_r_NaR(v)); }
HD double tau_r_NaR(double v) { return 1.0 / (a_r_NaR(v) + b_r_NaR(v)); }
HD double inf_s_NaR(double v) { return a_s_NaR(v) / (a_s_NaR(v) + b_s_NaR(v)); }
HD double tau_s_NaR(double v) { return 1.0 / (a_s_NaR(v) + b_s_NaR(v)); }
HD double inf_p_NaP(double v) { return (1 / (1.0 + mexp(-(v + 43.0) / 5.0))); }
HD double tau_p_NaP(double v) { return (5.0 / (a_p_NaP(v) + b_p_NaP(v))); }
HD double inf_ch_CaHVA(double v) { return (a_ch_CaHVA(v) / (a_ch_CaHVA(v) + b_ch_CaHVA(v))); }
HD double tau_ch_CaHVA(double v) { return (1.0 / (a_ch_CaHVA(v) + b_ch_CaHVA(v))); }
HD double inf_ci_CaHVA(double v) { return (a_ci_CaHVA(v) / (a_ci_CaHVA(v) + b_ci_CaHVA(v))); }
HD double tau_ci_CaHVA(double v) { return (1.0 / (a_ci_CaHVA(v) + b_ci_CaHVA(v))); }
HD double inf_cl_CaLVA(double v) { return (1.0 / (1.0 + mexp(-(v + 52.0) / 7.4))); }
HD double tau_cl_CaLVA(double v) {
  return ((3.0 + 1.0 / (mexp((v + 27.0) / 10.0) + mexp(-(v + 102.0) / 15.0))) / 0.85);
}
HD double inf_cm_CaLVA(double v) { return (1.0 / (1.0 + mexp((v + 80.0) / 5.0))); }
HD double tau_cm_CaLVA(double v) {
  return ((85.0 + 1.0 / (mexp((v + 48.0) / 4.0) + mexp(-(v + 407.0) / 50.0))) / 0.9);
}
HD double inf_n_KV(double v) { return (a_n_KV(v) / (a_n_KV(v) + b_n_KV(v))); }
HD double tau_n_KV(double v) { return (1.0 / (a_n_KV(v) + b_n_KV(v))); }
HD double inf_a_KA(double v) { return (1.0 / (1.0 + mexp(-(v + 38.0) / 17.0))); }
HD double tau_a_KA(double v) { return (1.0 / (a_a_KA(v) + b_a_KA(v))); }
HD double inf_b_KA(double v) { return (1.0 / (1.0 + mexp((v + 78.8) / 8.4))); }
HD double tau_b_KA(double v) { return (1.0 / (a_b_KA(v) + b_b_KA(v))); }
HD double inf_c_KC(double v, double ca) { return (a_c_KC(v, ca) / (a_c_KC(v, ca) + b_c_KC(v, ca))); }
HD double tau_c_KC(double v, double ca) { return (1.0 / (a_c_KC(v, ca) + b_c_KC(v, ca))); }
HD double inf_sl_Kslow(double v) { return (1.0 / (1.0 + mexp(-(v + 35.0) / 6.0))); }
HD double tau_sl_Kslow(double v) { return (1.0 / (a_sl_Kslow(v) + b_sl_Kslow(v))); }
HD double inf_hf_HCN1(double v) { return (r_HCN1(v) * (1.0 / (1.0 + mexp((v + 72.5) * 0.11)))); }
HD double inf_hs_HCN1(double v) { return ((1.0 - r_HCN1(v)) * (1.0 / (1.0 + mexp((v + 72.5) * 0.11)))); }
HD double tau_hf_HCN1(double v) { return (mexp((0.0137 * v + 3.37) * 2.3)); }
HD double tau_hs_HCN1(double v) { return (mexp((0.0145 * v + 4.06) * 2.3)); }
HD double inf_hf_HCN2(double v) { return (r_HCN2(v) * (1.0 / (1.0 + mexp((v + 81.9) * 0.16)))); }
HD double inf_hs_HCN2(double v) { return ((1.0 - r_HCN2(v)) * (1.0 / (1.0 + mexp((v + 81.9) * 0.16)))); }
HD double tau_hf_HCN2(double v) { return (mexp((0.027 * v + 5.6) * 2.3)); }
HD double tau_hs_HCN2(double v) { return (mexp((0.015 * v + 5.3) * 2.3)); }

// steady state of every gate at (v, ca); KAHP all closed in state c1
// (go_ion.cu:609-632)
HD void steady_state(double v, double ca, double *g) {
  g[m_NaT] = inf_m_NaT(v);
  g[h_NaT] = inf_h_NaT(v);
  g[r_NaR] = inf_r_NaR(v);
  g[s_NaR] = inf_s_NaR(v);
  g[p_NaP] = inf_p_NaP(v);
  g[n_KV] = inf_n_KV(v);
  g[a_KA] = inf_a_KA(v);
  g[b_KA] = inf_b_KA(v);
  g[c_KC] = inf_c_KC(v, ca);
  g[sl_Kslow] = inf_sl_Kslow(v);
  g[ch_CaHVA] = inf_ch_CaHVA(v);
  g[ci_CaHVA] = inf_ci_CaHVA(v);
  g[cl_CaLVA] = inf_cl_CaLVA(v);
  g[cm_CaLVA] = inf_cm_CaLVA(v);
  g[hf_HCN1] = inf_hf_HCN1(v);
  g[hf_HCN2] = inf_hf_HCN2(v);
  g[hs_HCN1] = inf_hs_HCN1(v);
  g[hs_HCN2] = inf_hs_HCN2(v);
  g[o1_KAHP] = g[o2_KAHP] = g[c2_KAHP] = g[c3_KAHP] = g[c4_KAHP] = 0.0;
  g[c1_KAHP] = 1.0;
}

HD void cn(double &x, double inf, double tau, double DT) { x = (2.0 * DT * inf + (2.0 * tau - DT) * x) / (2.0 * tau + DT); }
// exponential Euler, the BE path's gate update: x = inf + (x - inf) exp(-DT / tau)
HD void ee(double &x, double inf, double tau, double DT) { x = inf + (x - inf) * mexp(-DT / tau); }

// go_update_ion_RKC_exp_imp (go_ion.cu:379-437): both Ca pools by RK4 with
// their currents at the old voltage and gates (the second pool against the
// reversal from before this update), the second pool's Nernst reversal at
// 23 C, then every voltage-gated state by Crank-Nicolson at the mid-step
// voltage (c_KC at the mid-step Ca of the first pool).  cond holds density x
// area, so the Ca current densities divide by the area again.
// The Ca currents are taken at v_ca and the gates move at (v_a + v_b) / 2:
// the RKC path has v_ca = v_a = V before the update, v_b = the new V; the CN
// path (go_update_ion_exp_imp, go_ion.cu:319-377) v_ca = v_b = V_old, v_a = V.
HD void update_ion_at(double v_ca, double v_a, double v_b, double area, const double *c, double &ca, double &ca2,
                      double &rev, double DT, double *g) {
  const double v_val = (v_a + v_b) / 2.0;
  double Ca_val = ca;
  const double Ca2_val = ca2;
  const double I_Ca1 = 1e-3 * c[g_CaHVA] / area * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA] * (v_ca - V_CA);
  const double I_Ca2 = 1e-3 * c[g_CaLVA] / area * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA] * (v_ca - (rev));
  const double den = 2.0 * FARADAY * SHELL1_D;
  double k1 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val - CA1_0));
  double k2 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k1 / 2.0 - CA1_0));
  double k3 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k2 / 2.0 - CA1_0));
  double k4 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k3 - CA1_0));
  ca += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  k1 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val - CA1_0));
  k2 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k1 / 2.0 - CA1_0));
  k3 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k2 / 2.0 - CA1_0));
  k4 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k3 - CA1_0));
  ca2 += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  rev = (1e3) * (8.313424 * (23.0 + 273.15)) / (2 * FARADAY) * log(CA1_OUT / ca2);
  Ca_val = (Ca_val + ca) / 2.0;
  cn(g[m_NaT], inf_m_NaT(v_val), tau_m_NaT(v_val), DT);
  cn(g[h_NaT], inf_h_NaT(v_val), tau_h_NaT(v_val), DT);
  cn(g[r_NaR], inf_r_NaR(v_val), tau_r_NaR(v_val), DT);
  cn(g[s_NaR], inf_s_NaR(v_val), tau_s_NaR(v_val), DT);
  cn(g[p_NaP], inf_p_NaP(v_val), tau_p_NaP(v_val), DT);
  cn(g[n_KV], inf_n_KV(v_val), tau_n_KV(v_val), DT);
  cn(g[a_KA], inf_a_KA(v_val), tau_a_KA(v_val), DT);
  cn(g[b_KA], inf_b_KA(v_val), tau_b_KA(v_val), DT);
  cn(g[sl_Kslow], inf_sl_Kslow(v_val), tau_sl_Kslow(v_val), DT);
  cn(g[c_KC], inf_c_KC(v_val, Ca_val), tau_c_KC(v_val, Ca_val), DT);
  cn(g[ch_CaHVA], inf_ch_CaHVA(v_val), tau_ch_CaHVA(v_val), DT);
  cn(g[ci_CaHVA], inf_ci_CaHVA(v_val), tau_ci_CaHVA(v_val), DT);
  cn(g[cl_CaLVA], inf_cl_CaLVA(v_val), tau_cl_CaLVA(v_val), DT);
  cn(g[cm_CaLVA], inf_cm_CaLVA(v_val), tau_cm_CaLVA(v_val), DT);
  cn(g[hf_HCN1], inf_hf_HCN1(v_val), tau_hf_HCN1(v_val), DT);
  cn(g[hf_HCN2], inf_hf_HCN2(v_val), tau_hf_HCN2(v_val), DT);
  cn(g[hs_HCN1], inf_hs_HCN1(v_val), tau_hs_HCN1(v_val), DT);
  cn(g[hs_HCN2], inf_hs_HCN2(v_val), tau_hs_HCN2(v_val), DT);
}

HD void update_ion(double vold, double vnew, double area, const double *c, double &ca, double &ca2, double &rev,
                   double DT, double *g) {
  update_ion_at(vold, vold, vnew, area, c, ca, ca2, rev, DT, g);
}

// One step of the KAHP scheme on x = (c1, c2, c3, c4, o1, o2), the system
// solved by Gaussian elimination with partial pivoting as go_ion.cu:146-198 /
// :252-304 write it (strictly larger |pivot| wins, rows swapped from the
// pivot column on).  BE = false: the Crank-Nicolson step of
// go_KAHP_update_2order (go_ion.cu:88-213),
// (A(ca) - 2/DT) x' = -(A(ca_old) + 2/DT) x.  BE = true: the implicit-Euler
// step of go_KAHP_update (go_ion.cu:214-317), (I - DT A(ca)) x' = x with the
// matrix formed as the reference forms it (every entry times -DT, then 1
// added to the diagonal); ca_old is unused.  Every loop is unrolled so the
// system stays in registers (the elimination sits in the same body as the
// matrix build for that); a row exchange is a conditional swap with constant
// indices.  Returns false, leaving the state alone, on a pivot below 1e-12
// (the reference returns early there).
template <bool BE>
HD bool kahp_solve(double ca_old, double ca, double DT, double *g) {
  const double i1 = 80e-3, i2 = 80e-3, i3 = 200e-3, a1 = 1.0, a2 = 100e-3, b1 = 160e-3, b2 = 1.2;
  double A[6][6], r[6];
  if constexpr (BE) {
    const double d1 = 200 * ca / 3.0, d2 = 160 * ca / 3.0, d3 = 80 * ca / 3.0;
    const double A0[6][6] = {{-d1, i1, 0.0, 0.0, 0.0, 0.0},
                             {d1, -d2 - i1, i2, 0.0, 0.0, 0.0},
                             {0.0, d2, -(d3 + i2 + b1), i3, a1, 0.0},
                             {0.0, 0.0, d3, -(i3 + b2), 0.0, a2},
                             {0.0, 0.0, b1, 0.0, -a1, 0.0},
                             {0.0, 0.0, 0.0, b2, 0.0, -a2}};
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) A[i][j] = A0[i][j] * -DT;
#pragma unroll
    for (int i = 0; i < 6; i++) A[i][i] += 1.0;
    r[0] = g[c1_KAHP];
    r[1] = g[c2_KAHP];
    r[2] = g[c3_KAHP];
    r[3] = g[c4_KAHP];
    r[4] = g[o1_KAHP];
    r[5] = g[o2_KAHP];
  } else {
    const double l_dt = 2.0 / DT;
    double d1 = 200 * ca_old / 3.0, d2 = 160 * ca_old / 3.0, d3 = 80 * ca_old / 3.0;
    const double An[6][6] = {{-d1 + l_dt, i1, 0.0, 0.0, 0.0, 0.0},
                             {d1, -d2 - i1 + l_dt, i2, 0.0, 0.0, 0.0},
                             {0.0, d2, -(d3 + i2 + b1) + l_dt, i3, a1, 0.0},
                             {0.0, 0.0, d3, -(i3 + b2) + l_dt, 0.0, a2},
                             {0.0, 0.0, b1, 0.0, -a1 + l_dt, 0.0},
                             {0.0, 0.0, 0.0, b2, 0.0, -a2 + l_dt}};
    const double x[6] = {g[c1_KAHP], g[c2_KAHP], g[c3_KAHP], g[c4_KAHP], g[o1_KAHP], g[o2_KAHP]};
#pragma unroll
    for (int i = 0; i < 6; i++) {
      r[i] = 0.0;
#pragma unroll
      for (int j = 0; j < 6; j++) r[i] -= x[j] * An[i][j];
    }
    d1 = 200 * ca / 3.0;
    d2 = 160 * ca / 3.0;
    d3 = 80 * ca / 3.0;
    const double A1[6][6] = {{-d1 - l_dt, i1, 0.0, 0.0, 0.0, 0.0},
                             {d1, -d2 - i1 - l_dt, i2, 0.0, 0.0, 0.0},
                             {0.0, d2, -(d3 + i2 + b1) - l_dt, i3, a1, 0.0},
                             {0.0, 0.0, d3, -(i3 + b2) - l_dt, 0.0, a2},
                             {0.0, 0.0, b1, 0.0, -a1 - l_dt, 0.0},
                             {0.0, 0.0, 0.0, b2, 0.0, -a2 - l_dt}};
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) A[i][j] = A1[i][j];
  }
#pragma unroll
  for (int i = 0; i < 5; i++) {
    int piv = i;
    double pmax = fabs(A[i][i]);
#pragma unroll
    for (int j = i + 1; j < 6; j++)
      if (fabs(A[j][i]) > pmax) {
        piv = j;
        pmax = fabs(A[j][i]);
      }
    if (fabs(pmax) < 1.0e-12) return false;
#pragma unroll
    for (int j = i + 1; j < 6; j++)
      if (piv == j) {
#pragma unroll
        for (int k = i; k < 6; k++) {
          const double t = A[i][k];
          A[i][k] = A[j][k];
          A[j][k] = t;
        }
        const double t = r[i];
        r[i] = r[j];
        r[j] = t;
      }
#pragma unroll
    for (int j = i + 1; j < 6; j++) {
      const double w = A[j][i] / A[i][i];
      A[j][i] = 0.0;
#pragma unroll
      for (int k = i + 1; k < 6; k++) A[j][k] = A[j][k] - A[i][k] * w;
      r[j] = r[j] - r[i] * w;
    }
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
#pragma unroll
    for (int j = i + 1; j < 6; j++) r[i] = r[i] - A[i][j] * r[j];
    r[i] = r[i] / A[i][i];
  }
  g[c1_KAHP] = r[0];
  g[c2_KAHP] = r[1];
  g[c3_KAHP] = r[2];
  g[c4_KAHP] = r[3];
  g[o1_KAHP] = r[4];
  g[o2_KAHP] = r[5];
  return true;
}

HD bool kahp_step(double ca_old, double ca, double DT, double *g) { return kahp_solve<false>(ca_old, ca, DT, g); }
HD bool kahp_be(double ca, double DT, double *g) { return kahp_solve<true>(ca, ca, DT, g); }

// The BE ion update after the KAHP step (go_update_ion, go_ion.cu:438-511):
// Ca currents at V with the present gates and reversal, every gate by
// exponential Euler at V (c_KC at the present Ca), both Ca pools by RK4 and
// the reversal from the new second pool.
HD void update_ion_be(double v, double area, const double *c, double &ca, double &ca2, double &rev, double DT,
                      double *g) {
  const double Ca_val = ca, Ca2_val = ca2;
  const double I_Ca1 = 1e-3 * c[g_CaHVA] / area * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA] * (v - V_CA);
  const double I_Ca2 = 1e-3 * c[g_CaLVA] / area * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA] * (v - (rev));
  ee(g[m_NaT], inf_m_NaT(v), tau_m_NaT(v), DT);
  ee(g[h_NaT], inf_h_NaT(v), tau_h_NaT(v), DT);
  ee(g[r_NaR], inf_r_NaR(v), tau_r_NaR(v), DT);
  ee(g[s_NaR], inf_s_NaR(v), tau_s_NaR(v), DT);
  ee(g[p_NaP], inf_p_NaP(v), tau_p_NaP(v), DT);
  ee(g[n_KV], inf_n_KV(v), tau_n_KV(v), DT);
  ee(g[a_KA], inf_a_KA(v), tau_a_KA(v), DT);
  ee(g[b_KA], inf_b_KA(v), tau_b_KA(v), DT);
  ee(g[c_KC], inf_c_KC(v, Ca_val), tau_c_KC(v, Ca_val), DT);
  ee(g[sl_Kslow], inf_sl_Kslow(v), tau_sl_Kslow(v), DT);
  ee(g[ch_CaHVA], inf_ch_CaHVA(v), tau_ch_CaHVA(v), DT);
  ee(g[ci_CaHVA], inf_ci_CaHVA(v), tau_ci_CaHVA(v), DT);
  ee(g[cl_CaLVA], inf_cl_CaLVA(v), tau_cl_CaLVA(v), DT);
  ee(g[cm_CaLVA], inf_cm_CaLVA(v), tau_cm_CaLVA(v), DT);
  ee(g[hf_HCN1], inf_hf_HCN1(v), tau_hf_HCN1(v), DT);
  ee(g[hf_HCN2], inf_hf_HCN2(v), tau_hf_HCN2(v), DT);
  ee(g[hs_HCN1], inf_hs_HCN1(v), tau_hs_HCN1(v), DT);
  ee(g[hs_HCN2], inf_hs_HCN2(v), tau_hs_HCN2(v), DT);
  const double den = 2.0 * FARADAY * SHELL1_D;
  double k1 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val - CA1_0));
  double k2 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k1 / 2.0 - CA1_0));
  double k3 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k2 / 2.0 - CA1_0));
  double k4 = DT * (-I_Ca1 / den - B_CA1 * (Ca_val + k3 - CA1_0));
  ca += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  k1 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val - CA1_0));
  k2 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k1 / 2.0 - CA1_0));
  k3 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k2 / 2.0 - CA1_0));
  k4 = DT * (-I_Ca2 / den - B_CA1 * (Ca2_val + k3 - CA1_0));
  ca2 += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  rev = (1e3) * (8.313424 * (23.0 + 273.15)) / (2 * FARADAY) * log(CA1_OUT / ca2);
}

// go_update_matrix (solve_rkc.cu:288-314), terms in the reference's order
HD double current(double v, double iext, double rev, const double *c, const double *g) {
  return (-iext
          + c[g_leak] * (v - V_LEAK)
          + c[g_NaT] * (v - V_NA) * g[m_NaT] * g[m_NaT] * g[m_NaT] * g[h_NaT]
          + c[g_NaR] * (v - V_NA) * g[r_NaR] * g[s_NaR]
          + c[g_NaP] * (v - V_NA) * g[p_NaP]
          + c[g_CaHVA] * (v - V_CA) * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA]
          + c[g_CaLVA] * (v - rev) * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA]
          + c[g_KAHP] * (v - V_K) * (g[o1_KAHP] + g[o2_KAHP])
          + c[g_KV] * (v - V_K) * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
          + c[g_KA] * (v - V_K) * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
          + c[g_KC] * (v - V_K) * g[c_KC]
          + c[g_Kslow] * (v - V_K) * g[sl_Kslow]
          + c[g_HCN1] * (v - V_H) * (g[hf_HCN1] + g[hs_HCN1])
          + c[g_HCN2] * (v - V_H) * (g[hf_HCN2] + g[hs_HCN2]));
}
// CN path sums before halving (go_cnm_vec_initialize / go_update_matrix,
// solve_cnm.cu:269-300, 320-351)
HD void cn_ge(const double *c, const double *g, double rev, double iext, double &gs, double &es) {
  gs = (c[g_leak]
        + c[g_NaT] * g[m_NaT] * g[m_NaT] * g[m_NaT] * g[h_NaT]
        + c[g_NaR] * g[r_NaR] * g[s_NaR]
        + c[g_NaP] * g[p_NaP]
        + c[g_CaHVA] * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA]
        + c[g_CaLVA] * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA]
        + c[g_KAHP] * (g[o1_KAHP] + g[o2_KAHP])
        + c[g_KV] * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KA] * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KC] * g[c_KC]
        + c[g_Kslow] * g[sl_Kslow]
        + c[g_HCN1] * (g[hf_HCN1] + g[hs_HCN1])
        + c[g_HCN2] * (g[hf_HCN2] + g[hs_HCN2]));
  es = (c[g_leak] * V_LEAK + iext
        + c[g_NaT] * V_NA * g[m_NaT] * g[m_NaT] * g[m_NaT] * g[h_NaT]
        + c[g_NaR] * V_NA * g[r_NaR] * g[s_NaR]
        + c[g_NaP] * V_NA * g[p_NaP]
        + c[g_CaHVA] * V_CA * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA]
        + c[g_CaLVA] * (rev) * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA]
        + c[g_KAHP] * V_K * (g[o1_KAHP] + g[o2_KAHP])
        + c[g_KV] * V_K * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KA] * V_K * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KC] * V_K * g[c_KC]
        + c[g_Kslow] * V_K * g[sl_Kslow]
        + c[g_HCN1] * V_H * (g[hf_HCN1] + g[hs_HCN1])
        + c[g_HCN2] * V_H * (g[hf_HCN2] + g[hs_HCN2]));
}
// BE path: the system diagonal addend and right-hand side with the
// capacitive part first (go_update_matrix, solve_bem.cu:323-347)
HD void be_ge(const double *c, const double *g, double rev, double iext, double cmdt, double v, double &gs,
              double &es) {
  gs = (cmdt
        + c[g_leak]
        + c[g_NaT] * g[m_NaT] * g[m_NaT] * g[m_NaT] * g[h_NaT]
        + c[g_NaR] * g[r_NaR] * g[s_NaR]
        + c[g_NaP] * g[p_NaP]
        + c[g_CaHVA] * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA]
        + c[g_CaLVA] * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA]
        + c[g_KAHP] * (g[o1_KAHP] + g[o2_KAHP])
        + c[g_KV] * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KA] * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KC] * g[c_KC]
        + c[g_Kslow] * g[sl_Kslow]
        + c[g_HCN1] * (g[hf_HCN1] + g[hs_HCN1])
        + c[g_HCN2] * (g[hf_HCN2] + g[hs_HCN2]));
  es = (cmdt * v
        + c[g_leak] * V_LEAK + iext
        + c[g_NaT] * V_NA * g[m_NaT] * g[m_NaT] * g[m_NaT] * g[h_NaT]
        + c[g_NaR] * V_NA * g[r_NaR] * g[s_NaR]
        + c[g_NaP] * V_NA * g[p_NaP]
        + c[g_CaHVA] * V_CA * g[ch_CaHVA] * g[ch_CaHVA] * g[ci_CaHVA]
        + c[g_CaLVA] * (rev) * g[cl_CaLVA] * g[cl_CaLVA] * g[cm_CaLVA]
        + c[g_KAHP] * V_K * (g[o1_KAHP] + g[o2_KAHP])
        + c[g_KV] * V_K * g[n_KV] * g[n_KV] * g[n_KV] * g[n_KV]
        + c[g_KA] * V_K * g[a_KA] * g[a_KA] * g[a_KA] * g[b_KA]
        + c[g_KC] * V_K * g[c_KC]
        + c[g_Kslow] * V_K * g[sl_Kslow]
        + c[g_HCN1] * V_H * (g[hf_HCN1] + g[hs_HCN1])
        + c[g_HCN2] * V_H * (g[hf_HCN2] + g[hs_HCN2]));
}
#undef HD

}  // namespace go
}  // namespace cbrkc
