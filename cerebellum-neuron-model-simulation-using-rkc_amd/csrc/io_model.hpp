// Inferior-olive cell model on the RKC path: three compartments (dendrite,
// soma, axon hillock) with per-area densities, coupled to the four grid
// neighbours' dendrites by nonlinear gap junctions.
//
// Constants and formulas follow the reference cell model:
//   constants            io.cuh:11-58
//   rate functions       io_ion.cu:3-54
//   ion update (RKC)     io_ion.cu:90-127
//   frozen current b     solve_rkc.cu:412-434
//   gap current          gap.cu:149-160 (added to b, solve_rkc.cu:400-410)
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "exp_acc.hpp"

namespace cbrkc {
namespace io {

constexpr int COMP = 3;   // io.cuh:11
constexpr int DEND = 0, SOMA = 1, AXON = 2;
constexpr double V_INIT = -60.0, V_LEAK = 10.0, V_NA = 55.0, V_CA = 120.0, V_K = -75.0, V_H = -43.0;
constexpr double G_LEAK = 0.016, SOMA_G_CAL = 0.83, SOMA_G_NA = 120.0, SOMA_G_KDR = 9.0, SOMA_G_K = 5.0;
constexpr double DNDR_G_H = 0.15, DNDR_G_CAH = 4.5, DNDR_G_KCA = 35.0, AXON_G_NA = 240.0, AXON_G_K = 20.0;
constexpr double G_INT = 0.13, G_P_DS = 0.200, G_P_SA = 0.869, G_C = 0.04;
constexpr double CA_PH = 3.0, BETA_CA = 0.075;
constexpr double STIM = 5.5;              // uA/cm^2 on the dendrite, io.cu:253
constexpr double STIM_T0 = 800.0, STIM_T1 = 820.0;

enum Gate { k_CaL, l_CaL, m_Na, h_Na, n_Kdr, p_Kdr, x_K, r_CaH, s_KCa, q_H, NGATE };
enum Cond { g_leak, g_CaL, g_Na, g_Kdr, g_K, g_CaH, g_KCa, g_H, NCOND };

template <typename T> __host__ __device__ inline T alpha_r(T v) { return T(1.7) / (T(1) + mexp((-(v - T(5))) / T(13.9))); }
template <typename T> __host__ __device__ inline T alpha_s(T x) { return fmin(T(0.00002) * x, T(0.01)); }
template <typename T> __host__ __device__ inline T alpha_x(T v) { return (T(0.13) * (v + T(25))) / (T(1) - mexp(-(v + T(25)) / T(10))); }
template <typename T> __host__ __device__ inline T beta_r(T v) { return (T(0.02) * (v + T(8.5))) / (mexp((v + T(8.5)) / T(5)) - T(1)); }
template <typename T> __host__ __device__ inline T beta_s() { return T(0.015); }
template <typename T> __host__ __device__ inline T beta_x(T v) { return T(1.69) * mexp(T(-0.0125) * (v + T(35))); }

// h_inf and tau_h are defined for the soma (1) and the axon (2) only; the
// dendrite keeps 1 for both (io_ion.cu:17-21, 37-41)
template <typename T> __host__ __device__ inline T h_inf(T v, int c) {
  T h = T(1);
  if (c == 1) h = T(1) / (T(1) + mexp((-v - T(70)) / T(-5.8)));
  if (c == 2) h = T(1) / (T(1) + mexp((-v - T(60)) / T(-5.8)));
  return h;
}
template <typename T> __host__ __device__ inline T tau_h(T v, int c) {
  T h = T(1);
  if (c == 1) h = T(3) * mexp((-v - T(40)) / T(33));
  if (c == 2) h = T(1.5) * mexp((-v - T(40)) / T(33));
  return h;
}
template <typename T> __host__ __device__ inline T k_inf(T v) { return T(1) / (T(1) + mexp((-v - T(61)) / T(4.2))); }
template <typename T> __host__ __device__ inline T l_inf(T v) { return T(1) / (T(1) + mexp((v + T(85.5)) / T(8.5))); }
template <typename T> __host__ __device__ inline T m_inf(T v) { return T(1) / (T(1) + mexp((-v - T(30)) / T(5.5))); }
template <typename T> __host__ __device__ inline T n_inf(T v) { return T(1) / (T(1) + mexp((-v - T(3)) / T(10))); }
template <typename T> __host__ __device__ inline T p_inf(T v) { return T(1) / (T(1) + mexp((-v - T(51)) / T(-12))); }
template <typename T> __host__ __device__ inline T q_inf(T v) { return T(1) / (T(1) + mexp((v + T(80)) / T(4))); }
template <typename T> __host__ __device__ inline T r_inf(T v) { return alpha_r(v) / (alpha_r(v) + beta_r(v)); }
template <typename T> __host__ __device__ inline T s_inf(T x) { return alpha_s(x) / (alpha_s(x) + beta_s<T>()); }
template <typename T> __host__ __device__ inline T x_inf(T v) { return alpha_x(v) / (alpha_x(v) + beta_x(v)); }
template <typename T> __host__ __device__ inline T tau_k(T) { return T(1); }
template <typename T> __host__ __device__ inline T tau_l(T v) {
  return ((T(20) * mexp((v + T(160)) / T(30))) / (T(1) + mexp((v + T(84)) / T(7.3)))) + T(35);
}
template <typename T> __host__ __device__ inline T tau_n(T v) { return T(47) * mexp((-(-v - T(50))) / T(900)) + T(5); }
template <typename T> __host__ __device__ inline T tau_q(T v) { return T(1) / (mexp(T(-0.086) * v - T(14.6)) + mexp(T(0.07) * v - T(1.87))); }
template <typename T> __host__ __device__ inline T tau_r(T v) { return T(5) / (alpha_r(v) + beta_r(v)); }
template <typename T> __host__ __device__ inline T tau_s(T x) { return T(1) / (alpha_s(x) + beta_s<T>()); }
template <typename T> __host__ __device__ inline T tau_x(T v) { return T(1) / (alpha_x(v) + beta_x(v)); }

template <typename T> __host__ __device__ inline void cn(T &x, T inf, T tau, T dt) {
  x = (T(2) * dt * inf + (T(2) * tau - dt) * x) / (T(2) * tau + dt);
}

// Ca RK4 with the Ca current at v_ca and the present gate, then every gate by
// Crank-Nicolson at v_val (m_Na instantaneous, s_KCa at the mid-step Ca).
// io_update_ion_RKC (io_ion.cu:90-127) uses v_ca = V before the update and
// v_val = the mid-step voltage; io_update_ion_2nd (:53-89, the CN path)
// v_ca = V and v_val = 1.5 V - 0.5 V_old.
template <typename T>
__host__ __device__ inline void update_ion_at(T v_ca, T v_val, T &ca, T gCaH, int comp, T DT, T *g) {
  T Ca_val = ca;
  const T I_Ca1 = gCaH * g[r_CaH] * g[r_CaH] * (v_ca - T(V_CA));
  const T k1 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val));
  const T k2 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k1 / T(2)));
  const T k3 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k2 / T(2)));
  const T k4 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k3));
  ca += (k1 + k2 * T(2) + k3 * T(2) + k4) / T(6);
  Ca_val = (Ca_val + ca) / T(2);
  cn(g[k_CaL], k_inf(v_val), tau_k(v_val), DT);
  cn(g[l_CaL], l_inf(v_val), tau_l(v_val), DT);
  cn(g[n_Kdr], n_inf(v_val), tau_n(v_val), DT);
  cn(g[p_Kdr], p_inf(v_val), tau_n(v_val), DT);
  cn(g[x_K], x_inf(v_val), tau_x(v_val), DT);
  cn(g[r_CaH], r_inf(v_val), tau_r(v_val), DT);
  cn(g[q_H], q_inf(v_val), tau_q(v_val), DT);
  cn(g[s_KCa], s_inf(Ca_val), tau_s(Ca_val), DT);
  g[m_Na] = m_inf(v_val);
  cn(g[h_Na], h_inf(v_val, comp), tau_h(v_val, comp), DT);
}

template <typename T>
__host__ __device__ inline void update_ion(T vold, T vnew, T &ca, T gCaH, int comp, T DT, T *g) {
  update_ion_at(vold, T(0.5) * (vnew + vold), ca, gCaH, comp, DT, g);
}

// exponential Euler, the BE path's gate update: x = inf + (x - inf) exp(-dt / tau)
template <typename T> __host__ __device__ inline void ee(T &x, T inf, T tau, T dt) { x = inf + (x - inf) * mexp(-dt / tau); }

// io_update_ion (io_ion.cu:128-161), the BE path: every gate by exponential
// Euler at V (s_KCa at the present Ca, m_Na instantaneous), then the Ca RK4
// with the Ca current of the already-updated r_CaH gate.
template <typename T>
__host__ __device__ inline void update_ion_be(T v, T &ca, T gCaH, int comp, T DT, T *g) {
  const T Ca_val = ca;
  ee(g[k_CaL], k_inf(v), tau_k(v), DT);
  ee(g[l_CaL], l_inf(v), tau_l(v), DT);
  ee(g[n_Kdr], n_inf(v), tau_n(v), DT);
  ee(g[p_Kdr], p_inf(v), tau_n(v), DT);
  ee(g[x_K], x_inf(v), tau_x(v), DT);
  ee(g[r_CaH], r_inf(v), tau_r(v), DT);
  ee(g[q_H], q_inf(v), tau_q(v), DT);
  ee(g[s_KCa], s_inf(Ca_val), tau_s(Ca_val), DT);
  g[m_Na] = m_inf(v);
  ee(g[h_Na], h_inf(v, comp), tau_h(v, comp), DT);
  const T I_Ca1 = gCaH * g[r_CaH] * g[r_CaH] * (v - T(V_CA));
  const T k1 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val));
  const T k2 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k1 / T(2)));
  const T k3 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k2 / T(2)));
  const T k4 = DT * (-I_Ca1 * T(CA_PH) - T(BETA_CA) * (Ca_val + k3));
  ca += (k1 + k2 * T(2) + k3 * T(2) + k4) / T(6);
}

// CN path sums before halving (io_cnm_vec_initialize / io_update_matrix,
// solve_cnm.cu:751-768, 793-810)
__host__ __device__ inline void cn_ge(const double *c, const double *g, double iext, double &gs, double &es) {
  gs = (c[g_leak]
        + c[g_CaL] * g[k_CaL] * g[k_CaL] * g[k_CaL] * g[l_CaL]
        + c[g_Na] * g[m_Na] * g[m_Na] * g[m_Na] * g[h_Na]
        + c[g_Kdr] * g[n_Kdr] * g[p_Kdr]
        + c[g_K] * g[x_K] * g[x_K] * g[x_K] * g[x_K]
        + c[g_CaH] * g[r_CaH] * g[r_CaH]
        + c[g_KCa] * g[s_KCa]
        + c[g_H] * g[q_H]);
  es = (c[g_leak] * V_LEAK + iext
        + c[g_CaL] * V_CA * g[k_CaL] * g[k_CaL] * g[k_CaL] * g[l_CaL]
        + c[g_Na] * V_NA * g[m_Na] * g[m_Na] * g[m_Na] * g[h_Na]
        + c[g_Kdr] * V_K * g[n_Kdr] * g[p_Kdr]
        + c[g_K] * V_K * g[x_K] * g[x_K] * g[x_K] * g[x_K]
        + c[g_CaH] * V_CA * g[r_CaH] * g[r_CaH]
        + c[g_KCa] * V_K * g[s_KCa]
        + c[g_H] * V_H * g[q_H]);
}

// BE path: the system diagonal addend and right-hand side with the
// capacitive part first (io_update_matrix, solve_bem.cu:478-508)
__host__ __device__ inline void be_ge(const double *c, const double *g, double iext, double cmdt, double v,
                                      double &gs, double &es) {
  gs = (cmdt
        + c[g_leak]
        + c[g_CaL] * g[k_CaL] * g[k_CaL] * g[k_CaL] * g[l_CaL]
        + c[g_Na] * g[m_Na] * g[m_Na] * g[m_Na] * g[h_Na]
        + c[g_Kdr] * g[n_Kdr] * g[p_Kdr]
        + c[g_K] * g[x_K] * g[x_K] * g[x_K] * g[x_K]
        + c[g_CaH] * g[r_CaH] * g[r_CaH]
        + c[g_KCa] * g[s_KCa]
        + c[g_H] * g[q_H]);
  es = (cmdt * v
        + c[g_leak] * V_LEAK + iext
        + c[g_CaL] * V_CA * g[k_CaL] * g[k_CaL] * g[k_CaL] * g[l_CaL]
        + c[g_Na] * V_NA * g[m_Na] * g[m_Na] * g[m_Na] * g[h_Na]
        + c[g_Kdr] * V_K * g[n_Kdr] * g[p_Kdr]
        + c[g_K] * V_K * g[x_K] * g[x_K] * g[x_K] * g[x_K]
        + c[g_CaH] * V_CA * g[r_CaH] * g[r_CaH]
        + c[g_KCa] * V_K * g[s_KCa]
        + c[g_H] * V_H * g[q_H]);
}

// io_update_matrix (solve_rkc.cu:412-434), terms in the reference's order
template <typename T>
__host__ __device__ inline T current(T v, T iext, const T *c, const T *g) {
  return (-iext
          + c[g_leak] * (v - T(V_LEAK))
          + c[g_CaL] * (v - T(V_CA)) * g[k_CaL] * g[k_CaL] * g[k_CaL] * g[l_CaL]
          + c[g_Na] * (v - T(V_NA)) * g[m_Na] * g[m_Na] * g[m_Na] * g[h_Na]
          + c[g_Kdr] * (v - T(V_K)) * g[n_Kdr] * g[p_Kdr]
          + c[g_K] * (v - T(V_K)) * g[x_K] * g[x_K] * g[x_K] * g[x_K]
          + c[g_CaH] * (v - T(V_CA)) * g[r_CaH] * g[r_CaH]
          + c[g_KCa] * (v - T(V_K)) * g[s_KCa]
          + c[g_H] * (v - T(V_H)) * g[q_H]);
}

// gap current into a dendrite, diff = V_post - V_pre (gap.cu:149-160)
template <typename T> __host__ __device__ inline T gap_current(T diff) {
  return T(G_C) * (T(0.8) * mexp(T(-0.01) * diff * diff) + T(0.2)) * (diff);
}

// the four gap-junction partners of cell id on an nx x ny grid, or -1
// (gap.cu:4-133; the grid row is id / ny, as the reference computes it)
__host__ __device__ inline void neighbours(int id, int nx, int ny, int out[4]) {
  const int idx = id % nx, idy = id / ny;
  const int l = id - 1, u = id - nx, r = id + 1, d = id + nx;
  if (idx == 0 && idy == 0) { out[0] = -1; out[1] = -1; out[2] = r; out[3] = d; }
  else if (idx == nx - 1 && idy == 0) { out[0] = l; out[1] = -1; out[2] = -1; out[3] = d; }
  else if (idx == 0 && idy == ny - 1) { out[0] = -1; out[1] = u; out[2] = r; out[3] = -1; }
  else if (idx == nx - 1 && idy == ny - 1) { out[0] = l; out[1] = u; out[2] = -1; out[3] = -1; }
  else if (idy == 0) { out[0] = l; out[1] = -1; out[2] = r; out[3] = d; }
  else if (idy == ny - 1) { out[0] = l; out[1] = u; out[2] = r; out[3] = -1; }
  else if (idx == 0) { out[0] = -1; out[1] = u; out[2] = r; out[3] = d; }
  else if (idx == nx - 1) { out[0] = l; out[1] = u; out[2] = -1; out[3] = d; }
  else { out[0] = l; out[1] = u; out[2] = r; out[3] = d; }
}

}  // namespace io
}  // namespace cbrkc
