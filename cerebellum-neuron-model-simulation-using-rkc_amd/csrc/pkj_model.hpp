// Purkinje-cell channel model on the RKC path, shared by the FP64 and mixed
// engines.  Written for CDNA: every function is a straight-line __host__
// __device__ expression, templated on the arithmetic type so the same text
// serves the FP64 population engine and the FP32 per-cell engine.
//
// Model constants and formulas follow the reference cell model:
//   constants            pkj.cuh:11-82
//   rate functions       pkj_ion.cu:4-91
//   ion update (RKC)     pkj_ion.cu:304-358
//   frozen current b     solve_rkc.cu:369-395
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <utility>

namespace cbrkc {
namespace pkj {

constexpr int COMP = 1600;          // pkj.cuh:11
constexpr int SOMA = 0, MAIN = 1, SMOOTH = 2, SPINY = 3;
constexpr double V_INIT = -68.0, V_LEAK = -80.0, V_NA = 45.0, V_CA = 135.0,
                 V_KH = -30.0, V_K = -85.0;
constexpr double G_GJ = 2.5e-6;     // mS, pkj.cuh:26
constexpr int GJ_EACH = 26;
constexpr double SPINE_DENS = 13.0e4, SPINE_AREA = 1.33e-8;
constexpr double CM = 1.64, RM_S = 10.0, RM_D = 30.0, RI = 0.25;
constexpr double B_CA = 10.0, SHELLD = 0.2e-4, CA_OUT = 2.4, CA_0 = 4.0e-5, FARADAY = 9.6485e4;
constexpr double CURRENT = 0.1e-3;  // uA (0.1 nA), pkj.cuh:15

// gate slots, pkj.cuh:86-87
enum Gate { m_NaF, h_NaF, m_NaP, m_CaP, h_CaP, m_CaT, h_CaT, m_KA, h_KA, m_KC, m_Kh1, m_Kh2,
            m_Kdr, h_Kdr, m_KM, z_KC, m_K2, z_K2, NGATE };
// conductance slots, pkj.cuh:90-91
enum Cond { g_leak, g_NaF, g_NaP, g_CaP, g_CaT, g_Kh, g_Kdr, g_KM, g_KA, g_KC, g_K2, NCOND };

// densities [mS/cm^2] by compartment class (soma, main dendrite, rest),
// order NaF NaP CaP CaT KH KDR KM KA KC K2 (pkj.cuh:28-62)
constexpr double DENSITY[3][10] = {
    {7500.0, 1.0, 0.0, 0.5, 0.3, 600.0, 0.040, 15.0, 0.0, 0.0},
    {0.0, 0.0, 4.5, 0.5, 0.0, 60.0, 0.01, 2.0, 80.0, 0.39},
    {0.0, 0.0, 4.5, 0.5, 0.0, 0.0, 0.013, 0.0, 80.0, 0.39},
};

// c[g_NaF + Q] = DENSITY[row][Q] * area with the densities as immediates:
// selecting among literals keeps the table out of memory, so no dependent
// load sits in the ion update.
template <typename T, int Q>
__host__ __device__ inline T density_of(int row) {
  constexpr T d0 = T(DENSITY[0][Q]), d1 = T(DENSITY[1][Q]), d2 = T(DENSITY[2][Q]);
  return row == 0 ? d0 : (row == 1 ? d1 : d2);
}
template <typename T, size_t... Q>
__host__ __device__ inline void densities(int row, T area, T *c, std::index_sequence<Q...>) {
  ((c[1 + Q] = density_of<T, (int)Q>(row) * area), ...);
}

template <typename T> __host__ __device__ inline T ex(T x);
template <> __host__ __device__ inline double ex<double>(double x) { return exp(x); }
template <> __host__ __device__ inline float ex<float>(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __expf(x);
#else
  return expf(x);
#endif
}

// A / (B + exp((v + C) / D)), the reference's generic rate shape
template <typename T>
__host__ __device__ inline T lf(T v, T A, T B, T C, T D) { return A / (B + ex<T>((v + C) / D)); }

// alpha/beta pair of every two-rate gate, evaluated once; the reference
// evaluates them separately for tau and inf but the values are identical.
template <typename T> struct Rates { T a, b; };

template <typename T>
__host__ __device__ inline void cn_gate(T &x, T inf, T tau, T DT) {
  // (2 dt inf + (2 tau - dt) x) / (2 tau + dt): CN step of dx/dt = (inf - x)/tau
  x = (T(2.0) * DT * inf + (T(2.0) * tau - DT) * x) / (T(2.0) * tau + DT);
}

// exponential Euler, the BE path's gate update (pkj_ion.cu:112-129):
// x = inf + (x - inf) exp(-dt / tau)
template <typename T>
__host__ __device__ inline void ee_gate(T &x, T inf, T tau, T DT) {
  x = inf + (x - inf) * ex<T>(-DT / tau);
}

// one gate step by Crank-Nicolson (EE = false) or exponential Euler (EE = true)
template <typename T, bool EE>
__host__ __device__ inline void step_gate(T &x, T inf, T tau, T DT) {
  if constexpr (EE) ee_gate<T>(x, inf, tau, DT);
  else cn_gate<T>(x, inf, tau, DT);
}

template <typename T, bool EE = false>
__host__ __device__ inline void ab_gate(T &x, T a, T b, T DT) {
  T tau = T(1.0) / (a + b);
  T inf = a / (a + b);
  step_gate<T, EE>(x, inf, tau, DT);
}

template <typename T> __host__ __device__ inline T kdr_alpha(T u) {
  // -0.0235 u / (exp(-u/12) - 1), removable singularity at u = 0 (pkj_ion.cu:53-54)
  return (fabs(u) > T(1e-6)) ? (T(-0.0235) * u / (ex<T>(-u / T(12.0)) - T(1.0))) : T(0.0235 * 12.0);
}

// Steady state of every gate at voltage v and raw Ca (pkj_ion.cu:360-395).
template <typename T>
__host__ __device__ inline void steady_state(T v, T ca, T *g) {
  auto inf2 = [](T a, T b) { return a / (a + b); };
  g[m_NaF] = inf2(lf<T>(v, 35.0, 0.0, 5.0, -10.0), lf<T>(v, 7.0, 0.0, 65.0, 20.0));
  g[h_NaF] = inf2(lf<T>(v, 0.225, 1.0, 80.0, 10.0), lf<T>(v, 7.5, 0.0, -3.0, -18.0));
  g[m_NaP] = inf2(lf<T>(v, 200.0, 1.0, -18.0, -16.0), lf<T>(v, 25.0, 1.0, 58.0, 8.0));
  g[m_CaP] = inf2(lf<T>(v, 8.5, 1.0, -8.0, -12.5), lf<T>(v, 35.0, 1.0, 74.0, 14.5));
  g[h_CaP] = inf2(lf<T>(v, 0.0015, 1.0, 29.0, 8.0), lf<T>(v, 0.0055, 1.0, 23.0, -8.0));
  g[m_CaT] = inf2(lf<T>(v, 2.60, 1.0, 21.0, -8.0), lf<T>(v, 0.180, 1.0, 40.0, 4.0));
  g[h_CaT] = inf2(lf<T>(v, 0.0025, 1.0, 40.0, 8.0), lf<T>(v, 0.190, 1.0, 50.0, -10.0));
  g[m_Kh1] = T(0.8) / (T(1.0) + ex<T>((v + T(82.0)) / T(7.0)));
  g[m_Kh2] = T(0.2) / (T(1.0) + ex<T>((v + T(82.0)) / T(7.0)));
  {
    T a = kdr_alpha<T>(v - T(8.0));
    T b = T(5.0) * ex<T>(-(v + T(127.0)) / T(30.0));
    g[m_Kdr] = a / (a + b);
  }
  g[h_Kdr] = T(1.0) / (T(1.0) + ex<T>((v + T(25.0)) / T(4.0)));
  g[m_KM] = T(1.0) / (T(1.0) + ex<T>(-(v + T(35.0)) / T(10.0)));
  g[m_KA] = inf2(lf<T>(v, 1.40, 1.0, 27.0, -12.0), lf<T>(v, 0.490, 1.0, 30.0, 4.0));
  g[h_KA] = inf2(lf<T>(v, 0.0175, 1.0, 50.0, 8.0), lf<T>(v, 1.30, 1.0, 13.0, -10.0));
  g[m_KC] = inf2(T(7.5), lf<T>(v, 0.110, 0.0, -35.0, 14.9));
  g[z_KC] = T(1.0) / (T(1.0) + T(4.00) / ca);
  g[m_K2] = inf2(T(25.0), lf<T>(v, 0.075, 0.0, 25.0, 6.0));
  g[z_K2] = T(1.0) / (T(1.0) + T(0.20) / ca);
}

// Update of all 18 gates at voltage vm over dt with the quantised Ca caq:
// Crank-Nicolson (EE = false, pkj_ion.cu:337-354, vm the interval midpoint)
// or exponential Euler (EE = true, pkj_ion.cu:112-129, the BE path).
template <typename T, bool EE>
__host__ __device__ inline void gates_upd(T vm, T caq, T DT, T *g) {
  ab_gate<T, EE>(g[m_NaF], lf<T>(vm, 35.0, 0.0, 5.0, -10.0), lf<T>(vm, 7.0, 0.0, 65.0, 20.0), DT);
  ab_gate<T, EE>(g[h_NaF], lf<T>(vm, 0.225, 1.0, 80.0, 10.0), lf<T>(vm, 7.5, 0.0, -3.0, -18.0), DT);
  ab_gate<T, EE>(g[m_NaP], lf<T>(vm, 200.0, 1.0, -18.0, -16.0), lf<T>(vm, 25.0, 1.0, 58.0, 8.0), DT);
  ab_gate<T, EE>(g[m_CaP], lf<T>(vm, 8.5, 1.0, -8.0, -12.5), lf<T>(vm, 35.0, 1.0, 74.0, 14.5), DT);
  ab_gate<T, EE>(g[h_CaP], lf<T>(vm, 0.0015, 1.0, 29.0, 8.0), lf<T>(vm, 0.0055, 1.0, 23.0, -8.0), DT);
  ab_gate<T, EE>(g[m_CaT], lf<T>(vm, 2.60, 1.0, 21.0, -8.0), lf<T>(vm, 0.180, 1.0, 40.0, 4.0), DT);
  ab_gate<T, EE>(g[h_CaT], lf<T>(vm, 0.0025, 1.0, 40.0, 8.0), lf<T>(vm, 0.190, 1.0, 50.0, -10.0), DT);
  {
    T e = ex<T>((vm + T(82.0)) / T(7.0));
    step_gate<T, EE>(g[m_Kh1], T(0.8) / (T(1.0) + e), T(7.6), DT);
    step_gate<T, EE>(g[m_Kh2], T(0.2) / (T(1.0) + e), T(36.8), DT);
  }
  {
    T a = kdr_alpha<T>(vm - T(8.0));
    T b = T(5.0) * ex<T>(-(vm + T(127.0)) / T(30.0));
    T inf = a / (a + b);
    T a2 = kdr_alpha<T>(vm + T(12.0));
    T b2 = T(5.0) * ex<T>(-(vm + T(147.0)) / T(30.0));
    T tau = T(1.0) / (a2 + b2);
    step_gate<T, EE>(g[m_Kdr], inf, tau, DT);
  }
  step_gate<T, EE>(g[h_Kdr], T(1.0) / (T(1.0) + ex<T>((vm + T(25.0)) / T(4.0))),
             (vm < T(-25.0)) ? T(1200.0) : T(10.0), DT);
  step_gate<T, EE>(g[m_KM], T(1.0) / (T(1.0) + ex<T>(-(vm + T(35.0)) / T(10.0))),
             T(200.0) / (T(3.3) * (ex<T>((vm + T(35.0)) / T(20.0))) + ex<T>(-(vm + T(35.0)) / T(20.0))), DT);
  ab_gate<T, EE>(g[m_KA], lf<T>(vm, 1.40, 1.0, 27.0, -12.0), lf<T>(vm, 0.490, 1.0, 30.0, 4.0), DT);
  ab_gate<T, EE>(g[h_KA], lf<T>(vm, 0.0175, 1.0, 50.0, 8.0), lf<T>(vm, 1.30, 1.0, 13.0, -10.0), DT);
  ab_gate<T, EE>(g[m_KC], T(7.5), lf<T>(vm, 0.110, 0.0, -35.0, 14.9), DT);
  step_gate<T, EE>(g[z_KC], T(1.0) / (T(1.0) + T(4.00) / caq), T(10.0), DT);
  ab_gate<T, EE>(g[m_K2], T(25.0), lf<T>(vm, 0.075, 0.0, 25.0, 6.0), DT);
  step_gate<T, EE>(g[z_K2], T(1.0) / (T(1.0) + T(0.20) / caq), T(10.0), DT);
}

template <typename T>
__host__ __device__ inline void gates_cn(T vm, T caq, T DT, T *g) {
  gates_upd<T, false>(vm, caq, DT, g);
}

// Ca shell update (RK4 over dt of dCa/dt = I - B (Ca - Ca0)), the reversal
// potential and the quantised Ca for KC/K2 (pkj_ion.cu:312-335).  Always
// FP64: the quantisation floor() is discontinuous in Ca.
//   fixed_rev: the reference keys the fixed somatic reversal on the global
//              compartment index being a multiple of 1599.
// Returns the quantised Ca; updates ca and rev.
__host__ __device__ inline double ca_rk4(double v_old, double &ca, double &rev, double shell,
                                         double gCaP_open, double gCaT_open, double DT, bool fixed_rev) {
  double Ca_val = ca;
  double I_Ca1 = 1e-3 * (v_old - rev) * (gCaP_open + gCaT_open);
  (I_Ca1 > 0.0) ? (I_Ca1 = 0.0) : (I_Ca1 = -I_Ca1 / (2.0 * FARADAY * shell));
  double k1 = DT * (I_Ca1 - B_CA * (Ca_val - CA_0));
  double k2 = DT * (I_Ca1 - B_CA * (Ca_val + k1 / 2.0 - CA_0));
  double k3 = DT * (I_Ca1 - B_CA * (Ca_val + k2 / 2.0 - CA_0));
  double k4 = DT * (I_Ca1 - B_CA * (Ca_val + k3 - CA_0));
  k1 = Ca_val + (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  ca = k1;
  Ca_val = 0.5 * (Ca_val + k1);
  rev = fixed_rev ? 12.5 * log(CA_OUT / CA_0) : 13.361624877 * log(CA_OUT / Ca_val);
  return floor((Ca_val * 1000 - 0.04) / 0.099986667) * 0.099986667 + 0.04;
}

// pkj_update_ion (pkj_ion.cu:94-151), the BE path: every gate by
// exponential Euler at V (KC / K2 at the quantised present Ca), then the Ca
// shell by RK4 with the present gates' Ca current clamped at zero from above
// and divided by shell x area -- the stored shell volume already carries the
// area, so the area enters twice, as in the reference -- and the reversal
// from the present Ca, fixed on compartments of the soma type.
__host__ __device__ inline void update_ion_be(double v, double &ca, double &rev, double shell, double area,
                                              bool soma, double gCaP_open, double gCaT_open, double DT,
                                              double *g) {
  const double Ca_val = ca;
  const double cafloor = floor((Ca_val * 1000 - 0.04) / 0.099986667) * 0.099986667 + 0.04;
  double I_Ca1 = 1e-3 * (v - rev) * (gCaP_open + gCaT_open);
  if (I_Ca1 > 0.0) I_Ca1 = 0.0;
  gates_upd<double, true>(v, cafloor, DT, g);
  const double l_shell = shell * area;
  const double k1 = DT * (-I_Ca1 / (2.0 * FARADAY * l_shell) - B_CA * (Ca_val - CA_0));
  const double k2 = DT * (-I_Ca1 / (2.0 * FARADAY * l_shell) - B_CA * (Ca_val + k1 / 2.0 - CA_0));
  const double k3 = DT * (-I_Ca1 / (2.0 * FARADAY * l_shell) - B_CA * (Ca_val + k2 / 2.0 - CA_0));
  const double k4 = DT * (-I_Ca1 / (2.0 * FARADAY * l_shell) - B_CA * (Ca_val + k3 - CA_0));
  ca += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  rev = soma ? 12.5 * log(CA_OUT / CA_0) : 13.361624877 * log(CA_OUT / Ca_val);
}

// ---------------------------------------------------------------------------
// FP32 fast path of the ion update for the mixed-precision engine.  Same
// model, rearranged for the gfx950 VALU: every division is one hardware
// reciprocal (v_rcp_f32) refined by one Newton step, every exp((v + C) / D)
// is one v_exp_f32 of (v + C) * (log2 e / D) with the constant folded, and
// the Crank-Nicolson step of a two-rate gate is written on the rates,
//   x' = (2 dt a + (2 - dt s) x) / (2 + dt s),   s = a + b,
// which is the reference's (2 dt inf + (2 tau - dt) x) / (2 tau + dt) with
// tau = 1/s, inf = a/s multiplied through by s: one reciprocal instead of
// three divisions per gate.  The Newton step (r + r (1 - x r), two FMAs)
// matters: the bare v_rcp_f32 result (1 ulp) carries a bias that the gate
// recurrence accumulates over thousands of updates; with it the mixed
// engine's voltage error against FP64 on the 1024-cell 200 ms workload fell
// from 0.103 to 0.085 mV (median per cell) at no measurable cost.
#ifdef __HIP_DEVICE_COMPILE__
__device__ inline float frcp(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }
#else
inline float frcp(float x) { return 1.0f / x; }
inline float fexp2(float x) { return exp2f(x); }
inline float flog2(float x) { return log2f(x); }
#endif

constexpr float LOG2E = 1.4426950408889634f;

// A / (B + exp((v + C) / D))
__host__ __device__ inline float lff(float v, float A, float B, float C, float D) {
  return A * frcp(B + fexp2((v + C) * (LOG2E / D)));
}

__host__ __device__ inline void cn_rates(float &x, float a, float b, float dt) {
  const float s = a + b;
  x = (2.0f * dt * a + (2.0f - dt * s) * x) * frcp(2.0f + dt * s);
}

__host__ __device__ inline void cn_inf_tau(float &x, float inf, float tau, float dt) {
  x = (2.0f * dt * inf + (2.0f * tau - dt) * x) * frcp(2.0f * tau + dt);
}

__host__ __device__ inline float kdr_alpha_f(float u) {
  return (fabsf(u) > 1e-6f) ? -0.0235f * u * frcp(fexp2(-u * (LOG2E / 12.0f)) - 1.0f) : 0.0235f * 12.0f;
}

// Channels present in a compartment class: bit Q set when DENSITY[row][Q]
// is nonzero (Q in the order NaF NaP CaP CaT KH KDR KM KA KC K2).  A channel
// of zero density adds exactly zero to the current and to the Ca influx, so
// the mixed engine neither updates nor streams its gates there: a dendrite
// advances 9 of the 18 gates, the soma 12, a main-dendrite compartment 13,
// with every voltage, current and Ca value unchanged.  (The skipped gates
// keep their initial values in the slab.)
enum ChanBit : unsigned {
  CH_NAF = 1u << 0, CH_NAP = 1u << 1, CH_CAP = 1u << 2, CH_CAT = 1u << 3, CH_KH = 1u << 4,
  CH_KDR = 1u << 5, CH_KM = 1u << 6, CH_KA = 1u << 7, CH_KC = 1u << 8, CH_K2 = 1u << 9, CH_ALL = 1023u
};
constexpr unsigned channels_of_row(int row) {
  unsigned m = 0;
  for (int q = 0; q < 10; q++)
    if (DENSITY[row][q] != 0.0) m |= 1u << q;
  return m;
}
// the channel a gate slot belongs to
constexpr unsigned chan_of_gate(int q) {
  return (q == m_NaF || q == h_NaF) ? CH_NAF
         : (q == m_NaP)             ? CH_NAP
         : (q == m_CaP || q == h_CaP) ? CH_CAP
         : (q == m_CaT || q == h_CaT) ? CH_CAT
         : (q == m_KA || q == h_KA)   ? CH_KA
         : (q == m_KC || q == z_KC)   ? CH_KC
         : (q == m_Kh1 || q == m_Kh2) ? CH_KH
         : (q == m_Kdr || q == h_Kdr) ? CH_KDR
         : (q == m_KM)                ? CH_KM
                                      : CH_K2;
}

template <unsigned M = CH_ALL>
__host__ __device__ inline void gates_cn_fast(float vm, float caq, float dt, float *g) {
  if constexpr ((M & CH_NAF) != 0) {
    cn_rates(g[m_NaF], lff(vm, 35.0f, 0.0f, 5.0f, -10.0f), lff(vm, 7.0f, 0.0f, 65.0f, 20.0f), dt);
    cn_rates(g[h_NaF], lff(vm, 0.225f, 1.0f, 80.0f, 10.0f), lff(vm, 7.5f, 0.0f, -3.0f, -18.0f), dt);
  }
  if constexpr ((M & CH_NAP) != 0)
    cn_rates(g[m_NaP], lff(vm, 200.0f, 1.0f, -18.0f, -16.0f), lff(vm, 25.0f, 1.0f, 58.0f, 8.0f), dt);
  if constexpr ((M & CH_CAP) != 0) {
    cn_rates(g[m_CaP], lff(vm, 8.5f, 1.0f, -8.0f, -12.5f), lff(vm, 35.0f, 1.0f, 74.0f, 14.5f), dt);
    cn_rates(g[h_CaP], lff(vm, 0.0015f, 1.0f, 29.0f, 8.0f), lff(vm, 0.0055f, 1.0f, 23.0f, -8.0f), dt);
  }
  if constexpr ((M & CH_CAT) != 0) {
    cn_rates(g[m_CaT], lff(vm, 2.60f, 1.0f, 21.0f, -8.0f), lff(vm, 0.180f, 1.0f, 40.0f, 4.0f), dt);
    cn_rates(g[h_CaT], lff(vm, 0.0025f, 1.0f, 40.0f, 8.0f), lff(vm, 0.190f, 1.0f, 50.0f, -10.0f), dt);
  }
  if constexpr ((M & CH_KH) != 0) {
    const float r = frcp(1.0f + fexp2((vm + 82.0f) * (LOG2E / 7.0f)));
    cn_inf_tau(g[m_Kh1], 0.8f * r, 7.6f, dt);
    cn_inf_tau(g[m_Kh2], 0.2f * r, 36.8f, dt);
  }
  if constexpr ((M & CH_KDR) != 0) {
    const float a = kdr_alpha_f(vm - 8.0f);
    const float b = 5.0f * fexp2(-(vm + 127.0f) * (LOG2E / 30.0f));
    const float a2 = kdr_alpha_f(vm + 12.0f);
    const float b2 = 5.0f * fexp2(-(vm + 147.0f) * (LOG2E / 30.0f));
    cn_inf_tau(g[m_Kdr], a * frcp(a + b), frcp(a2 + b2), dt);
    cn_inf_tau(g[h_Kdr], frcp(1.0f + fexp2((vm + 25.0f) * (LOG2E / 4.0f))), (vm < -25.0f) ? 1200.0f : 10.0f, dt);
  }
  if constexpr ((M & CH_KM) != 0) {
    const float e = fexp2((vm + 35.0f) * (LOG2E / 20.0f));   // exp((v + 35) / 20)
    const float ei = frcp(e);
    cn_inf_tau(g[m_KM], frcp(1.0f + ei * ei), 200.0f * frcp(3.3f * e + ei), dt);
  }
  if constexpr ((M & CH_KA) != 0) {
    cn_rates(g[m_KA], lff(vm, 1.40f, 1.0f, 27.0f, -12.0f), lff(vm, 0.490f, 1.0f, 30.0f, 4.0f), dt);
    cn_rates(g[h_KA], lff(vm, 0.0175f, 1.0f, 50.0f, 8.0f), lff(vm, 1.30f, 1.0f, 13.0f, -10.0f), dt);
  }
  if constexpr ((M & CH_KC) != 0) {
    cn_rates(g[m_KC], 7.5f, lff(vm, 0.110f, 0.0f, -35.0f, 14.9f), dt);
    cn_inf_tau(g[z_KC], frcp(1.0f + 4.00f * frcp(caq)), 10.0f, dt);
  }
  if constexpr ((M & CH_K2) != 0) {
    cn_rates(g[m_K2], 25.0f, lff(vm, 0.075f, 0.0f, 25.0f, 6.0f), dt);
    cn_inf_tau(g[z_K2], frcp(1.0f + 0.20f * frcp(caq)), 10.0f, dt);
  }
}

// Ca shell update for the mixed engine: the Ca state, its RK4 and its
// quantisation stay FP64 (the quantisation floor is discontinuous in Ca);
// divisions become products with host-computed reciprocals and the
// reversal potential uses an FP32 log.  inv2fs = 1 / (2 F shell).
__host__ __device__ inline double ca_rk4_mixed(float v_old, double &ca, float &rev, double inv2fs, float gopen,
                                               double DT, bool fixed_rev) {
  double Ca_val = ca;
  double I_Ca1 = 1e-3 * (double)((v_old - rev) * gopen);
  I_Ca1 = (I_Ca1 > 0.0) ? 0.0 : -I_Ca1 * inv2fs;
  const double k1 = DT * (I_Ca1 - B_CA * (Ca_val - CA_0));
  const double k2 = DT * (I_Ca1 - B_CA * (Ca_val + 0.5 * k1 - CA_0));
  const double k3 = DT * (I_Ca1 - B_CA * (Ca_val + 0.5 * k2 - CA_0));
  const double k4 = DT * (I_Ca1 - B_CA * (Ca_val + k3 - CA_0));
  const double cn = Ca_val + (k1 + k2 * 2.0 + k3 * 2.0 + k4) * (1.0 / 6.0);
  ca = cn;
  Ca_val = 0.5 * (Ca_val + cn);
  // 13.361624877 ln(2.4 / Ca) = 13.361624877 ln 2 log2(2.4 / Ca)
  rev = fixed_rev ? 137.52624801505297f /* 12.5 ln(2.4 / 4e-5) */
                  : 9.261572611192175f * flog2((float)(CA_OUT / Ca_val));
  return floor((Ca_val * 1000 - 0.04) * (1.0 / 0.099986667)) * 0.099986667 + 0.04;
}

// Frozen membrane current b at voltage v (solve_rkc.cu:369-395), summed in
// the reference's term order.
template <typename T>
__host__ __device__ inline T current(T v, T iext, T rev, const T *c, const T *g) {
  return (-iext
          + c[g_leak] * (v - T(V_LEAK))
          + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF] * (v - T(V_NA))
          + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP] * (v - T(V_NA))
          + c[g_CaP] * g[m_CaP] * g[h_CaP] * (v - rev)
          + c[g_CaT] * g[m_CaT] * g[h_CaT] * (v - rev)
          + c[g_Kh] * g[m_Kh1] * (v - T(V_KH))
          + c[g_Kh] * g[m_Kh2] * (v - T(V_KH))
          + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr] * (v - T(V_K))
          + c[g_KM] * g[m_KM] * (v - T(V_K))
          + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA] * (v - T(V_K))
          + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC] * (v - T(V_K))
          + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2] * (v - T(V_K)));
}

// current() over the channels of mask M only, in the same term order: the
// skipped terms are exact zeros (zero density), so the sum is unchanged
template <unsigned M, typename T>
__host__ __device__ inline T current_m(T v, T iext, T rev, const T *c, const T *g) {
  T s = -iext + c[g_leak] * (v - T(V_LEAK));
  if constexpr ((M & CH_NAF) != 0) s = s + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF] * (v - T(V_NA));
  if constexpr ((M & CH_NAP) != 0) s = s + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP] * (v - T(V_NA));
  if constexpr ((M & CH_CAP) != 0) s = s + c[g_CaP] * g[m_CaP] * g[h_CaP] * (v - rev);
  if constexpr ((M & CH_CAT) != 0) s = s + c[g_CaT] * g[m_CaT] * g[h_CaT] * (v - rev);
  if constexpr ((M & CH_KH) != 0) {
    s = s + c[g_Kh] * g[m_Kh1] * (v - T(V_KH));
    s = s + c[g_Kh] * g[m_Kh2] * (v - T(V_KH));
  }
  if constexpr ((M & CH_KDR) != 0) s = s + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr] * (v - T(V_K));
  if constexpr ((M & CH_KM) != 0) s = s + c[g_KM] * g[m_KM] * (v - T(V_K));
  if constexpr ((M & CH_KA) != 0) s = s + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA] * (v - T(V_K));
  if constexpr ((M & CH_KC) != 0) s = s + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC] * (v - T(V_K));
  if constexpr ((M & CH_K2) != 0) s = s + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2] * (v - T(V_K));
  return s;
}

// CN path: the ionic conductance sum and the conductance-weighted reversal
// sum plus the injected current, before halving, in the reference's term
// order (pkj_cnm_vec_initialize / pkj_update_matrix, solve_cnm.cu:429-456,
// 480-509)
__host__ __device__ inline void cn_ge(const double *c, const double *g, double rev, double iext, double &gs,
                                      double &es) {
  gs = (+c[g_leak]
        + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF]
        + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP]
        + c[g_CaP] * g[m_CaP] * g[h_CaP]
        + c[g_CaT] * g[m_CaT] * g[h_CaT]
        + c[g_Kh] * g[m_Kh1]
        + c[g_Kh] * g[m_Kh2]
        + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr]
        + c[g_KM] * g[m_KM]
        + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA]
        + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC]
        + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2]);
  es = (+c[g_leak] * (V_LEAK) + iext
        + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF] * (V_NA)
        + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP] * (V_NA)
        + c[g_CaP] * g[m_CaP] * g[h_CaP] * (rev)
        + c[g_CaT] * g[m_CaT] * g[h_CaT] * (rev)
        + c[g_Kh] * g[m_Kh1] * (V_KH)
        + c[g_Kh] * g[m_Kh2] * (V_KH)
        + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr] * (V_K)
        + c[g_KM] * g[m_KM] * (V_K)
        + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA] * (V_K)
        + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC] * (V_K)
        + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2] * (V_K));
}

// BE path: the system diagonal addend and right-hand side with the
// capacitive part first (pkj_update_matrix, solve_bem.cu:407-445)
__host__ __device__ inline void be_ge(const double *c, const double *g, double rev, double iext, double cmdt,
                                      double v, double &gs, double &es) {
  gs = (cmdt
        + c[g_leak]
        + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF]
        + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP]
        + c[g_CaP] * g[m_CaP] * g[h_CaP]
        + c[g_CaT] * g[m_CaT] * g[h_CaT]
        + c[g_Kh] * g[m_Kh1]
        + c[g_Kh] * g[m_Kh2]
        + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr]
        + c[g_KM] * g[m_KM]
        + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA]
        + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC]
        + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2]);
  es = (cmdt * v + iext
        + c[g_leak] * (V_LEAK)
        + c[g_NaF] * g[m_NaF] * g[m_NaF] * g[m_NaF] * g[h_NaF] * (V_NA)
        + c[g_NaP] * g[m_NaP] * g[m_NaP] * g[m_NaP] * (V_NA)
        + c[g_CaP] * g[m_CaP] * g[h_CaP] * (rev)
        + c[g_CaT] * g[m_CaT] * g[h_CaT] * (rev)
        + c[g_Kh] * g[m_Kh1] * (V_KH)
        + c[g_Kh] * g[m_Kh2] * (V_KH)
        + c[g_Kdr] * g[m_Kdr] * g[m_Kdr] * g[h_Kdr] * (V_K)
        + c[g_KM] * g[m_KM] * (V_K)
        + c[g_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[m_KA] * g[h_KA] * (V_K)
        + c[g_KC] * g[m_KC] * g[z_KC] * g[z_KC] * (V_K)
        + c[g_K2] * g[m_K2] * g[z_K2] * g[z_K2] * (V_K));
}

}  // namespace pkj
}  // namespace cbrkc
