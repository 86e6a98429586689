// Compartment state of the granule, Golgi and inferior-olive models in the
// FP64 structure-of-arrays layout ([field][nc] for gates and conductances),
// and the per-compartment ion update + frozen current of the RKC path, shared
// by the population-global RKC engine's kernels and the per-cell batched
// engine.
//   io_ion_current  io_update_ion_RKC (io_ion.cu:90-127) + io_update_matrix
//                   (solve_rkc.cu:412-434)
//   go_ion_current  go_update_ion_RKC_exp_imp (go_ion.cu:379-437),
//                   go_KAHP_update_2order (go_ion.cu:88-213) + go_update_matrix
//                   (solve_rkc.cu:288-314)
//   gr_ion_current  gr_Na_update_2order (gr_ion.cu:45-207),
//                   gr_update_ion_RKC_exp_imp (gr_ion.cu:461-480) +
//                   gr_update_matrix (solve_rkc.cu:226-249)
// Each updates the compartment's gates, Ca and stored voltage when do_ion is
// set and returns the frozen membrane current at vn, before any gap-junction
// or synaptic term (the callers add those in the reference's order).  Without
// with_iext the injected current counts as zero: the reference initialises
// RKC (RKC_vec_initialize) before the first set_current of the output loop
// (main.cu), so its initial frozen current has no stimulus.
#pragma once
#include <hip/hip_runtime.h>

#include "go_model.hpp"
#include "gr_model.hpp"
#include "io_model.hpp"
#include "synapse.hpp"

namespace cbrkc {

struct IoState {
  double *v, *ca, *iext;
  const double *cond;   // [NCOND][nc]
  double *gate;         // [NGATE][nc]
  const int8_t *compart;
  const int *nbr;       // [n][4] gap-junction partner cell or -1
  double *b;
  const double *syn_g;  // may be null
  double syn_e;
  int nc, n;
};

struct GoState {
  double *v, *ca, *ca2, *rev, *ca_old, *iext;
  const double *area;
  const double *cond;   // [NCOND][nc], density x area
  double *gate;         // [NGATE][nc]
  double *b;
  const double *syn_g;  // may be null
  double syn_e;
  int nc;
  const double *net_s;  // synaptic current of the macro step (granule input), may be null
};

struct GrState {
  double *v, *iext;
  const double *ca;
  const int8_t *compart;
  const double *cond;   // [NCOND][nc], density x area
  double *gate;         // [NGATE][nc]
  const gr::NaType *na; // per compartment type
  double *b;
  const double *syn_g;  // may be null
  double syn_e;
  int nc;
  SynIn net;            // mossy-fibre and Golgi input, per compartment
};

__device__ inline double io_ion_current(const IoState &S, int i, double vn, double DT, bool do_ion,
                                  bool with_iext = true) {
  using namespace io;
  const int nc = S.nc;
  double g[NGATE], c[NCOND];
#pragma unroll
  for (int k = 0; k < NGATE; k++) g[k] = S.gate[(size_t)k * nc + i];
#pragma unroll
  for (int k = 0; k < NCOND; k++) c[k] = S.cond[(size_t)k * nc + i];
  if (do_ion) {
    double ca = S.ca[i];
    update_ion<double>(S.v[i], vn, ca, c[g_CaH], S.compart[i], DT, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) S.gate[(size_t)k * nc + i] = g[k];
    S.ca[i] = ca;
    S.v[i] = vn;
  }
  return current<double>(vn, with_iext ? S.iext[i] : 0.0, c, g);
}

__device__ inline double go_ion_current(const GoState &S, int i, double vn, double DT, bool do_ion,
                                  bool with_iext = true) {
  using namespace go;
  const int nc = S.nc;
  double g[NGATE], c[NCOND];
#pragma unroll
  for (int k = 0; k < NGATE; k++) g[k] = S.gate[(size_t)k * nc + i];
#pragma unroll
  for (int k = 0; k < NCOND; k++) c[k] = S.cond[(size_t)k * nc + i];
  double rev = S.rev[i];
  if (do_ion) {
    double ca = S.ca[i], ca2 = S.ca2[i];
    update_ion(S.v[i], vn, S.area[i], c, ca, ca2, rev, DT, g);
    const double ca_prev = S.ca_old[i];
    kahp_step(ca_prev, ca, DT, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) S.gate[(size_t)k * nc + i] = g[k];
    S.ca[i] = ca;
    S.ca2[i] = ca2;
    S.rev[i] = rev;
    S.ca_old[i] = ca;
    S.v[i] = vn;
  }
  return current(vn, with_iext ? S.iext[i] : 0.0, rev, c, g);
}

__device__ inline double gr_ion_current(const GrState &S, int i, double vn, double DT, bool do_ion,
                                  bool with_iext = true) {
  using namespace gr;
  const int nc = S.nc;
  double g[NGATE], c[NCOND];
#pragma unroll
  for (int k = 0; k < NGATE; k++) g[k] = S.gate[(size_t)k * nc + i];
#pragma unroll
  for (int k = 0; k < NCOND; k++) c[k] = S.cond[(size_t)k * nc + i];
  if (do_ion) {
    const double vo = S.v[i];
    const int t = S.compart[i];
    if (t >= HILL) {
      const NaType T = S.na[t];
      double x[12];
      na_gather(g, x);
      if (na_step(T, vo, vn, DT, x)) na_scatter(x, g);
    }
    update_gates(vo, vn, S.ca[i], DT, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) S.gate[(size_t)k * nc + i] = g[k];
    S.v[i] = vn;
  }
  return current(vn, with_iext ? S.iext[i] : 0.0, c, g);
}

}  // namespace cbrkc
