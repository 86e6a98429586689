// Synapses between the populations (syn.cu): connection rules, the per-macro-
// step trace updates and the incoming-synapse view the engines couple.
//
//   kind          pre -> post                     rule (syn.cu)
//   SYN_MFGR      Poisson mossy fibre -> GR 0..3  4 per granule cell (:42-128)
//   SYN_GOGR      GO comp 30 -> GR comp 16        |dx|, |dy| < 32 um (:292-410)
//   SYN_GRGO      GR axon -> GO comp 0            |dx| < 16, |dy| < 220 um (:144-261)
//   SYN_GRPKJ     GR axon -> PKJ dendrite         |dx| < 120, |dy| < 220 um, 25 per pair (:440-644)
//   SYN_MLIPKJ    Poisson MLI -> PKJ comp 1599    weight 0, never updated or coupled (:666-759)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace cbrkc {

enum SynKind { SYN_MFGR = 0, SYN_GOGR, SYN_GRGO, SYN_GRPKJ, SYN_MLIPKJ };

// Connections and weights of one synapse set, built on the host.
struct SynHost {
  int kind = SYN_MFGR, n = 0;
  std::vector<int> pre, post;       // global compartment indices (pre unused for MF / MLI)
  std::vector<double> weight;
  double E = 0.0;                   // reversal potential [mV]
  double decay[3] = {0, 0, 0};      // exp(-0.125 / tau) per trace
};

// Builds a set from the two grids (nx_a x ny_a the first population named by
// the kind, nx_b x ny_b the second: MFGR a = GR; GOGR a = GO, b = GR; GRGO
// a = GR, b = GO; GRPKJ a = GR, b = PKJ, with x1600 the x column of
// Pkj_location_info.csv and seed the key of the contact jitter; MLIPKJ a =
// PKJ, b = GR).  Empty when either population is.
void build_syn(int kind, int nx_a, int ny_a, int nx_b, int ny_b, const double *x1600, uint64_t seed,
               SynHost &out);

// Device view of a set for the update kernel.
struct SynDev {
  int kind, n;
  const int *pre;
  double *x0, *x1, *x2, *val, *old_v;
  const double *weight;
  double d0, d1, d2;
  uint32_t k0, k1;
};

// One update after the macro step starting at t (xx_update, syn.cu), step =
// the number of updates done before (the Philox counter of the draws).
void syn_update(const SynDev &s, const double *pre_v, double t, uint32_t step, hipStream_t st);

// Incoming synapses of a population as a CSR by postsynaptic compartment:
// entry j of compartment i (row[i] <= j < row[i + 1]) is synapse
// slot[j] & 0x3fffffff of attached set slot[j] >> 30, with conductance
// val[set][k] [mS] and reversal E[set] [mV]; a compartment's entries come in
// set order, then synapse order (the reference's launch order, its atomics
// made sequential).
struct SynIn {
  const int *row = nullptr, *slot = nullptr;
  const double *val[2] = {nullptr, nullptr};
  double E[2] = {0.0, 0.0};
  __host__ __device__ bool on() const { return row != nullptr; }
};

__device__ inline double syn_entry(const SynIn &S, int j, double &E) {
  const int s = S.slot[j];
  const int t = s >> 30;
  E = S.E[t];
  return S.val[t][s & 0x3fffffff];
}

// sum over the compartment's entries of val (v - E), in entry order, from 0
__device__ inline double syn_current(const SynIn &S, int i, double v) {
  double acc = 0.0;
  for (int j = S.row[i]; j < S.row[i + 1]; j++) {
    double E;
    const double g = syn_entry(S, j, E);
    acc += g * (v - E);
  }
  return acc;
}

// s[i] = sum val (y - E) over the population (the Golgi / Purkinje RKC paths'
// synaptic current of a macro step, from its starting y)
void syn_macro_current(const SynIn &S, const double *y, double *s, int nc, hipStream_t st);

}  // namespace cbrkc
