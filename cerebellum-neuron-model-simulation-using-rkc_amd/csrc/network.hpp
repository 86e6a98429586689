// A network of granule, Golgi and Purkinje populations in the shape of the
// reference's network run (main.cu:182-311 with network_sim on, syn.cu:822-917)
// at a fixed in-degree instead of the reference's geometric rules:
//
//   kind        pre -> post                  receptor   target compartment
//   NET_MFGR    mossy fibre -> granule       AMPA       one per granule cell (its dendrite)
//   NET_GOGR    Golgi -> granule             GABA_A     the same
//   NET_GRGO    granule -> Golgi             AMPA       uniform in a compartment range
//   NET_GRPKJ   granule -> Purkinje          AMPA       uniform in a compartment range
//
// Every postsynaptic cell receives deg[kind] inputs; with total > 0 the first
// Purkinje cells receive one granule input more each, so the list holds exactly
// `total` synapses.  Presynaptic cells, weights and target compartments are
// Philox draws keyed by the seed with counter (post cell, input, kind, 0), so
// any synapse is available without generating the others.
//
// The conductance of a (compartment, receptor) slot is the sum of its
// synapses' weighted traces; the traces share the receptor's time constant, so
// the slot keeps one state: G <- G exp(-dt / tau) + the weights of the inputs
// whose presynaptic spike arrives at this interval (the reference's per-synapse
// trace update s <- s exp(-dt / tau) + spike, val = w s, syn.cu:87-111,
// 264-288, 413-437, 647-668, summed over a compartment's synapses).  A spike is
// an upward crossing of a threshold by a cell's spike compartment between two
// interval ends (the reference's pre V > 0 && old_v < 0 test); it arrives a
// whole number of intervals later (the delay), at the start of an interval.
// Mossy fibres fire with probability rate x interval per interval
// (mfgr_update, syn.cu:87-102), a Philox draw per (synapse, interval).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "engine.hpp"

namespace cbrkc {

enum NetKind { NET_MFGR = 0, NET_GOGR = 1, NET_GRGO = 2, NET_GRPKJ = 3, NET_NKIND = 4 };
enum NetPop { NP_GR = 0, NP_GO = 1, NP_PKJ = 2, NP_N = 3 };
enum NetReceptor { NR_AMPA = 0, NR_GABA = 1 };

constexpr int net_post_pop(int kind) { return kind <= NET_GOGR ? NP_GR : kind == NET_GRGO ? NP_GO : NP_PKJ; }
// presynaptic population (-1: mossy fibres, outside the populations)
constexpr int net_pre_pop(int kind) { return kind == NET_MFGR ? -1 : kind == NET_GOGR ? NP_GO : NP_GR; }
constexpr int net_receptor(int kind) { return kind == NET_GOGR ? NR_GABA : NR_AMPA; }

struct NetSpec {
  uint64_t seed = 4;
  int deg[NET_NKIND] = {4, 4, 100, 1392};
  long long total = 780404;
  double w_min_nS = 0.5, w_max_nS = 1.5;
  double tau_ms[2] = {1.5, 6.0};     // AMPA, GABA_A
  double e_mV[2] = {0.0, -82.0};     // E_GOGR -82 mV (syn.cuh)
  double delay_ms = 1.0;
  double mf_rate_hz = 5.0;
  double interval_ms = 0.125;        // the populations' macro step (param.h:12)
  int n[NP_N] = {0, 0, 0}, comp[NP_N] = {0, 0, 0};
  int tgt_lo[NP_N] = {0, 31, 0}, tgt_hi[NP_N] = {1, 131, 1599};   // [lo, hi) target compartments
  int spike_comp[NP_N] = {3, 30, 1599};   // granule: the chain end of a 4-compartment chain
  double thr_mV[NP_N] = {0.0, 0.0, -10.0};
};

// The synapse list in generation order (kind, then postsynaptic cell, then
// input): pre = presynaptic cell (mossy fibres: the fibre's index, one fibre
// per synapse), post = compartment index within the postsynaptic population.
struct NetList {
  std::vector<int8_t> kind;
  std::vector<int> pre, post;
  std::vector<double> w_nS;
  long long count[NET_NKIND] = {0, 0, 0, 0};
};

// checks the spec and the sizes; "" when they hold
std::string net_check(const NetSpec &s);
// the number of synapses the spec gives (-1 when the check fails)
long long net_size(const NetSpec &s);
// generates the list (Error on a spec that fails the check)
void net_generate(const NetSpec &s, NetList &out);

// The network on the device over three populations' engines (not owned; they
// must outlive it and must not take other synaptic input meanwhile).
class Network {
 public:
  Network(Engine *gr, Engine *go, Engine *pkj, const NetSpec &s, int device);
  ~Network();
  // nint intervals: per interval and population, the slots' conductances take
  // the decay and the arriving spikes, the population advances one output
  // interval of its solver (Engine::run of one interval), and its spikes go
  // into the ring; the three populations run concurrently for up to delay + 1
  // intervals between exchanges
  void run(int nint);
  // moves up to cap spikes (time order; granule, Golgi, Purkinje within a
  // time; cells ascending) out of the raster; returns how many
  int take_spikes(int cap, double *t, int *pop, int *cell);
  // the conductance [mS] of every compartment of a population for one
  // receptor (0 where it has no synapse)
  void conductance(int pop, int receptor, double *out) const;
  long long count(int kind) const { return list_count_[kind]; }
  long long slots(int pop) const { return post_[pop].nslot; }
  long intervals() const { return k_; }
  int delay_intervals() const { return delay_; }
  // host wall time [ms] of the runs so far: each population's own interval
  // loop (its engine's work and waits, in its thread), the exchanges, and the
  // whole of every run
  void times(double *pop3, double *exchange, double *total) const;

 private:
  struct Post {
    int nslot = 0;
    std::vector<int> slot_comp;        // compartment of each slot
    std::vector<int8_t> slot_rec;      // receptor of each slot
    int *row = nullptr, *pre = nullptr;   // [nslot + 1], [synapses]: ring bit or ~fibre
    double *w = nullptr;               // [synapses] mS
    int8_t *rec = nullptr;
    double *G = nullptr;               // [nslot]
    int *in_row = nullptr, *in_slot = nullptr;   // the engine's view (SynIn)
  };
  void interval(int p, long k);
  void free_all();

  Engine *eng_[NP_N];
  NetSpec spec_;
  int device_, delay_ = 8, words_ = 0, word_off_[NP_N] = {0, 0, 0}, nring_ = 0;
  long long list_count_[NET_NKIND] = {0, 0, 0, 0};
  Post post_[NP_N];
  uint32_t *ring_ = nullptr;           // [nring][words]
  double *prev_[NP_N] = {nullptr, nullptr, nullptr};
  long k_ = 0;
  std::vector<std::pair<double, std::pair<int, int>>> spikes_;
  double ms_pop_[NP_N] = {0.0, 0.0, 0.0}, ms_exchange_ = 0.0, ms_total_ = 0.0;
};

}  // namespace cbrkc
