// Fixed in-degree network between the granule, Golgi and Purkinje
// populations: host generator, device slot conductances, spike ring and the
// exchange loop (the reference's network drivers, syn.cu:822-917, and the
// per-macro-step order of main.cu:182-311 with network_sim on).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <thread>

#include "network.hpp"
#include "philox.hpp"

namespace cbrkc {
namespace {

constexpr uint32_t KEY_NET = 0x6e657477u;   // "netw": connections
constexpr uint32_t KEY_MF = 0x6d666472u;    // "mfdr": mossy-fibre firing

template <typename T>
T *dnew(size_t n) {
  void *p = nullptr;
  CB_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)));
  return (T *)p;
}

template <typename T>
T *dcopy(const std::vector<T> &h) {
  T *p = dnew<T>(h.size());
  if (!h.empty()) CB_HIP(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return p;
}

long long extra_inputs(const NetSpec &s) {
  long long base = 0;
  for (int k = 0; k < NET_NKIND; k++) base += (long long)s.deg[k] * s.n[net_post_pop(k)];
  return s.total > 0 ? s.total - base : 0;
}

// inputs of postsynaptic cell c for a kind
int inputs_of(const NetSpec &s, int kind, int c, long long extra) {
  return s.deg[kind] + ((kind == NET_GRPKJ && c < extra) ? 1 : 0);
}

// One slot per thread: G <- G decay(receptor) + the weights of the synapses
// whose presynaptic bit is set in the arriving ring slot (null: none) or whose
// mossy fibre fires in interval k, added in synapse order.
struct Deliver {
  int nslot;
  const int *row, *pre;
  const double *w;
  const int8_t *rec;
  double *G;
  double decay[2];
  double p_mf;
  uint32_t k0, k1;
};

__global__ void __launch_bounds__(256) k_net_deliver(Deliver D, const uint32_t *__restrict__ arrive, uint32_t k) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= D.nslot) return;
  double add = 0.0;
  for (int j = D.row[s]; j < D.row[s + 1]; j++) {
    const int q = D.pre[j];
    bool hit;
    if (q >= 0) {
      hit = arrive != nullptr && ((arrive[q >> 5] >> (q & 31)) & 1u);
    } else {   // mossy fibre ~q fires when rate x dt > u, u in (0, 1] (mfgr_update, syn.cu:90-92)
      const u32x4 r = philox4x32_10((uint32_t)~q, k, 0u, 0u, D.k0, D.k1);
      hit = D.p_mf > ((double)r.v[0] + 1.0) * (1.0 / 4294967296.0);
    }
    if (hit) add += D.w[j];
  }
  D.G[s] = D.G[s] * D.decay[D.rec[s]] + add;
}

// One thread per cell, 64 cells per block: bit c of the ring words is set when
// the spike compartment crossed thr upwards since the previous interval end;
// each wave writes its two words whole (so the slot needs no clearing).
__global__ void __launch_bounds__(64) k_net_mark(const double *__restrict__ vnew, int comp, int n, int spike_comp,
                                                 double thr, double *__restrict__ prev, uint32_t *__restrict__ bits) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  bool up = false;
  if (c < n) {
    const double x = vnew[(size_t)c * comp + spike_comp];
    up = x > thr && prev[c] <= thr;
    prev[c] = x;
  }
  const unsigned long long m = __ballot(up);
  const int w = c >> 5;
  if ((threadIdx.x & 31) == 0 && c < n) bits[w] = (uint32_t)(m >> (threadIdx.x & 32));
}

}  // namespace

std::string net_check(const NetSpec &s) {
  for (int p = 0; p < NP_N; p++) {
    if (s.n[p] < 0) return "negative cell count";
    if (s.n[p] > 0 && s.comp[p] <= 0) return "a population needs at least one compartment per cell";
    if (s.n[p] > 0) {
      if (s.spike_comp[p] < 0 || s.spike_comp[p] >= s.comp[p]) return "spike compartment out of range";
      if (s.tgt_lo[p] < 0 || s.tgt_hi[p] <= s.tgt_lo[p] || s.tgt_hi[p] > s.comp[p])
        return "target compartment range out of range";
    }
  }
  for (int k = 0; k < NET_NKIND; k++) {
    if (s.deg[k] < 0) return "negative in-degree";
    const int pre = net_pre_pop(k);
    if (s.deg[k] > 0 && s.n[net_post_pop(k)] > 0 && pre >= 0 && s.n[pre] == 0)
      return "an in-degree needs presynaptic cells";
  }
  const long long extra = extra_inputs(s);
  if (extra < 0) return "total is smaller than the in-degrees give";
  if (extra > s.n[NP_PKJ]) return "total exceeds the in-degrees by more than one input per Purkinje cell";
  if (extra > 0 && s.n[NP_GR] == 0) return "extra Purkinje inputs need granule cells";
  long long size = 0;
  for (int k = 0; k < NET_NKIND; k++) size += (long long)s.deg[k] * s.n[net_post_pop(k)];
  size += extra;
  if (size >= (1LL << 30)) return "too many synapses for int32 indices";
  if (!(s.w_min_nS >= 0.0) || !(s.w_max_nS >= s.w_min_nS)) return "weights need 0 <= w_min <= w_max";
  if (!(s.tau_ms[0] > 0.0) || !(s.tau_ms[1] > 0.0)) return "time constants must be > 0";
  if (!(s.mf_rate_hz >= 0.0)) return "the mossy-fibre rate must be >= 0";
  if (!(s.interval_ms > 0.0)) return "the interval must be > 0";
  const double d = s.delay_ms / s.interval_ms;
  if (!(d >= 1.0) || std::fabs(d - std::nearbyint(d)) > 1e-9 || d > 4096.0)
    return "the delay must be a whole number (1 to 4096) of intervals";
  return "";
}

long long net_size(const NetSpec &s) {
  if (!net_check(s).empty()) return -1;
  long long size = extra_inputs(s);
  for (int k = 0; k < NET_NKIND; k++) size += (long long)s.deg[k] * s.n[net_post_pop(k)];
  return size;
}

void net_generate(const NetSpec &s, NetList &out) {
  const std::string err = net_check(s);
  if (!err.empty()) throw Error(err, -2);
  out = NetList();
  const long long extra = extra_inputs(s);
  const size_t size = (size_t)net_size(s);
  out.kind.reserve(size);
  out.pre.reserve(size);
  out.post.reserve(size);
  out.w_nS.reserve(size);
  const uint32_t k0 = (uint32_t)s.seed ^ KEY_NET, k1 = (uint32_t)(s.seed >> 32);
  int fibre = 0;
  for (int kind = 0; kind < NET_NKIND; kind++) {
    const int post = net_post_pop(kind), pre = net_pre_pop(kind);
    const uint32_t span = (uint32_t)(s.tgt_hi[post] - s.tgt_lo[post]);
    for (int c = 0; c < s.n[post]; c++) {
      const int m = inputs_of(s, kind, c, extra);
      for (int q = 0; q < m; q++) {
        const u32x4 r = philox4x32_10((uint32_t)c, (uint32_t)q, (uint32_t)kind, 0u, k0, k1);
        const int src = pre < 0 ? fibre++ : (int)(((uint64_t)r.v[0] * (uint32_t)s.n[pre]) >> 32);
        const double w = s.w_min_nS + (s.w_max_nS - s.w_min_nS) * u01_53(r.v[1], r.v[2]);
        const int cp = s.tgt_lo[post] + (int)(((uint64_t)r.v[3] * span) >> 32);
        out.kind.push_back((int8_t)kind);
        out.pre.push_back(src);
        out.post.push_back(c * s.comp[post] + cp);
        out.w_nS.push_back(w);
        out.count[kind]++;
      }
    }
  }
}

Network::Network(Engine *gr, Engine *go, Engine *pkj, const NetSpec &s, int device) : spec_(s), device_(device) {
  eng_[NP_GR] = gr;
  eng_[NP_GO] = go;
  eng_[NP_PKJ] = pkj;
  for (int p = 0; p < NP_N; p++) {
    if (!eng_[p]) throw Error("the network needs all three populations", -2);
    if (!eng_[p]->device_vnew() || !eng_[p]->stream())
      throw Error("a population's engine exposes no device voltage or stream", -5);
    spec_.n[p] = eng_[p]->ncell();
    spec_.comp[p] = eng_[p]->nc() / std::max(1, eng_[p]->ncell());
  }
  const std::string err = net_check(spec_);
  if (!err.empty()) throw Error(err, -2);
  delay_ = (int)std::nearbyint(spec_.delay_ms / spec_.interval_ms);
  NetList L;
  net_generate(spec_, L);
  for (int k = 0; k < NET_NKIND; k++) list_count_[k] = L.count[k];
  // ring bit of a presynaptic cell: word-aligned regions per population
  for (int p = 0, w = 0; p < NP_N; p++) {
    word_off_[p] = w;
    w += (spec_.n[p] + 31) / 32;
    words_ = w;
  }
  nring_ = 2 * (delay_ + 1);
  try {
    CB_HIP(hipSetDevice(device_));
    // slots per postsynaptic population: (compartment, receptor) in
    // compartment order, receptors AMPA then GABA_A; a slot's synapses in
    // generation order (a counting sort, stable)
    for (int p = 0; p < NP_N; p++) {
      Post &P = post_[p];
      const int nc = spec_.n[p] * spec_.comp[p];
      std::vector<int> key_count((size_t)nc * 2 + 1, 0);
      std::vector<size_t> mine;
      for (size_t j = 0; j < L.kind.size(); j++)
        if (net_post_pop(L.kind[j]) == p) {
          mine.push_back(j);
          key_count[(size_t)L.post[j] * 2 + net_receptor(L.kind[j]) + 1]++;
        }
      std::vector<int> slot_of_key((size_t)nc * 2, -1), row(1, 0);
      for (size_t key = 0; key < (size_t)nc * 2; key++)
        if (key_count[key + 1] > 0) {
          slot_of_key[key] = (int)P.slot_comp.size();
          P.slot_comp.push_back((int)(key / 2));
          P.slot_rec.push_back((int8_t)(key % 2));
          row.push_back(row.back() + key_count[key + 1]);
        }
      P.nslot = (int)P.slot_comp.size();
      std::vector<int> fill(row.begin(), row.end() - 1), pre(mine.size());
      std::vector<double> w(mine.size());
      for (size_t j : mine) {
        const int sl = slot_of_key[(size_t)L.post[j] * 2 + net_receptor(L.kind[j])];
        const int at = fill[sl]++;
        const int src_pop = net_pre_pop(L.kind[j]);
        pre[at] = src_pop < 0 ? ~L.pre[j] : word_off_[src_pop] * 32 + L.pre[j];
        w[at] = L.w_nS[j] * 1e-6;   // nS -> mS
      }
      if (P.nslot == 0) continue;
      P.row = dcopy(row);
      P.pre = dcopy(pre);
      P.w = dcopy(w);
      P.rec = dcopy(P.slot_rec);
      P.G = dnew<double>(P.nslot);
      CB_HIP(hipMemset(P.G, 0, sizeof(double) * P.nslot));
      // the engine's incoming-synapse view: per compartment its slots
      // (receptor order), slot s of receptor r as entry (r << 30) | s of val[r]
      std::vector<int> in_row(nc + 1, 0), in_slot(P.nslot);
      for (int sl = 0; sl < P.nslot; sl++) in_row[P.slot_comp[sl] + 1]++;
      for (int i = 0; i < nc; i++) in_row[i + 1] += in_row[i];
      for (int sl = 0; sl < P.nslot; sl++) in_slot[sl] = ((int)P.slot_rec[sl] << 30) | sl;
      P.in_row = dcopy(in_row);
      P.in_slot = dcopy(in_slot);
    }
    ring_ = dnew<uint32_t>((size_t)nring_ * words_);
    CB_HIP(hipMemset(ring_, 0, sizeof(uint32_t) * (size_t)nring_ * words_));
    for (int p = 0; p < NP_N; p++) {
      prev_[p] = dnew<double>(spec_.n[p]);
      std::vector<double> h(spec_.n[p], 100.0);   // no crossing at the first interval end
      if (!h.empty()) CB_HIP(hipMemcpy(prev_[p], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    }
    for (int p = 0; p < NP_N; p++) {
      SynIn in;
      if (post_[p].nslot > 0) {
        in.row = post_[p].in_row;
        in.slot = post_[p].in_slot;
        in.val[0] = in.val[1] = post_[p].G;
        in.E[0] = spec_.e_mV[0];
        in.E[1] = spec_.e_mV[1];
      }
      eng_[p]->sync();
      eng_[p]->set_synapses(in);
    }
  } catch (...) {
    for (int p = 0; p < NP_N; p++) {
      try {
        eng_[p]->set_synapses(SynIn());
      } catch (...) {
      }
    }
    free_all();
    throw;
  }
}

void Network::free_all() {
  for (Post &P : post_)
    for (void *q : {(void *)P.row, (void *)P.pre, (void *)P.w, (void *)P.rec, (void *)P.G, (void *)P.in_row,
                    (void *)P.in_slot})
      if (q) (void)hipFree(q);
  for (double *q : prev_)
    if (q) (void)hipFree(q);
  if (ring_) (void)hipFree(ring_);
  ring_ = nullptr;
}

Network::~Network() {
  for (int p = 0; p < NP_N; p++) {
    try {
      eng_[p]->sync();
      eng_[p]->set_synapses(SynIn());
    } catch (...) {
    }
  }
  free_all();
}

// interval k of population p on its own stream
void Network::interval(int p, long k) {
  Engine *E = eng_[p];
  hipStream_t st = E->stream();
  const Post &P = post_[p];
  if (P.nslot > 0) {
    const long j = k - 1 - delay_;   // spikes of interval j (marked at its end) arrive now
    const uint32_t *arrive = j >= 0 ? ring_ + (size_t)(j % nring_) * words_ : nullptr;
    Deliver D{P.nslot,
              P.row,
              P.pre,
              P.w,
              P.rec,
              P.G,
              {std::exp(-spec_.interval_ms / spec_.tau_ms[0]), std::exp(-spec_.interval_ms / spec_.tau_ms[1])},
              spec_.mf_rate_hz * spec_.interval_ms * 1e-3,
              (uint32_t)spec_.seed ^ KEY_MF,
              (uint32_t)(spec_.seed >> 32)};
    k_net_deliver<<<(unsigned)((P.nslot + 255) / 256), 256, 0, st>>>(D, arrive, (uint32_t)k);
    CB_HIP(hipGetLastError());
  }
  // one output interval of the population's solver: RKC one macro step
  // (set_current at its start), CN / BE their sub-steps with set_current at
  // each (main.cu:193-200)
  E->run(1, k * spec_.interval_ms, -1, nullptr);
  const int n = spec_.n[p];
  if (n > 0) {
    k_net_mark<<<(unsigned)((n + 63) / 64), 64, 0, st>>>(E->device_vnew(), spec_.comp[p], n, spec_.spike_comp[p],
                                                         spec_.thr_mV[p], prev_[p],
                                                         ring_ + (size_t)(k % nring_) * words_ + word_off_[p]);
    CB_HIP(hipGetLastError());
  }
}

namespace {
double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

void Network::times(double *pop3, double *exchange, double *total) const {
  for (int p = 0; p < NP_N; p++) pop3[p] = ms_pop_[p];
  *exchange = ms_exchange_;
  *total = ms_total_;
}

void Network::run(int nint) {
  if (nint < 0) throw Error("negative interval count", -2);
  const auto t_run = std::chrono::steady_clock::now();
  std::vector<uint32_t> ring((size_t)nring_ * words_);
  while (nint > 0) {
    // delay + 1 intervals read only the ring slots of earlier chunks and
    // write slots that no reader of this chunk uses (the ring holds two
    // chunks), so the populations run them without exchanging
    const int len = std::min(nint, delay_ + 1);
    const long k0 = k_;
    std::exception_ptr fail[NP_N];
    double ms[NP_N] = {0.0, 0.0, 0.0};
    auto work = [&](int p) {
      const auto t0 = std::chrono::steady_clock::now();
      try {
        CB_HIP(hipSetDevice(device_));
        for (int i = 0; i < len; i++) interval(p, k0 + i);
        CB_HIP(hipStreamSynchronize(eng_[p]->stream()));
      } catch (...) {
        fail[p] = std::current_exception();
      }
      ms[p] = ms_since(t0);
    };
    std::thread th[NP_N - 1];
    for (int p = 1; p < NP_N; p++) th[p - 1] = std::thread(work, p);
    work(0);
    for (auto &t : th) t.join();
    for (int p = 0; p < NP_N; p++) {
      if (fail[p]) std::rethrow_exception(fail[p]);
      ms_pop_[p] += ms[p];
    }
    const auto t_ex = std::chrono::steady_clock::now();
    CB_HIP(hipMemcpy(ring.data(), ring_, sizeof(uint32_t) * ring.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < len; i++) {
      const long k = k0 + i;
      const uint32_t *slot = ring.data() + (size_t)(k % nring_) * words_;
      for (int p = 0; p < NP_N; p++)
        for (int c = 0; c < spec_.n[p]; c++)
          if ((slot[word_off_[p] + (c >> 5)] >> (c & 31)) & 1u)
            spikes_.push_back({(k + 1) * spec_.interval_ms, {p, c}});
    }
    ms_exchange_ += ms_since(t_ex);
    k_ += len;
    nint -= len;
  }
  ms_total_ += ms_since(t_run);
}

int Network::take_spikes(int cap, double *t, int *pop, int *cell) {
  const int k = (int)std::min<size_t>(spikes_.size(), (size_t)std::max(0, cap));
  for (int i = 0; i < k; i++) {
    t[i] = spikes_[i].first;
    pop[i] = spikes_[i].second.first;
    cell[i] = spikes_[i].second.second;
  }
  spikes_.erase(spikes_.begin(), spikes_.begin() + k);
  return k;
}

void Network::conductance(int pop, int receptor, double *out) const {
  if (pop < 0 || pop >= NP_N || receptor < 0 || receptor > 1) throw Error("no such population or receptor", -2);
  const Post &P = post_[pop];
  const int nc = spec_.n[pop] * spec_.comp[pop];
  std::fill(out, out + nc, 0.0);
  if (P.nslot == 0) return;
  std::vector<double> g(P.nslot);
  CB_HIP(hipSetDevice(device_));
  CB_HIP(hipMemcpy(g.data(), P.G, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
  for (int s = 0; s < P.nslot; s++)
    if (P.slot_rec[s] == receptor) out[P.slot_comp[s]] = g[s];
}

}  // namespace cbrkc
