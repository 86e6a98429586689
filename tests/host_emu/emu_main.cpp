// Host build of the batched engine's small-cell path, for the CPU tests: the
// device source of small_cells and k_io_gj and the host builders of the
// MixEngine (common_small, build_small, build_io_gj), taken verbatim from
// engine_cell.hip by build.sh, run with one host thread per lane.  A cell's G
// lanes meet at every shuffle and wave barrier (a barrier per cell group, as
// the lanes of one cell move together on the device), and every array is an
// exact-size heap block, so AddressSanitizer sees any access past an array.
//
// usage: emu CONFIG, CONFIG lines:
//   POP type morph n v0[0..n) amp[0..n)     (type 0 granule, 1 Golgi, 3 olive)
//   GJ path                                 (olive junction list)
//   NINT k                                  (intervals of 0.125 ms)
//   LANES g_gr g_go 0 g_io                  (0: the engine's default)
//   OUT path                                (recording [k][cells] doubles,
//                                            then spike bits [k][words])
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "go_model.hpp"
#include "gr_model.hpp"
#include "io_model.hpp"
#include "ion_state.hpp"
#include "morph.hpp"

thread_local emu_dim3 threadIdx;
thread_local emu_dim3 blockIdx;

namespace {

// sense-counting barrier of n threads that yields while it waits
struct Barrier {
  int n = 0;
  std::atomic<int> waiting{0};
  std::atomic<long> gen{0};
  void wait() {
    const long g = gen.load(std::memory_order_acquire);
    if (waiting.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
      waiting.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
    } else {
      while (gen.load(std::memory_order_acquire) == g) std::this_thread::yield();
    }
  }
};

// the lanes of one cell: a barrier and the exchange slots of a shuffle
struct Group {
  Barrier bar;
  double xs[64];
};
thread_local Group *t_group = nullptr;
thread_local int t_lane = 0;

}  // namespace

void emu_wave_barrier() {
  if (t_group && t_group->bar.n > 1) t_group->bar.wait();
}

double emu_shfl_xor(double x, int off) {
  const int G = t_group ? t_group->bar.n : 1;
  const int base = t_lane - t_lane % G, partner = t_lane ^ off;
  if (partner < base || partner >= base + G) {
    std::fprintf(stderr, "shuffle from lane %d to lane %d outside its cell's %d lanes\n", t_lane, partner, G);
    std::abort();
  }
  if (G == 1) return x;
  t_group->xs[t_lane] = x;
  t_group->bar.wait();
  const double y = t_group->xs[partner];
  t_group->bar.wait();
  return y;
}

namespace cbrkc {
namespace {
#include "ctl.inc"
#include "small_body.inc"
#include "io_gj.inc"
}  // namespace

// the MixEngine's small-population builders over host memory
struct HostMix {
  static constexpr int T_GR = 0, T_GO = 1, T_PKJ = 2, T_IO = 3;
  struct Small {
    std::vector<void *> mut;
    std::vector<size_t> mut_bytes;
  };
  MixSpec spec_;
  int n_[4] = {0, 0, 0, 0};
  SmallPop pop_[4] = {};
  Small small_[4];
  const int *io_row_ = nullptr, *io_col_ = nullptr;
  const double *io_g_ = nullptr;
  bool io_gj_on_ = false;
  std::vector<void *> bufs_;

  ~HostMix() {
    for (void *p : bufs_) std::free(p);
  }
  template <typename U>
  U *put(const std::vector<U> &h) {
    U *d = (U *)std::malloc(sizeof(U) * h.size());
    if (h.size()) std::memcpy((void *)d, h.data(), sizeof(U) * h.size());
    bufs_.push_back((void *)d);
    return d;
  }
  template <typename U>
  U *mut(const std::vector<U> &h, int t) {
    U *d = put(h);
    small_[t].mut.push_back((void *)d);
    small_[t].mut_bytes.push_back(sizeof(U) * h.size());
    return d;
  }
#include "nint.inc"
#include "mix_host.inc"
};

// one wavefront on one work item (population P, first cell c0)
template <bool INIT>
void run_item(const SmallPop &P, int c0, int nsteps) {
  std::vector<double> lds(2 * SMALL_LDS, std::nan(""));
  const int G = P.G;
  std::vector<std::unique_ptr<Group>> groups(64 / G);
  for (auto &g : groups) {
    g.reset(new Group);
    g->bar.n = G;
  }
  std::vector<std::thread> th;
  for (int lane = 0; lane < 64; lane++)
    th.emplace_back([&, lane]() {
      threadIdx = {(unsigned)lane, 0, 0};
      t_lane = lane;
      t_group = groups[lane / G].get();
      small_cells<INIT>(P, c0, nsteps, lds.data());
    });
  for (auto &t : th) t.join();
}

}  // namespace cbrkc

using namespace cbrkc;

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: emu CONFIG\n");
    return 2;
  }
  std::ifstream in(argv[1]);
  std::string line, out_path;
  MixSpec S;   // the C-ABI's defaults (cb_mix_config_default)
  S.pop[0].comp = 4;
  S.pop[0].stim_comp = 1;
  S.pop[0].rec_comp = 3;
  S.pop[0].thr = 10.0;
  S.pop[1].stim_comp = 30;
  S.pop[1].rec_comp = 30;
  S.pop[1].thr = 10.0;
  S.pop[3].stim_comp = 0;
  S.pop[3].rec_comp = 2;
  S.pop[3].thr = 10.0;
  int nint = 0;
  int lanes[4] = {0, 0, 0, 0};
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "POP") {
      int t, n;
      std::string path;
      ss >> t >> path >> n;
      S.pop[t].n = n;
      S.pop[t].morph = path;
      S.pop[t].v0.resize(n);
      S.pop[t].amp.resize(n);
      for (int c = 0; c < n; c++) ss >> S.pop[t].v0[c];
      for (int c = 0; c < n; c++) ss >> S.pop[t].amp[c];
    } else if (key == "GJ") {
      ss >> S.io_gj;
    } else if (key == "NINT") {
      ss >> nint;
    } else if (key == "OUT") {
      ss >> out_path;
    } else if (key == "LANES") {
      for (int t = 0; t < 4; t++) ss >> lanes[t];
    }
  }
  try {
    HostMix M;
    M.spec_ = S;
    for (int t = 0; t < 4; t++) M.n_[t] = S.pop[t].n;
    if (M.n_[2]) throw Error("the host build covers the small cell types only");
    for (int t : {0, 1, 3})
      if (M.n_[t]) M.build_small(t);
    if (!S.io_gj.empty() && M.n_[3]) M.build_io_gj();
    for (int t : {0, 1, 3})
      if (lanes[t]) M.pop_[t].G = lanes[t];
    // the engine's item order: Golgi, granule, olive; 64 / G cells an item
    auto items = [&](auto fn) {
      for (int t : {1, 0, 3}) {
        if (!M.n_[t]) continue;
        const int W = 64 / M.pop_[t].G;
        for (int c0 = 0; c0 < M.n_[t]; c0 += W) fn(t, c0);
      }
    };
    items([&](int t, int c0) { run_item<true>(M.pop_[t], c0, 0); });
    int tot = 0, words = 0, off[4], woff[4];
    for (int t = 0; t < 4; t++) {
      off[t] = tot;
      woff[t] = words;
      tot += M.n_[t];
      words += (M.n_[t] + 31) / 32;
    }
    std::vector<double> rec((size_t)nint * tot, 0.0);
    std::vector<uint32_t> spk((size_t)nint * words, 0u);
    // granule and Golgi cells: one launch over all intervals (the engine's
    // launch without olive coupling); the recording of interval s at s * n
    for (int t : {0, 1}) {
      if (!M.n_[t]) continue;
      std::vector<double> r((size_t)nint * M.n_[t], std::nan(""));
      M.pop_[t].rec = r.data();
      M.pop_[t].spk = spk.data() + woff[t];
      M.pop_[t].spk_words = words;
      const int W = 64 / M.pop_[t].G;
      for (int c0 = 0; c0 < M.n_[t]; c0 += W) run_item<false>(M.pop_[t], c0, nint);
      for (int l = 0; l < nint; l++)
        for (int c = 0; c < M.n_[t]; c++) rec[(size_t)l * tot + off[t] + c] = r[(size_t)l * M.n_[t] + c];
      for (int c = 0; c < M.n_[t]; c++) {
        const int s = (int)M.pop_[t].ctl[(size_t)c * NCTL + C_STATUS];
        if (s) {
          std::fprintf(stderr, "type %d cell %d: RKC status %d\n", t, c, s);
          return 3;
        }
      }
    }
    // olive cells: per interval the junction currents, then one launch of
    // one interval (the engine's sequence with coupling)
    for (int l = 0; l < nint && M.n_[3]; l++) {
      if (M.io_gj_on_) {
        const int n = M.n_[3];
        for (unsigned b = 0; b < (unsigned)((n + 255) / 256); b++)
          for (unsigned x = 0; x < 256; x++) {
            blockIdx = {b, 0, 0};
            threadIdx = {x, 0, 0};
            k_io_gj(M.io_row_, M.io_col_, M.io_g_, M.pop_[3].vnew, io::COMP, (double *)M.pop_[3].gj, n);
          }
      }
      std::vector<double> r(M.n_[3], std::nan(""));
      M.pop_[3].rec = r.data();
      M.pop_[3].spk = spk.data() + (size_t)l * words + woff[3];
      M.pop_[3].spk_words = words;
      const int W = 64 / M.pop_[3].G;
      for (int c0 = 0; c0 < M.n_[3]; c0 += W) run_item<false>(M.pop_[3], c0, 1);
      for (int c = 0; c < M.n_[3]; c++) rec[(size_t)l * tot + off[3] + c] = r[c];
      for (int t : {3})
        for (int c = 0; c < M.n_[t]; c++) {
          const int s = (int)M.pop_[t].ctl[(size_t)c * NCTL + C_STATUS];
          if (s) {
            std::fprintf(stderr, "type %d cell %d: RKC status %d in interval %d\n", t, c, s, l);
            return 3;
          }
        }
    }
    std::FILE *f = std::fopen(out_path.c_str(), "wb");
    if (!f) throw Error("cannot write " + out_path);
    std::fwrite(rec.data(), sizeof(double), rec.size(), f);
    std::fwrite(spk.data(), sizeof(uint32_t), spk.size(), f);
    std::fclose(f);
    std::printf("ok %d intervals, %d cells, %d words\n", nint, tot, words);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
