// Engine interface behind the C-ABI population handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "morph.hpp"
#include "synapse.hpp"

#define CB_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      throw cbrkc::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                         __FILE__ + ":" + std::to_string(__LINE__));                      \
    }                                                                                     \
  } while (0)

namespace cbrkc {

struct Error {
  std::string msg;
  int code = -1;
  explicit Error(std::string m, int c = -1) : msg(std::move(m)), code(c) {}
};

// RKC status beyond the reference's idid values (1, 2, 3, 4, 6): the state
// of a cell became NaN or infinite
constexpr int NONFINITE = 7;

// field ids shared with the C-ABI (cb_pop_read)
enum Field {
  F_VNEW = 0,     // interpolated V at the end of the last macro step (vec[v_new])
  F_V = 1,        // V at the last ion update (elem[v])
  F_CA = 2,
  F_REV = 3,
  F_B = 4,
  F_IEXT = 5,
  F_Y = 6,        // RKC vec[y]
  F_YN = 7,       // RKC vec[yn_r]
  F_FN = 8,       // RKC vec[fn]
  F_CA2 = 9,      // Golgi cells: the second (CaLVA) Ca pool
  F_CA_OLD = 10,  // Golgi cells: Ca at the previous ion update (KAHP scheme)
  F_VOLD = 11,    // CN solver: V of the previous step (vec[cn_v_old])
  F_CTL = 50,     // per-cell engines: the per-cell controller records (ncell x 32 doubles)
  F_GATE0 = 100,  // + gate index
};

struct Stats {
  long steps = 0, rejects = 0, stages = 0, rho_iters = 0, ion_updates = 0, macro_steps = 0;
};

struct StimSpec {
  int comp = 1599;        // compartment (local index) receiving the step current
  double t0 = 550.0, t1 = 1300.0;
  std::vector<double> amp;  // per cell, uA
};

// Poisson conductance drive on every compartment: rate_hz events per second,
// each adding g_event (mS) that decays with tau_ms; reversal e_syn (mV).
struct PoissonSpec {
  bool on = false;
  double rate_hz = 5.0, g_event = 1e-6, tau_ms = 2.0, e_syn = 0.0;
  unsigned long long seed = 0;
};

class Engine {
 public:
  virtual ~Engine() { free_raster(); }
  virtual void set_current(double t) = 0;      // xx_set_current
  virtual void advance() = 0;                  // one H_MAX macro step (xx_solve_by_rkc)
  virtual void read(int field, double *dst) = 0;  // nc values
  virtual Stats stats() const = 0;
  virtual int nc() const = 0;
  virtual int ncell() const = 0;
  virtual void controller(double *others14, double *hwork8) const { (void)others14; (void)hwork8; }
  virtual void sync() {}
  // picks the launch mapping by timing short trials from the current state
  // (the state is left unchanged); returns the chosen threads per cell
  virtual int autotune(int trial_steps, double t0) {
    (void)trial_steps;
    (void)t0;
    return 0;
  }
  // selects the mapping with this many threads per cell; false if none
  virtual bool select_mapping(int threads) { return threads == 0; }
  // a description of the launch mapping in use (and of the last autotune trial)
  virtual std::string mapping() const { return "one thread per compartment"; }
  // event-timed duration of the main kernel (ms) and its launch count
  virtual void kernel_time(double *ms, long *launches, bool reset) {
    *ms = 0.0;
    *launches = 0;
    (void)reset;
  }
  // many macro steps without host round trips; soma V and spikes are
  // buffered on the device.  Default: loop over advance().
  virtual void run(int nsteps, double t0, int rec_comp, double *rec_out) ;
  // run() without any device allocation: the per-call device memory (work
  // counter, recording staging) lives in the library's static per-device
  // scratch array.  Only the per-cell engines provide it.
  virtual void run_noalloc(int nsteps, double t0, int rec_comp, double *rec_out) {
    (void)nsteps;
    (void)t0;
    (void)rec_comp;
    (void)rec_out;
    throw Error("the allocation-free run is provided by the per-cell RKC engines", -5);
  }
  // couples the population's incoming synapses (in.on() false: none); the
  // device arrays stay owned by the caller and must outlive the coupling
  virtual void set_synapses(const SynIn &in) {
    if (in.on()) throw Error("synaptic input is not provided for this population's engine", -5);
  }
  // V at the last ion update / solve on the device (elem[v]), the
  // presynaptic voltage read by synapse updates, and the stream the engine
  // works on
  virtual const double *device_v() const { return nullptr; }
  virtual hipStream_t stream() const { return nullptr; }
  // the output voltage on the device (what read(F_VNEW) returns), or null
  virtual const double *device_vnew() const { return nullptr; }

  // Raster: with comp >= 0 every run detects, on the device after each
  // output interval, the upward crossings of thr by compartment comp of every
  // cell and appends (time at the interval's end, cell) to the spike list;
  // comp < 0 stops it.
  void set_raster(int comp, double thr);
  // moves up to cap spikes (time order, cells ascending within a time) out of
  // the list; returns how many
  int take_spikes(int cap, double *t_out, int *cell_out);

 protected:
  int raster_comp_ = -1;
  double raster_thr_ = 0.0;
  double *raster_prev_ = nullptr;     // device, one per cell, 100 mV at the start
  std::vector<std::pair<double, int>> spikes_;
  // the previous-value array, allocated at the first use
  double *raster_prev();
  // appends the spikes of nsteps intervals of bit words (words per interval)
  void append_spikes(const std::vector<uint32_t> &bits, int nsteps, int words, double t0, double interval);
  // the device recorder after output interval k of a run: recorded values
  // into rec (null: none), spike bits into bits (null: none)
  void record_interval(hipStream_t st, const double *v, double *rec, int rec_comp, uint32_t *bits);
  void free_raster();
};

// hipStreamCreate re-seeds the C library's rand() state.  The reference
// draws every population's initial voltages from one rand() sequence seeded
// once (main.cu:37) and creates no streams, so a stream created between two
// populations' draws would change the later draws.  This creates the stream
// with rand()'s state parked on a private buffer and restores the caller's
// state afterwards (glibc rand() and random() share that state).
void create_stream(hipStream_t *st);

// Tolerances and the operator-splitting interval of one population.
struct SolverSpec {
  double atol = 1e-6, rtol = 1e-6;
  double ion_interval = -1.0;   // <= 0: 2^log10(atol) (solve_rkc.cu:1265)
  bool refresh_fn = false;      // recompute fn = F(y_n) after each ion update (off in the reference)
  std::vector<double> rho_floor;  // per cell: lower bound applied to every spectral-radius estimate (empty: none)
  double hmin = -1.0;           // <= 0: the cell type's H_MIN (solve_rkc.cuh:24-27)
  double interval = 0.125;      // output / exchange interval, also the largest step (the reference's H_MAX)
  double hmin_or(double dflt) const { return hmin > 0.0 ? hmin : dflt; }
};

// RKC stage coefficients per stage count m = 2..mtab: row m holds mtab
// entries of 4 doubles; entry 0 = {mu~_1}, entry j = {mu_j, nu_j, mu~_j,
// a_{j-1}} for j = 1..m-1 (solve_rkc.cu:815-906).
std::vector<double> rkc_stage_table(int mtab);

// The batched engine over an uneven mix of cell types (granule, Golgi,
// Purkinje, inferior olive; index = the C-ABI's cell type): every cell on
// its own adaptive RKC step, all of them in one kernel per launch.
struct MixPopSpec {
  int n = 0;
  std::string morph;
  int comp = 0;                     // granule chain length (others: the format's)
  std::vector<double> v0, amp;      // per cell
  int stim_comp = 0, rec_comp = 0;
  double thr = 0.0;
};
struct MixSpec {
  MixPopSpec pop[4];                // CB_GR, CB_GO, CB_PKJ, CB_IO order (0, 1, 2, 3)
  std::string io_gj;                // "pre,post,g_nS" pairs, or empty
  double interval = 0.125;
  double pkj_atol = 1e-6, pkj_rtol = 1e-6;
  bool pkj_stable = true;           // ion interval 2^-8 ms, fn refresh, spectral-radius floor
};

class Mix {
 public:
  virtual ~Mix() {}
  // nint intervals; rec_out [nint][total cells], cells in type order;
  // spk_out [nint][words()] spike bits; either may be null
  virtual void run(int nint, double *rec_out, uint32_t *spk_out) = 0;
  virtual int ncell(int type) const = 0;
  virtual int total_cells() const = 0;
  virtual int words() const = 0;
  virtual int word_off(int type) const = 0;
  virtual int lanes(int type) const = 0;   // lanes (threads) per cell
  virtual Stats stats(int type) const = 0;
  virtual void autotune(int trial_intervals) = 0;
  virtual void kernel_time(double *ms, long *launches, bool reset) = 0;
  virtual void use_graph(bool on) = 0;
};
Mix *make_mix(const MixSpec &spec);

// synthetic inputs in the reference formats (synth.cpp)
bool synth_gr_chains(const std::string &path, const std::string &template_path, uint64_t seed, int ncell,
                     const std::vector<int> &types, std::string &err);
bool synth_go_trees(const std::string &path, const std::string &template_path, uint64_t seed, int ncell,
                    std::string &err);
bool synth_io_gj(const std::string &path, uint64_t seed, int ncell, int k, double g_min, double g_max,
                 std::string &err);

Engine *make_global_pkj(const PkjHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn);
Engine *make_global_io(const IoHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn);
Engine *make_global_go(const GoHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn);
Engine *make_global_gr(const GrHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn);
// the Crank-Nicolson engines (solve_cnm.cu), FP64, population-wide CG / BiCG;
// direct: an exact per-cell tree (Hines) solve instead, for populations whose
// cells are uncoupled trees (throws Error code -5 otherwise); dt > 0: the step
// [ms] in place of the reference's 0.025, a whole fraction of H_MAX
Engine *make_cn_pkj(const PkjHost &host, const StimSpec &stim, bool direct = false, double dt = 0.0);
Engine *make_cn_io(const IoHost &host, const StimSpec &stim, bool direct = false, double dt = 0.0);
Engine *make_cn_go(const GoHost &host, const StimSpec &stim, bool direct = false, double dt = 0.0);
Engine *make_cn_gr(const GrHost &host, const StimSpec &stim, bool direct = false, double dt = 0.0);
// the backward-Euler solver of solve_bem.cu on the same Krylov core
Engine *make_be_pkj(const PkjHost &h, const StimSpec &s, bool direct = false, double dt = 0.0);
Engine *make_be_io(const IoHost &h, const StimSpec &s, bool direct = false, double dt = 0.0);
Engine *make_be_go(const GoHost &h, const StimSpec &s, bool direct = false, double dt = 0.0);
Engine *make_be_gr(const GrHost &h, const StimSpec &s, bool direct = false, double dt = 0.0);
Engine *make_cell_pkj(const PkjHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn,
                      bool mixed);

}  // namespace cbrkc
