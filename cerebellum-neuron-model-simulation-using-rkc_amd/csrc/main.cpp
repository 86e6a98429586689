// Drop-in for the reference's `main` (main.cu:35-355), on the C-ABI:
//
//   ./main S_gr gx gy S_go ox oy S_pkj px py S_io ix iy [options]
//
// S_* is "BE", "CN" or "RKC", matched by prefix as main.cu:60-102 does; the
// cell count of a population is x * y and 0 disables it.  Output files are the
// reference's: xx_v_by_<S>.csv (t, then the recorded compartment of every
// cell, "%lf,"), xx_raster_by_<S>.csv ("%lf,%d" per upward threshold
// crossing), an empty debug_file.csv and one appended comptime.csv row.
//
// Options (all optional; the defaults are the reference's behaviour):
//   --ts MS          simulated time [ms] (default 1500, param.h:7)
//   --engine E       global (default: FP64, population-global controller, the
//                    reference's semantics), cell (FP64, one controller per
//                    cell) or mixed (FP32 stages/gates, FP64 control)
//   --stable         the safeguards for synthetic trees (ion interval 2^-8 ms,
//                    fn refresh, spectral-radius floor)
//   --chunk K        macro steps per device launch between host writes (64)
//   --network        update the synapses after every macro step and inject
//                    no step current (the reference's network_sim, main.cu:38,
//                    false there, which skips every xx_set_current,
//                    main.cu:192-196, 220-223, 248-251, 279-283); the synapse
//                    sets are created and coupled either way, as
//                    main.cu:139-145 does, so without it they stay silent
//   --tree           CN and BE populations solve each step exactly, per cell
//                    on its tree (Hines elimination), instead of the
//                    reference's 100-iteration CG / BiCG; cells coupled by gap
//                    junctions (Purkinje grids with Pkj_GJ.csv, Golgi grids)
//                    are refused
//
// This build provides all four populations with all three solvers.  Each
// population advances one H_MAX = 0.125 ms output interval per row, in
// sub-steps of its solver's DT (RKC: one step of H_MAX; CN and BE: 0.025 ms,
// main.cu:193-200).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cbrkc.h"

namespace {

const char *kSolverNames[3] = {"BE", "CN", "RKC"};
const char *kTypeNames[4] = {"gr", "go", "pkj", "io"};
const int kSolverArg[4] = {1, 4, 7, 10};
constexpr double H_MAX = 0.125;   // param.h:12

int solver_of(const char *s) {
  if (std::strncmp(s, "BE", 2) == 0) return CB_SOLVER_BE;
  if (std::strncmp(s, "CN", 2) == 0) return CB_SOLVER_CN;
  if (std::strncmp(s, "RKC", 3) == 0) return CB_SOLVER_RKC;
  return -1;
}

// a solver's step: one RKC call covers H_MAX, CN and BE step 0.025 ms (param.h:10-12)
double dt_of(int solver) { return solver == CB_SOLVER_RKC ? H_MAX : 0.025; }

// xx_initialize + xx_solve_initialize of a granule, Golgi or olive
// population; with --tree a CN / BE population takes the exact per-cell
// solve, with --network it has no step current
int init_pop(bool tree, bool network, int type, int solver, int nx, int ny, const char *name, cb_pop **out) {
  if ((tree && solver != CB_SOLVER_RKC) || network) {
    cb_pop_config cfg;
    const int rc = cb_config_default(&cfg, type, solver, nx, ny);
    if (rc != CB_OK) return rc;
    if (tree && solver != CB_SOLVER_RKC) cfg.engine = CB_ENGINE_TREE_FP64;
    if (network) cfg.stim_comp = CB_STIM_NONE;
    return cb_pop_create(&cfg, out);
  }
  if (type == CB_GR) return cb_gr_initialize(nx, ny, name, out);
  if (type == CB_GO) return cb_go_initialize(nx, ny, name, out);
  return cb_io_initialize(nx, ny, name, out);
}

int check(int rc, const char *what) {
  if (rc != CB_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, cb_last_error());
    std::exit(1);
  }
  return rc;
}

}  // namespace

int main(int argc, char *argv[]) {
  std::printf("argc = %d\n", argc);
  for (int i = 0; i < argc; i++) std::printf("argv[%d] = %s\n", i, argv[i]);
  if (argc < 13) {
    std::printf("usage: %s S_gr gx gy S_go ox oy S_pkj px py S_io ix iy [--ts MS] [--engine global|cell|mixed] "
                "[--stable] [--chunk K] [--network] [--tree]\n", argv[0]);
    return 0;
  }
  double ts = 1500.0;
  int engine = CB_ENGINE_GLOBAL_FP64, chunk = 64;
  bool stable = false, network = false, tree = false;
  for (int i = 13; i < argc; i++) {
    std::string o = argv[i];
    if (o == "--ts" && i + 1 < argc) {
      ts = std::atof(argv[++i]);
    } else if (o == "--engine" && i + 1 < argc) {
      std::string e = argv[++i];
      engine = e == "cell" ? CB_ENGINE_CELL_FP64 : e == "mixed" ? CB_ENGINE_CELL_MIXED : CB_ENGINE_GLOBAL_FP64;
    } else if (o == "--stable") {
      stable = true;
    } else if (o == "--chunk" && i + 1 < argc) {
      chunk = std::atoi(argv[++i]);
      if (chunk < 1) chunk = 1;
    } else if (o == "--network") {
      network = true;
    } else if (o == "--tree") {
      tree = true;
    } else {
      std::printf("unknown option %s\n", o.c_str());
      return 1;
    }
  }

  int solver[4], nx[4], ny[4];
  std::printf("GR  solver -> %s\nGO  solver -> %s\nPKJ solver -> %s\nIO  solver -> %s\n\n", argv[1], argv[4],
              argv[7], argv[10]);
  std::FILE *fv[4] = {nullptr, nullptr, nullptr, nullptr}, *fr[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < 4; k++) {
    solver[k] = solver_of(argv[kSolverArg[k]]);
    if (solver[k] < 0) {
      std::printf("Error!! -> set solver name in argv [ %d ]\n", kSolverArg[k]);
      return 0;
    }
    const std::string s = kSolverNames[solver[k]];
    fv[k] = std::fopen((std::string(kTypeNames[k]) + "_v_by_" + s + ".csv").c_str(), "w");
    fr[k] = std::fopen((std::string(kTypeNames[k]) + "_raster_by_" + s + ".csv").c_str(), "w");
    nx[k] = std::atoi(argv[kSolverArg[k] + 1]);
    ny[k] = std::atoi(argv[kSolverArg[k] + 2]);
  }
  std::FILE *dbg = std::fopen("debug_file.csv", "w");
  if (dbg) std::fclose(dbg);

  // the reference seeds libc rand() with 0 before any population draws its
  // initial voltages (main.cu:37), and initialises GR, GO, PKJ, IO in that
  // order (main.cu:104-137)
  std::srand(0);
  const int ngr = nx[CB_GR] * ny[CB_GR];
  cb_pop *grp = nullptr;
  if (ngr > 0) check(init_pop(tree, network, CB_GR, solver[CB_GR], nx[CB_GR], ny[CB_GR], argv[1], &grp), "gr_initialize");
  const int ngo = nx[CB_GO] * ny[CB_GO];
  cb_pop *gop = nullptr;
  if (ngo > 0) check(init_pop(tree, network, CB_GO, solver[CB_GO], nx[CB_GO], ny[CB_GO], argv[4], &gop), "go_initialize");
  const int n = nx[CB_PKJ] * ny[CB_PKJ];
  cb_pop *pkj = nullptr;
  int comp = 0;
  if (n > 0) {
    cb_pop_config cfg;
    check(cb_config_default(&cfg, CB_PKJ, solver[CB_PKJ], nx[CB_PKJ], ny[CB_PKJ]), "config");
    if (solver[CB_PKJ] == CB_SOLVER_RKC) cfg.engine = engine;
    if (tree && solver[CB_PKJ] != CB_SOLVER_RKC) cfg.engine = CB_ENGINE_TREE_FP64;
    if (network) cfg.stim_comp = CB_STIM_NONE;
    if (stable && solver[CB_PKJ] == CB_SOLVER_RKC) {
      cfg.ion_interval_ms = std::ldexp(1.0, -8);
      cfg.refresh_fn = 1;
      cfg.rho_floor = 1;
    }
    check(cb_pop_create(&cfg, &pkj), "pkj_initialize");
    int nn = 0;
    check(cb_pop_size(pkj, &nn, &comp), "pkj size");
    comp /= nn;
  }
  // the inferior olive draws its initial voltages after the Purkinje cells,
  // as io_initialize follows pkj_initialize in main.cu:125-137
  const int nio = nx[CB_IO] * ny[CB_IO];
  cb_pop *iop = nullptr;
  if (nio > 0) check(init_pop(tree, network, CB_IO, solver[CB_IO], nx[CB_IO], ny[CB_IO], argv[10], &iop), "io_initialize");
  std::printf("\n");

  // create synapse (main.cu:139-145): each Poisson / jittered set draws its
  // seed from rand() when it is not empty, in this order, as the reference
  // seeds its cuRAND states
  cb_syn *mfgr = nullptr, *gogr = nullptr, *grgo = nullptr, *grpkj = nullptr, *mlipkj = nullptr;
  {
    const uint64_t mf_seed = ngr > 0 ? (uint64_t)std::rand() : 0;
    check(cb_mfgr_create(grp, mf_seed, &mfgr), "mfgr_create");
    check(cb_gogr_create(gop, grp, &gogr), "gogr_create");
    check(cb_grgo_create(grp, gop, &grgo), "grgo_create");
    const uint64_t pk_seed = (ngr > 0 && n > 0) ? (uint64_t)std::rand() : 0;
    check(cb_grpkj_create(grp, pkj, "./Pkj_location_info.csv", pk_seed, &grpkj), "grpkj_create");
    const uint64_t mli_seed = (ngr > 0 && n > 0) ? (uint64_t)std::rand() : 0;
    check(cb_mlipkj_create(pkj, grp, mli_seed, &mlipkj), "mlipkj_create");
    int nsyn[5] = {0, 0, 0, 0, 0};
    const cb_syn *sets[5] = {mfgr, gogr, grgo, grpkj, mlipkj};
    const char *names[5] = {"mfgr", "gogr", "grgo", "grpkj", "mlipkj"};
    for (int k = 0; k < 5; k++) {
      check(cb_syn_size(sets[k], &nsyn[k]), "syn_size");
      std::printf("# of %s = %d\n", names[k], nsyn[k]);
    }
    // mfgr_create / mlipkj_create open these files, which stay empty
    // (mf_output_file and mli_output_file are not called, main.cu:215)
    if (nsyn[0] > 0) {
      std::FILE *f = std::fopen("MF_RASTER.csv", "w");
      if (f) std::fclose(f);
    }
    if (nsyn[4] > 0) {
      std::FILE *f = std::fopen("MLI_RASTER.csv", "w");
      if (f) std::fclose(f);
    }
  }
  if (network) chunk = 1;   // the synapses are updated between macro steps

  // the rasters of xx_output_file are found on the device (cb_pop_set_raster):
  // granule compartment 77 above +10 mV (gr.cu:268-276), Golgi 30 above +10
  // (go.cu:284-292), Purkinje 1599 above -10 (pkj.cu:403-409), olive 2 above
  // +10 (io.cu:276-284)
  if (grp) check(cb_pop_set_raster(grp, 77, 10.0), "gr raster");
  if (gop) check(cb_pop_set_raster(gop, 30, 10.0), "go raster");
  if (pkj) check(cb_pop_set_raster(pkj, 1599, -10.0), "pkj raster");
  if (iop) check(cb_pop_set_raster(iop, 2, 10.0), "io raster");

  // one population's chunk of k output intervals: the recorded values come
  // back once per chunk (per = values per cell and row), the raster from the
  // device spike list; rows and raster lines in the reference's formats
  auto chunk_out = [&](cb_pop *p, const char *what, int k, double t, int ncell, int per, int rec_comp,
                       int nan_index, std::FILE *fv_, std::FILE *fr_, std::vector<double> &rec) -> bool {
    rec.assign((size_t)k * ncell * per, 0.0);
    check(cb_pop_run(p, k, t, rec_comp, rec.data()), what);
    std::vector<double> st(256);
    std::vector<int> sc(256);
    std::vector<std::pair<double, int>> spikes;
    for (;;) {
      int got = 0;
      check(cb_pop_spikes(p, (int)st.size(), st.data(), sc.data(), &got), "spikes");
      for (int q = 0; q < got; q++) spikes.push_back({st[q], sc[q]});
      if (got < (int)st.size()) break;
    }
    size_t next = 0;
    for (int s = 0; s < k; s++) {
      const double ts_ = t + (s + 1) * H_MAX;
      const double *row = rec.data() + (size_t)s * ncell * per;
      if (std::isnan(row[nan_index])) return false;
      std::fprintf(fv_, "%lf,", ts_);
      for (size_t c = 0; c < (size_t)ncell * per; c++) std::fprintf(fv_, "%lf,", row[c]);
      std::fprintf(fv_, "\n");
      while (next < spikes.size() && spikes[next].first <= ts_) {
        std::fprintf(fr_, "%lf,%d\n", spikes[next].first, spikes[next].second);
        next++;
      }
    }
    return true;
  };

  // the macro-step loop of main.cu:182-311 for the populations this build
  // provides, in the reference's order (GR, GO, PKJ, IO); each population
  // advances a chunk of macro steps per call
  std::vector<double> rec;
  const auto start = std::chrono::steady_clock::now();
  auto half = start;
  const long total = (long)std::floor((ts - 0.0001) / H_MAX) + 1;   // while (t < TS - 0.0001)
  long done = 0;
  double t = 0.0;
  while (done < total) {
    const int k = (int)std::min<long>(chunk, total - done);
    // gr_output_file (gr.cu:256-277): V of compartment 16 per cell
    if (ngr > 0 && !chunk_out(grp, "gr_solve_update_v", k, t, ngr, 1, 16, 0, fv[CB_GR], fr[CB_GR], rec)) {
      std::printf("isnan in gr, time=%f\n", t);
      return 1;
    }
    // go_output_file (go.cu:270-293): V of compartment 30 per cell
    if (ngo > 0 && !chunk_out(gop, "go_solve_update_v", k, t, ngo, 1, 30, 0, fv[CB_GO], fr[CB_GO], rec)) {
      std::printf("isnan in go, time=%f\n", t);
      return 1;
    }
    // pkj_output_file (pkj.cu:370-412): V of compartment 1599 per cell
    if (n > 0 && !chunk_out(pkj, "pkj_solve_update_v", k, t, n, 1, 1599, 0, fv[CB_PKJ], fr[CB_PKJ], rec)) {
      std::printf("isnan in pkj, time=%f\n", t);
      return 1;
    }
    // io_output_file (io.cu:259-285): all three compartments of every cell;
    // the reference tests compartment 1 of cell 0 for NaN
    if (nio > 0 && !chunk_out(iop, "io_solve_update_v", k, t, nio, 3, CB_REC_ALL, 1, fv[CB_IO], fr[CB_IO], rec)) {
      std::printf("isnan,time=%f\n", t);
      return 1;
    }
    // Update synapses (main.cu:302-306), after every population's macro step
    if (network) {
      check(cb_gr_synapse_update(t, mfgr, gogr), "gr_synapse_update");
      check(cb_go_synapse_update(t, grgo), "go_synapse_update");
      check(cb_pkj_synapse_update(t, grpkj, mlipkj), "pkj_synapse_update");
    }
    for (int s = 0; s < k; s++) {
      if ((done + s) % 400 == 0) std::printf("time:%f\n", (done + s) * H_MAX);
      if (t + (s + 1) * H_MAX >= 500.0 && 500.0 + H_MAX > t + (s + 1) * H_MAX)
        half = std::chrono::steady_clock::now();
    }
    done += k;
    t = done * H_MAX;
  }
  const auto end = std::chrono::steady_clock::now();
  const double secs = std::chrono::duration<double>(end - start).count();
  std::printf("time:%f \ncomputational time = %.2f \n", t, secs);
  for (int k = 0; k < 4; k++) {
    if (fv[k]) std::fclose(fv[k]);
    if (fr[k]) std::fclose(fr[k]);
  }

  std::FILE *tp = std::fopen("comptime.csv", "a");
  if (tp) {
    const int sim = (int)ts;
    std::fprintf(tp, "Gr,%d,%s,Go,%d,%s,Pkj,%d,%s,Io,%d,%s,", ngr, argv[1], ngo, argv[4], n,
                 argv[7], nio, argv[10]);
    if (ts <= 500.0)
      std::fprintf(tp, "CompTime,%lf,SimTime,%d\n", secs, sim);
    else
      std::fprintf(tp, "AllCompTime,%lf,0.5s-endCompTime,%lf,SimTime,%d\n", secs,
                   std::chrono::duration<double>(end - half).count(), sim);
    std::fclose(tp);
  }
  // synapse_finalize (main.cu:325-331), before the populations
  for (cb_syn *x : {mfgr, gogr, grgo, grpkj, mlipkj}) check(cb_syn_finalize(x), "syn_finalize");
  if (grp) check(cb_gr_finalize(grp), "gr_finalize");
  if (gop) check(cb_go_finalize(gop), "go_finalize");
  if (pkj) check(cb_pkj_finalize(pkj), "pkj_finalize");
  if (iop) check(cb_io_finalize(iop), "io_finalize");
  return 0;
}
