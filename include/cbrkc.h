/*
 * cbrkc — MI355X-native RKC time stepping for cerebellar multi-compartment
 * neuron populations.  C-ABI boundary: plain pointers, sizes and status
 * codes, no device or framework types.
 *
 * A population handle (cb_pop) stands for the reference's pair
 * (neuron_t, neuron_solve_t) of one cell type (param.h:27-46): it owns the
 * compartment state, the cable matrix and the solver state.  Unlike the
 * reference, which keeps solver state in function-local statics
 * (solve_rkc.cu:1258-1265), every handle is independent.
 *
 * All functions return CB_OK (0) or a negative CB_E_* code; the message of
 * the last failure on the calling thread is available from cb_last_error().
 */
#ifndef CBRKC_H
#define CBRKC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CB_OK 0
#define CB_E_ARG (-1)      /* bad argument or configuration */
#define CB_E_IO (-2)       /* missing or malformed input file */
#define CB_E_HIP (-3)      /* HIP runtime failure */
#define CB_E_SOLVER (-4)   /* RKC failure: a zero error weight (the reference's idid 3) or a NaN /
                              infinite state (the reference's main stops on a NaN voltage); idid 4
                              (h below hmin) and 6 (power method not converged) fall through to an
                              accepted step, as in the reference */
#define CB_E_UNSUP (-5)    /* combination not provided */

/* cell types, param.h:15 */
enum { CB_GR = 0, CB_GO = 1, CB_PKJ = 2, CB_IO = 3 };
/* solvers, param.h:14 (argv names "BE", "CN", "RKC") */
enum { CB_SOLVER_BE = 0, CB_SOLVER_CN = 1, CB_SOLVER_RKC = 2 };
/* engines for the RKC solver */
enum {
  CB_ENGINE_GLOBAL_FP64 = 0,  /* one controller for the whole population, FP64 (reference semantics) */
  CB_ENGINE_CELL_FP64 = 1,    /* one controller per cell, persistent launch, FP64 */
  CB_ENGINE_CELL_MIXED = 2,   /* one controller per cell, FP32 stages/gates, compensated V, FP64 control */
  CB_ENGINE_TREE_FP64 = 3     /* CN and BE only: the exact per-cell tree (Hines) solve in place of the
                                 reference's 100-iteration CG / BiCG (cg_cusparse_crs, solve_cnm.cu:20-98);
                                 cells must be uncoupled trees of at most 1820 compartments (CB_E_UNSUP
                                 otherwise); no counterpart in the reference, which has only the Krylov
                                 solves */
};
/* fields readable with cb_pop_read (nc doubles each) */
enum {
  CB_F_VNEW = 0,   /* V at the end of the last macro step (vec[v_new]) */
  CB_F_V = 1,      /* V at the last ion update (elem[v]) */
  CB_F_CA = 2,
  CB_F_REV = 3,    /* Ca reversal potential */
  CB_F_B = 4,      /* frozen current */
  CB_F_IEXT = 5,
  CB_F_Y = 6,
  CB_F_YN = 7,
  CB_F_FN = 8,
  CB_F_CA2 = 9,    /* Golgi cells: second (CaLVA) Ca pool (go->ca2) */
  CB_F_CA_OLD = 10, /* Golgi cells: Ca at the previous ion update (go->ca_old) */
  CB_F_VOLD = 11,  /* CN solver: V of the previous step (vec[cn_v_old]) */
  CB_F_GATE0 = 100 /* + gate index in the cell type's gate order */
};

/* cb_pop_config.stim_comp: no injected current at all */
#define CB_STIM_NONE (-2)

typedef struct cb_pop cb_pop;

typedef struct cb_pop_config {
  int cell_type;            /* CB_PKJ, ... */
  int solver;               /* CB_SOLVER_* */
  int engine;               /* CB_ENGINE_* (CN and BE: CB_ENGINE_GLOBAL_FP64, the reference's Krylov
                               solve, or CB_ENGINE_TREE_FP64) */
  int nx, ny;               /* cell grid; nx * ny cells */
  const char *morph_path;   /* morphology file; NULL: "./Pkj.txt" */
  const char *gj_path;      /* gap-junction pair list of a Purkinje grid (nx * ny > 1): NULL reads
                               "./Pkj_GJ.csv", and a missing file is an error (pkj_solve.cu:86-87);
                               "" builds the grid without gap junctions */
  double atol, rtol;        /* <= 0: the reference's value for the cell type */
  const double *v0;         /* per-cell initial V [mV]; NULL: -68 + 2 (U - 0.5) from libc rand() */
  int stim_comp;            /* local compartment of the step current; -1: reference protocol;
                               CB_STIM_NONE: no injected current, as the reference runs with
                               network_sim on (main.cu:192-196, 220-223, 248-251, 279-283) */
  double stim_t0, stim_t1;  /* [ms] window of the step current */
  const double *stim_amp;   /* per-cell amplitude [uA]; NULL: reference amplitude */
  int poisson_on;           /* dendritic Poisson conductance drive on every compartment */
  double poisson_rate_hz, poisson_g_mS, poisson_tau_ms, poisson_e_mV;
  uint64_t poisson_seed;
  double ion_interval_ms;   /* accepted time between ion updates; <= 0: 2^log10(atol) (solve_rkc.cu:1265) */
  int refresh_fn;           /* 1: recompute fn = F(y_n) after each ion update (the reference keeps
                               the stale value, solve_rkc.cu:1229-1250); 0: reference behaviour */
  int rho_floor;            /* 1: spectral-radius estimates are raised to at least 1.2 x the exact
                               cable spectral radius (see cb_pkj_cable_spectral_radius); 0: reference */
  int device;               /* HIP device ordinal */
  double step_ms;           /* CN and BE: the step [ms]; <= 0: the reference's 0.025 (param.h:10-11);
                               otherwise 0.125 / step_ms must be a whole number from 1 to 4096 (no
                               counterpart in the reference, whose step is a compile-time constant) */
  double hmin_ms;           /* RKC: the smallest step [ms]; <= 0: the reference's H_MIN of the cell type
                               (2^-10, granule 2^-12, solve_rkc.cuh:24-27); at most H_MAX = 0.125 */
  int comp;                 /* granule cells: compartments per cell of a chain file (one block of comp
                               rows per cell, as cb_synth_gr_chains writes); <= 0: the format's 578
                               (gr.cuh:11).  Must be <= 0 for the other types, whose files fix it.
                               A chain needs an explicit stimulus compartment or CB_STIM_NONE. */
} cb_pop_config;

/* Fills a config with the reference defaults for one cell type and solver. */
int cb_config_default(cb_pop_config *cfg, int cell_type, int solver, int nx, int ny);

/* Creates a population and initialises its solver: xx_initialize +
 * xx_solve_initialize + RKC_vec_initialize (pkj.cu:240-306,
 * pkj_solve.cu:286-366, solve_rkc.cu:756-791); for CN the halved matrix and
 * xx_cnm_vec_initialize (solve_cnm.cu:420-465), for BE the unhalved matrix
 * (pkj_solve.cu:238-245).  CB_E_ARG for an unknown solver, CB_E_UNSUP for
 * CN / BE with a per-cell engine or the Poisson drive. */
int cb_pop_create(const cb_pop_config *cfg, cb_pop **out);

/* Releases every resource of the population: xx_finalize + xx_solve_finalize
 * (pkj.cu:308-331, pkj_solve.cu:368-393). */
int cb_pop_destroy(cb_pop *p);

/* Sets the external current for time t [ms]: xx_set_current (pkj.cu:333-344). */
int cb_pop_set_current(cb_pop *p, double t);

/* Advances one step of the population's solver: xx_solve_update_v
 * (pkj_solve.cu:395-409): RKC one macro step of H_MAX = 0.125 ms
 * (solve_rkc.cu:1252-1299), CN one 0.025 ms step (xx_solve_by_cnm,
 * solve_cnm.cu), BE one 0.025 ms step (xx_solve_by_bem, solve_bem.cu). */
int cb_pop_advance(cb_pop *p);

/* nsteps output intervals of H_MAX from t0 (RKC: set_current(t0 + k H_MAX);
 * advance; CN / BE: the five 0.025 ms sub-steps of main.cu:193-200, each with
 * set_current(t + l_dt)), recording compartment rec_comp of every cell after
 * each interval into rec_out[k * ncell + c] (rec_out may be NULL), or, with
 * rec_comp = CB_REC_ALL, every compartment into rec_out[k * nc + i] (the
 * olive's output row, io.cu:259-285; not the per-cell engines).  The main
 * loop of main.cu:182-311 for one population: the recording (and the raster,
 * see cb_pop_set_raster) stay on the device and are copied back once per
 * call instead of all of V after every step (main.cu:263-273). */
#define CB_REC_ALL (-1)
int cb_pop_run(cb_pop *p, int nsteps, double t0, int rec_comp, double *rec_out);

/* Raster on the device: from now on every cb_pop_run finds, after each output
 * interval, the upward crossings of thr [mV] by compartment comp of every cell
 * (previous value 100 mV at the start, main.cu:162-169) and appends (time of
 * the interval's end, cell) to the population's spike list; comp < 0 stops
 * it.  The per-cell engines detect on the recorded compartment, so a run of
 * theirs that records must record comp.  Replaces the host-side raster of
 * xx_output_file (pkj.cu:403-409, gr.cu:268-276, go.cu:284-292,
 * io.cu:276-284). */
int cb_pop_set_raster(cb_pop *p, int comp, double thr);

/* Moves up to cap spikes out of the population's list, in time order and
 * cells ascending within a time: t_out[i], cell_out[i]; *n_out = how many. */
int cb_pop_spikes(cb_pop *p, int cap, double *t_out, int *cell_out, int *n_out);

/* cb_pop_run without any device allocation or free, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The per-call device
 * memory (the work counter and the recording before its copy to rec_out)
 * lives in a file-scope device array of 8 MiB per device, so nsteps x ncell
 * must stay within 1048544 recorded doubles (CB_E_ARG beyond).  All calls on
 * one device share that array: calls that overlap on one device (several
 * host threads, or populations on different streams used concurrently)
 * race.  Per-cell RKC engines only (CB_E_UNSUP otherwise); results equal
 * cb_pop_run's. */
int cb_pop_run_noalloc(cb_pop *p, int nsteps, double t0, int rec_comp, double *rec_out);

/* Copies nc values of a state field to host memory. */
int cb_pop_read(cb_pop *p, int field, double *dst);

/* Population sizes. */
int cb_pop_size(const cb_pop *p, int *ncell, int *nc);

/* Counters: accepted steps, rejected steps, RHS stage sweeps, power-method
 * iterations, ion updates, macro steps (6 values). */
int cb_pop_stats(const cb_pop *p, int64_t *out6);

/* Device time of the engine's main kernel, measured with HIP events on the
 * stream it is launched on: accumulated milliseconds and launch count since
 * the last reset (reset != 0 clears them after reading).  No counterpart in
 * the reference, which only times the whole run with clock() (main.cu:178-189). */
int cb_pop_kernel_time(cb_pop *p, double *ms, int64_t *launches, int reset);

/* Per-cell engines: times every launch mapping (threads per cell x
 * compartments per thread) on a trial of trial_steps macro steps from the
 * current state, keeps the fastest and restores the state; *threads gets the
 * chosen threads per cell (0 for engines with a single mapping).  When
 * *threads > 0 on entry, that mapping is selected without a trial
 * (CB_E_ARG if the engine has no such mapping).  No counterpart in the
 * reference. */
int cb_pop_autotune(cb_pop *p, int trial_steps, int *threads);

/* Tree-solve engine (CB_ENGINE_TREE_FP64): when every cell is at most 64
 * unbranched runs, both the solve by runs and the level-by-level solve apply;
 * cb_pop_autotune times both on trial_steps solves of a copy of the current
 * system and keeps the faster (*threads = 64, one wavefront per cell).
 *
 * A text description of the launch mapping in use and of the last autotune
 * trial, NUL-terminated in buf (CB_E_ARG when len is too short).  No
 * counterpart in the reference. */
int cb_pop_mapping(const cb_pop *p, char *buf, int len);

/* Global-controller scalars: h_others[14] and h_work[8] (param.h:22,
 * solve_rkc.cu:453-476).  CB_E_UNSUP for per-cell engines. */
int cb_pop_controller(const cb_pop *p, double *others14, double *hwork8);

/* Writes one seeded synthetic Purkinje tree per cell, in the Pkj.txt format,
 * to path: soma sphere of radius soma_r_um, binary branching, segment
 * lengths U(lmin_um, lmax_um), diameters tapering with depth from d0_um to
 * d1_um; cell c uses seed + c and its diameters are scaled by dscale[c]
 * (dscale NULL: 1). */
int cb_synth_pkj_morph(const char *path, uint64_t seed, int ncell, double soma_r_um, double lmin_um,
                       double lmax_um, double d0_um, double d1_um, const double *dscale);

/* Host only: spectral radius of each cell's cable operator C^-1 L (the RKC
 * Jacobian with the ionic current frozen) for the ncell cells of a Pkj.txt
 * morphology (one block replicated, or one block per cell) into rho[ncell].
 * The reference estimates it with a loosely converged power method
 * (solve_rkc.cu:562-754), which can fall short of the true value on
 * synthetic trees; 1.2 x this value is the floor applied when the
 * configuration sets rho_floor. */
int cb_pkj_cable_spectral_radius(const char *morph_path, int ncell, double *rho);

/* Philox4x32-10 block (Salmon et al., SC'11), the generator behind every
 * seeded synthetic input. */
void cb_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Message of the last failure on this thread ("" if none). */
const char *cb_last_error(void);

/* ---- entry points named after the reference's per-type interface ---- */

/* pkj_initialize + pkj_solve_initialize (pkj.cu:240, pkj_solve.cu:286):
 * solver "BE", "CN" or "RKC" matched by prefix as main.cu:60-102 does;
 * reads ./Pkj.txt and ./Pkj_GJ.csv. */
int cb_pkj_initialize(int nx, int ny, const char *solver, cb_pop **out);
/* pkj_solve_update_v (pkj_solve.cu:395-409) without network synapses. */
int cb_pkj_solve_update_v(cb_pop *p);
/* pkj_set_current (pkj.cu:333-344). */
int cb_pkj_set_current(cb_pop *p, double t);
/* pkj_output_file (pkj.cu:370-412): appends "t,V_1599(cell0),...,\n" to
 * v_csv and "t,cell\n" for every upward crossing of -10 mV to raster_csv;
 * memv is the population V (nc values), raster_prev one value per cell. */
int cb_pkj_output_file(const cb_pop *p, const double *memv, double *raster_prev, double t, const char *v_csv,
                       const char *raster_csv);
/* pkj_finalize + pkj_solve_finalize. */
int cb_pkj_finalize(cb_pop *p);

/* gr_initialize + gr_solve_initialize (gr.cu:151-218, gr_solve.cu:170-259):
 * reads ./Gr.txt; initial voltages from libc rand() as gr_ion.cu:642-644
 * draws them. */
int cb_gr_initialize(int nx, int ny, const char *solver, cb_pop **out);
/* gr_solve_update_v (gr_solve.cu:292-307) without mossy-fibre / Golgi synapses. */
int cb_gr_solve_update_v(cb_pop *p);
/* gr_set_current (gr.cu:245-254): 0.01 nA into compartment 16, [50, 100) ms. */
int cb_gr_set_current(cb_pop *p, double t);
/* gr_output_file (gr.cu:256-277): appends "t,V16 of every cell,\n" and a
 * "t,cell" row per upward crossing of +10 mV at compartment 77. */
int cb_gr_output_file(const cb_pop *p, const double *memv, double *raster_prev, double t, const char *v_csv,
                      const char *raster_csv);
/* gr_finalize + gr_solve_finalize. */
int cb_gr_finalize(cb_pop *p);

/* go_initialize + go_solve_initialize (go.cu:160-230, go_solve.cu:248-333):
 * reads ./Go_131.txt; the grid gap junctions between the cells' compartments 0
 * are part of the cable matrix (go_solve.cu:52-200); initial voltages from
 * libc rand() as go_ion.cu:597-600 draws them. */
int cb_go_initialize(int nx, int ny, const char *solver, cb_pop **out);
/* go_solve_update_v (go_solve.cu:361-374) without granule-cell synapses. */
int cb_go_solve_update_v(cb_pop *p);
/* go_set_current (go.cu:258-268): 0.2 nA into compartment 30, [800, 1100) ms. */
int cb_go_set_current(cb_pop *p, double t);
/* go_output_file (go.cu:270-293): appends "t,V30 of every cell,\n" and a
 * "t,cell" row per upward crossing of +10 mV. */
int cb_go_output_file(const cb_pop *p, const double *memv, double *raster_prev, double t, const char *v_csv,
                      const char *raster_csv);
/* go_finalize + go_solve_finalize. */
int cb_go_finalize(cb_pop *p);

/* io_initialize + io_solve_initialize + io_gap_create (io.cu:186-229,
 * io_solve.cu:192-262, gap.cu:136-147): reads ./IO.txt; initial voltages from
 * libc rand() as io_ion.cu:168-170 draws them. */
int cb_io_initialize(int nx, int ny, const char *solver, cb_pop **out);
/* io_solve_update_v (io_solve.cu:296-312) with the grid gap junctions. */
int cb_io_solve_update_v(cb_pop *p);
/* io_set_current (io.cu:247-257). */
int cb_io_set_current(cb_pop *p, double t);
/* io_output_file (io.cu:259-285): appends "t,V0,V1,V2 of cell 0,...,\n" to
 * v_csv and "t,cell\n" for every upward crossing of +10 mV by compartment 2
 * (the axon hillock) to raster_csv; memv holds nc values. */
int cb_io_output_file(const cb_pop *p, const double *memv, double *raster_prev, double t, const char *v_csv,
                      const char *raster_csv);
/* io_finalize + io_solve_finalize + io_gap_finalize. */
int cb_io_finalize(cb_pop *p);

/* ---- synapses between the populations (syn.cu) ----
 * A set is created between existing populations (a NULL population gives an
 * empty set) and couples itself into its postsynaptic population's solver
 * as the reference's xx_solve_update_v does with the synapse_t it is passed:
 * RKC granule cells b += val (y - E) at each ion update, Golgi / Purkinje
 * cells the macro step's val (y - E) from its start at each ion update
 * (Purkinje: x exp(-dt / 0.6)); CN gamma2 / omega2 += val / 2, val E / 2; BE
 * diagonal / b += val, val E (granule and Golgi cells only); the per-cell
 * Purkinje engines (CB_ENGINE_CELL_*) take the RKC Purkinje term and then
 * advance one macro step per call.  Finalize every set before its
 * populations.  The random draws (mossy-fibre and inhibitory
 * Poisson inputs, contact jitter) come from Philox4x32-10 keyed by seed; the
 * reference's cuRAND streams are not reproduced. */
typedef struct cb_syn cb_syn;

/* set fields for cb_syn_read: decaying traces (MF: AMPA, NMDA; GoGr: GABAA,
 * GABAB; GrGo: AMPA, NMDA1, NMDA2; GrPkj: AMPA; MliPkj: GABA), weight,
 * conductance val [mS], last presynaptic V (syn.cuh:63-68) */
enum { CB_SYN_TRACE0 = 0, CB_SYN_TRACE1 = 1, CB_SYN_TRACE2 = 2, CB_SYN_WEIGHT = 3, CB_SYN_VAL = 4,
       CB_SYN_OLD_V = 5 };
/* set kinds for cb_syn_build */
enum { CB_SYN_MFGR = 0, CB_SYN_GOGR = 1, CB_SYN_GRGO = 2, CB_SYN_GRPKJ = 3, CB_SYN_MLIPKJ = 4 };

/* mfgr_create (syn.cu:104-128): 4 mossy-fibre inputs per granule cell. */
int cb_mfgr_create(cb_pop *gr, uint64_t seed, cb_syn **out);
/* gogr_create (syn.cu:339-410): Golgi compartment 30 -> granule compartment 16. */
int cb_gogr_create(cb_pop *go, cb_pop *gr, cb_syn **out);
/* grgo_create (syn.cu:181-261): granule axon -> Golgi compartment 0. */
int cb_grgo_create(cb_pop *gr, cb_pop *go, cb_syn **out);
/* grpkj_create (syn.cu:519-644): reads the compartment x positions from
 * location_csv (NULL: "./Pkj_location_info.csv"), 25 synapses per pair. */
int cb_grpkj_create(cb_pop *gr, cb_pop *pkj, const char *location_csv, uint64_t seed, cb_syn **out);
/* mlipkj_create (syn.cu:690-738): one inert synapse per Purkinje cell. */
int cb_mlipkj_create(cb_pop *pkj, cb_pop *gr, uint64_t seed, cb_syn **out);
/* mfgr_finalize / gogr_finalize / ... (syn.cu:130-141 etc.): uncouples and frees. */
int cb_syn_finalize(cb_syn *s);
/* gr_synapse_update (syn.cu:821-855): mossy-fibre draws, Golgi spikes. */
int cb_gr_synapse_update(double t, cb_syn *mfgr, cb_syn *gogr);
/* go_synapse_update (syn.cu:857-885): granule spikes onto Golgi cells. */
int cb_go_synapse_update(double t, cb_syn *grgo);
/* pkj_synapse_update (syn.cu:887-917): granule spikes onto Purkinje cells; the
 * inhibitory set is not updated, as in the reference. */
int cb_pkj_synapse_update(double t, cb_syn *grpkj, cb_syn *mlipkj);
int cb_syn_size(const cb_syn *s, int *n);
/* pre / post global compartment of every synapse (syn.cuh:62 comp). */
int cb_syn_comp(const cb_syn *s, int *pre, int *post);
/* n values of one CB_SYN_* field. */
int cb_syn_read(const cb_syn *s, int field, double *dst);
/* The connection rules alone, on the host (no device): the set of the given
 * kind between an nx_a x ny_a grid and an nx_b x ny_b grid (MFGR a = GR;
 * GOGR a = GO, b = GR; GRGO a = GR, b = GO; GRPKJ a = GR, b = PKJ with x1600
 * the location file's x column; MLIPKJ a = PKJ, b = GR).  With pre, post and
 * weight NULL only *n is set; otherwise they receive *n entries. */
int cb_syn_build(int kind, int nx_a, int ny_a, int nx_b, int ny_b, const double *x1600, uint64_t seed, int *n,
                 int *pre, int *post, double *weight);

/* ---- fixed in-degree network between populations ----
 * The reference's network run (main.cu:182-311 with network_sim on: per
 * macro step every population advances, then the synapse sets update from
 * the presynaptic voltages, syn.cu:822-917) over a granule, a Golgi and a
 * Purkinje population, with a seeded fixed in-degree connectivity in place of
 * the reference's geometric rules (syn.cu:144-261, 292-410, 440-644):
 *   mossy fibre -> granule (AMPA), Golgi -> granule (GABA_A), granule ->
 *   Golgi (AMPA), granule -> Purkinje (AMPA);
 * weights U(w_min, w_max) nS; one decaying conductance per (compartment,
 * receptor), G <- G exp(-dt / tau) + the weights of the spikes arriving in
 * the interval (the reference's per-synapse traces summed per compartment,
 * syn.cu:87-111, 264-288, 413-437, 647-668); a spike (an upward crossing of
 * the spike compartment's threshold between interval ends) arrives delay_ms
 * later; mossy fibres fire with probability rate x dt per interval
 * (mfgr_update, syn.cu:87-102; Philox4x32-10 in place of cuRAND).  The
 * conductances couple into each population's engine as the reference's sets
 * do (solve_rkc.cu:200-224, 253-287, 318-367).  The interval is the macro
 * step, H_MAX = 0.125 ms. */
typedef struct cb_net cb_net;
typedef struct cb_net_config {
  uint64_t seed;
  int deg_mf_gr, deg_go_gr, deg_gr_go, deg_gr_pkj;  /* inputs per postsynaptic cell */
  int64_t total;            /* > 0: the first (total - the in-degrees' sum) Purkinje cells take one more
                               granule input each, so the network holds exactly total synapses */
  double w_min_nS, w_max_nS;
  double tau_ampa_ms, tau_gaba_ms;
  double e_ampa_mV, e_gaba_mV;
  double delay_ms;          /* a whole number (1 to 4096) of 0.125 ms intervals */
  double mf_rate_hz;
  int gr_target_comp;       /* the granule compartment of the mossy-fibre and Golgi inputs */
  int go_target_lo, go_target_hi;    /* [lo, hi): Golgi compartments of granule inputs (uniform) */
  int pkj_target_lo, pkj_target_hi;  /* [lo, hi): Purkinje compartments of granule inputs (uniform) */
  int gr_spike_comp, go_spike_comp, pkj_spike_comp;
  double gr_thr_mV, go_thr_mV, pkj_thr_mV;
} cb_net_config;

/* 4 mossy-fibre and 4 Golgi inputs per granule cell, 100 granule inputs per
 * Golgi cell, 1392 per Purkinje cell, total 780404 (2048 / 512 / 512 cells),
 * w U(0.5, 1.5) nS, AMPA 1.5 ms / 0 mV, GABA_A 6 ms / -82 mV (E_GOGR,
 * syn.cuh), delay 1 ms, mossy fibres 5 Hz; targets: granule compartment 0
 * (the chain's dendrite), Golgi 31..130 (the Go_131.txt dendrites), Purkinje
 * 0..1598; spikes: granule chain end 3 and Golgi 30 through 0 mV (the
 * reference's pre V > 0 && old_v < 0 test), Purkinje 1599 through -10 mV. */
int cb_net_config_default(cb_net_config *cfg);
/* The synapse list alone, on the host (no device), for populations of n_gr,
 * n_go, n_pkj cells of comp_gr, comp_go, comp_pkj compartments: in
 * generation order (kind CB_NET_MFGR..CB_NET_GRPKJ, then postsynaptic cell,
 * then input) the kind, presynaptic cell (mossy fibres: the fibre, one per
 * synapse), postsynaptic compartment within its population and weight [nS].
 * With every array NULL only *n is set; otherwise they receive *n <= cap
 * entries (CB_E_ARG when cap is smaller).  No counterpart in the reference. */
enum { CB_NET_MFGR = 0, CB_NET_GOGR = 1, CB_NET_GRGO = 2, CB_NET_GRPKJ = 3 };
int cb_net_generate(const cb_net_config *cfg, int n_gr, int n_go, int n_pkj, int comp_gr, int comp_go,
                    int comp_pkj, int64_t cap, int8_t *kind, int32_t *pre, int32_t *post, double *w_nS,
                    int64_t *n);
/* Builds the network over three populations of one device and couples it
 * into them: the mfgr / gogr / grgo / grpkj _create calls of main.cu:140-158
 * in one.  Each population keeps its own solver and engine (RKC on any
 * engine; CN or BE, global Krylov or tree solve) and couples the
 * conductances as that solver's reference path does (the BE Purkinje path
 * couples none, as in the reference).  The populations take no other
 * synapse sets meanwhile and are destroyed after the network. */
int cb_net_create(cb_pop *gr, cb_pop *go, cb_pop *pkj, const cb_net_config *cfg, cb_net **out);
/* nint intervals of 0.125 ms: per interval every population advances one
 * output interval of its solver (RKC: set_current, one macro step; CN / BE:
 * the sub-steps of main.cu:193-200) with the conductances of the interval,
 * and its spikes enter the ring; the three
 * populations run concurrently for up to delay + 1 intervals between
 * exchanges (the network's own delay bounds what an interval can depend on). */
int cb_net_run(cb_net *net, int nint);
/* moves up to cap raster entries out of the network: time [ms] at the end of
 * the interval, population (0 granule, 1 Golgi, 2 Purkinje) and cell, in
 * time order, populations in that order within a time, cells ascending */
int cb_net_spikes(cb_net *net, int cap, double *t_out, int *pop_out, int *cell_out, int *n_out);
/* the conductance [mS] of every compartment of one population (0 granule, 1
 * Golgi, 2 Purkinje) for one receptor (0 AMPA, 1 GABA_A); 0 where none */
int cb_net_conductance(const cb_net *net, int pop, int receptor, double *dst);
/* synapses per kind (4), slots per population (3), intervals run, delay in intervals */
int cb_net_stats(const cb_net *net, int64_t *out9);
/* host wall time [ms] of the runs so far: each population's own interval loop
 * (its engine's work and waits; the populations run concurrently), the
 * exchanges between chunks, and the runs in total.  No counterpart in the
 * reference. */
int cb_net_time(const cb_net *net, double *pop_ms3, double *exchange_ms, double *total_ms);
int cb_net_destroy(cb_net *net);

/* ---- batched engine over an uneven mix of cell types ----
 * Granule, Golgi, Purkinje and inferior-olive cells advanced together, each
 * cell on its own adaptive RKC step (the reference's algorithm applied to a
 * population of one cell, solve_rkc.cu:562-1362), all of them in one kernel
 * per launch from device work counters: a workgroup per Purkinje cell
 * (mixed precision, as CB_ENGINE_CELL_MIXED), G lanes per granule / Golgi /
 * olive cell (FP64; 64 / G cells per wavefront).  Olive cells exchange
 * gap-junction currents once per interval (each pair's current
 * g (0.8 exp(-0.01 dV^2) + 0.2) dV, the reference's junction shape of
 * gap.cu:149-160 with a per-pair conductance, frozen over the interval);
 * chunks of intervals run as a captured HIP graph.  Constant per-cell
 * currents, no synapses.  No counterpart in the reference, which advances
 * one population of one type per call (main.cu:182-311). */
typedef struct cb_mix cb_mix;
typedef struct cb_mix_pop {
  int n;                    /* cells (0: none of this type) */
  const char *morph_path;   /* the type's file format (Gr.txt, Go_131.txt, Pkj.txt, IO.txt): one block of
                               rows per cell, or one block for all */
  int comp;                 /* granule: compartments per cell (<= 578); ignored for the other types */
  const double *v0;         /* per-cell initial V [mV] (required) */
  const double *amp;        /* per-cell constant current into stim_comp [uA; olive uA/cm^2]; NULL: none */
  int stim_comp, rec_comp;  /* stimulated compartment; compartment recorded and tested for spikes */
  double thr;               /* spike threshold on rec_comp [mV] (upward crossing between intervals) */
} cb_mix_pop;
typedef struct cb_mix_config {
  cb_mix_pop pop[4];        /* indexed by CB_GR, CB_GO, CB_PKJ, CB_IO */
  const char *io_gj_path;   /* olive junctions, lines "pre,post,g_nS" (cells of the olive population);
                               NULL: uncoupled */
  double interval_ms;       /* exchange and output interval, also the largest step (0.125 by default) */
  double pkj_atol, pkj_rtol;/* Purkinje tolerances (the other types use the reference's TOL_xx) */
  int pkj_stable;           /* 1: the Purkinje safeguards (ion interval 2^-8 ms, fn refresh,
                               spectral-radius floor) */
  int device;
} cb_mix_config;

/* Reference-like defaults: interval 0.125 ms, Purkinje 1e-6 / 1e-6 with the
 * safeguards, the reference's stimulated and raster compartments and
 * thresholds per type (granule: chain end, +10 mV; Golgi: 30, +10 mV;
 * Purkinje: 1599, -10 mV; olive: stimulus on 0, raster on 2, +10 mV). */
int cb_mix_config_default(cb_mix_config *cfg);
/* Builds every population and runs RKC_vec_initialize for every cell. */
int cb_mix_create(const cb_mix_config *cfg, cb_mix **out);
/* nint intervals: rec_out[k * total + off + c] = V of rec_comp of cell c of
 * the type at offset off (types in the order GR, GO, PKJ, IO) after interval
 * k; spk_out[k * words + word_off[type] + c / 32] bit c % 32 = an upward
 * threshold crossing of that cell in interval k.  Either may be NULL. */
int cb_mix_run(cb_mix *m, int nint, double *rec_out, uint32_t *spk_out);
/* cells per type, each type's first spike word, words per interval, lanes
 * (threads) per cell of each type */
int cb_mix_layout(const cb_mix *m, int *ncell4, int *word_off4, int *words, int *lanes4);
/* counters of one type (as cb_pop_stats) */
int cb_mix_stats(const cb_mix *m, int type, int64_t *out6);
/* Picks the lanes per cell of each small type by timing trial intervals from
 * the current state (restored afterwards); lanes4 receives the mapping. */
int cb_mix_autotune(cb_mix *m, int trial_intervals, int *lanes4);
/* event-timed duration of the launches of the runs so far (ms) and their count */
int cb_mix_kernel_time(cb_mix *m, double *ms, int64_t *launches, int reset);
/* 0: enqueue every chunk launch by launch instead of replaying its graph */
int cb_mix_use_graph(cb_mix *m, int on);
int cb_mix_destroy(cb_mix *m);

/* Seeded synthetic inputs in the reference formats (host only):
 * granule chains of comp compartments (types[k] names the template row type of
 * compartment k; NULL: dendrite 3, soma 4, axon initial segment 6, axon 7),
 * Ra, Cm and densities from the template's first row of that type, diameter
 * and length from it times U(0.8, 1.25); Golgi trees with the template's
 * topology, types and densities, every diameter and length times
 * U(0.8, 1.25); olive gap junctions, every cell joined to k distinct random
 * partners with g ~ U(g_min, g_max) nS, lines "pre,post,g_nS" (pre < post). */
int cb_synth_gr_chains(const char *path, const char *template_path, uint64_t seed, int ncell, const int *types,
                       int comp);
int cb_synth_go_trees(const char *path, const char *template_path, uint64_t seed, int ncell);
int cb_synth_io_gj(const char *path, uint64_t seed, int ncell, int k, double g_min_nS, double g_max_nS);

#ifdef __cplusplus
}
#endif
#endif
