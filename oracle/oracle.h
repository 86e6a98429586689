/*
 * CPU FP64 oracle for the RKC time-stepping path of the cerebellar
 * multi-compartment simulator.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is a plain-C restatement of the reference
 * algorithm, written from the reference sources read as text, and it is used
 * only as the checker (tests, the smoke check and the CPU-baseline leg of the
 * benchmark).  Nothing in the product links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  The arithmetic
 * order of each expression follows the reference expression, with
 * floating-point contraction disabled (-ffp-contract=off), so that the HIP
 * FP64 path can be compared against it at the ulp level for short runs.
 *
 * Pinning status of each part is recorded in the header of the file that
 * implements it.
 */
#ifndef CB_ORACLE_H
#define CB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cell types, param.h:15 */
enum { OR_GR = 0, OR_GO = 1, OR_PKJ = 2, OR_IO = 3 };

/* RKC vector slots, param.h:21 */
enum { OR_Y = 0, OR_YN, OR_FN, OR_VT1, OR_VT2, OR_EDE, OR_VNEW, OR_NVEC };
/* scalar controller slots, param.h:22 */
enum { OR_MMAX = 0, OR_ABSH, OR_SPRAD, OR_NACCPT, OR_ATOL, OR_RTOL, OR_IDID,
       OR_GT, OR_GT_OLD, OR_LT, OR_TEND, OR_ERROLD, OR_HOLD, OR_NSTSIG,
       OR_NOTHERS };

#define OR_MAX_COND 16
#define OR_MAX_ION 24

/* synapse kinds (syn.cu), in the order the reference creates them */
enum { OR_SYN_MFGR = 0, OR_SYN_GOGR, OR_SYN_GRGO, OR_SYN_GRPKJ, OR_SYN_MLIPKJ };

struct or_pop;
/* one synapse set (the reference's synapse_t, syn.cuh:76-82): global pre and
   post compartments, the decaying traces, weight, conductance val and the
   last presynaptic voltage */
typedef struct or_syn {
  int kind, n;
  int *pre, *post;
  double *x[3];
  double *weight, *val, *old_v;
  double E;                      /* reversal potential [mV] */
  const struct or_pop *pre_pop;  /* spike source (GoGr: Golgi; GrGo, GrPkj: granule) */
  uint32_t key[2];               /* Philox key of the random draws */
  long step;                     /* updates done (Philox counter) */
} or_syn;

typedef struct or_pop {
  int type;           /* OR_PKJ, ... */
  int n, nx, ny;      /* cells */
  int comp;           /* compartments per cell */
  int nc;             /* n * comp */
  int ncond, nion;
  /* compartment state, structure of arrays (param.h:27-35 elem/ion/cond) */
  double *v, *ca, *cm, *area, *iext, *shell, *rev;
  double *ca2, *ca_old; /* GO: second Ca pool and the Ca of the previous ion update */
  /* CN path (or_cn.c): previous voltage, half conductance / reversal sums of
     the last step, the halved cable matrix (reference sign), the position of
     each row's diagonal and the transpose's structure */
  double *cn_vold, *cn_gamma1, *cn_omega1, *cn_val;
  int *cn_dpos, *cn_trow, *cn_tcol, *cn_tsrc;
  int cn_direct;      /* 1: exact tree-elimination solve in place of CG / BiCG */
  double cn_dt;       /* CN / BE step [ms]; 0: the reference's 0.025 (param.h:10-11) */
  int *compart;       /* compartment type */
  double *cond[OR_MAX_COND];
  double *ion[OR_MAX_ION];
  /* cable matrix in CSR (signed for RKC: off-diagonal +g, diagonal -sum g) */
  int nnz;
  int *row, *col;
  double *val;
  double *b;          /* frozen ionic + external + synaptic current */
  double *dammy;
  double *vec[OR_NVEC];
  double *syn_g;      /* optional per-compartment synaptic conductance [mS] */
  double syn_e;       /* its reversal potential [mV] */
  /* RKC controller (solve_rkc.cu:437-560) */
  double h_work[8];
  double others[OR_NOTHERS];
  int last, newspc, jacatt;
  double sum_dt, integ;
  double hmin, hmax;
  /* stimulus (pkj.cu:333-344 shape: one compartment per cell, a window) */
  int stim_comp;
  double stim_t0, stim_t1;
  double *stim_amp;   /* per cell */
  /* statistics */
  long n_steps, n_rejects, n_stages, n_rho_iters, n_ion_updates;
  int nthreads;
  /* global index of compartment 0 of this population: a one-cell oracle
     standing in for cell c of a larger population uses c * comp, so the
     index-keyed quirk (pkj_ion.cu:329) and the noise counters line up */
  long gid_offset;
  int refresh_fn;     /* 1: recompute fn = F(y_n) after every ion update */
  double rho_floor;   /* lower bound applied to every spectral-radius estimate (0: none) */
  /* incoming synapse sets (or_syn.c), in the reference's launch order, and
     the macro step's synaptic current of the Golgi / Purkinje RKC paths */
  int n_syn_in;
  or_syn *syn_in[2];
  double *syn_s;
} or_pop;

/* ---- deterministic sums (or_rkc.c) ---- */
double or_det_sum(const double *x, long n);

/* ---- generic RKC (or_rkc.c), solve_rkc.cu ---- */
void or_dydt(or_pop *p, const double *y, double *dy);
int  or_rkc_init(or_pop *p);                /* RKC_vec_initialize */
int  or_rkc_advance(or_pop *p);             /* xx_solve_by_rkc, one H_MAX macro step */
void or_pop_free(or_pop *p);

/* ---- Purkinje (or_pkj.c) ---- */
or_pop *or_pkj_create(const char *morph_file, const char *gj_file, int nx, int ny,
                      const double *v0_per_cell, double atol, double rtol,
                      int use_libc_rand);
void or_pkj_update_b(or_pop *p, const double *vv);
void or_pkj_update_ion(or_pop *p, const double *vnew, double dt);
double or_pkj_rate(int which, double v, double ca);

/* ---- inferior olive (or_io.c) ---- */
or_pop *or_io_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand);
void or_io_update_b(or_pop *p, const double *vv, int add_gap);
void or_io_update_ion(or_pop *p, const double *vnew, double dt);
double or_io_rate(int which, double v, int comp);
void or_io_neighbours(int id, int nx, int ny, int *out4);

/* ---- Golgi cells (or_go.c) ---- */
or_pop *or_go_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand);
void or_go_update_b(or_pop *p, const double *vv);
void or_go_update_ion(or_pop *p, const double *vnew, double dt);
double or_go_rate(int which, double v, double ca);
int or_go_kahp(double ca_old, double ca, double dt, double *st6);

/* ---- granule cells (or_gr.c) ---- */
or_pop *or_gr_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand);
or_pop *or_gr_create_comp(const char *param_file, int comp, int nx, int ny, const double *v0_per_cell, double atol,
                          double rtol, int use_libc_rand);
void or_gr_update_b(or_pop *p, const double *vv);
void or_gr_update_ion(or_pop *p, const double *vnew, double dt);
double or_gr_rate(int which, double v, double ca);
int or_gr_na_step(int type, double v_old, double v_new, double dt, double *x12);

/* ---- CN path (or_cn.c) and the models' CN pieces ---- */
int or_cn_init(or_pop *p);
int or_cn_step(or_pop *p);
int or_cn_advance(or_pop *p, double t);
int or_cn_solve_test(or_pop *p, const double *diag, const double *b, double *x, int sym);
int or_cn_set_direct(or_pop *p, int on);
int or_cn_set_dt(or_pop *p, double dt);
int or_cn_direct_test(or_pop *p, const double *diag, const double *b, double *x);
void or_pkj_cn_ion(or_pop *p, double dt);
void or_pkj_cn_ge(const or_pop *p, int id, double *gsum, double *esum);
void or_io_cn_ion(or_pop *p, double dt);
void or_io_cn_ge(const or_pop *p, int id, double *gsum, double *esum);
double or_io_gap_into(const or_pop *p, int cell);
void or_go_cn_ion(or_pop *p, double dt);
void or_go_cn_ge(const or_pop *p, int id, double *gsum, double *esum);
void or_gr_cn_ion(or_pop *p, double dt);
void or_gr_cn_ge(const or_pop *p, int id, double *gsum, double *esum);

/* backward-Euler path (or_cn.c and the cell-type files) */
int or_be_init(or_pop *p);
int or_be_step(or_pop *p);
int or_be_advance(or_pop *p, double t);
int or_be_assemble_test(or_pop *p, double *diag, double *b);
void or_pkj_be_ion(or_pop *p, double dt);
void or_pkj_be_ge(const or_pop *p, int id, double dt, double *gsum, double *esum);
void or_io_be_ion(or_pop *p, double dt);
void or_io_be_ge(const or_pop *p, int id, double dt, double *gsum, double *esum);
double or_io_gap_add(const or_pop *p, int cell, double base);
void or_go_be_ion(or_pop *p, double dt);
void or_go_be_ge(const or_pop *p, int id, double dt, double *gsum, double *esum);
int or_go_kahp_be(double ca, double dt, double *st6);
void or_gr_be_ion(or_pop *p, double dt);
void or_gr_be_ge(const or_pop *p, int id, double dt, double *gsum, double *esum);
int or_gr_na_be(int type, double v, double dt, double *x12);

/* ---- stimulus ---- */
void or_set_stim(or_pop *p, int comp, double t0, double t1, const double *amp_per_cell);
void or_set_current(or_pop *p, double t);

/* ---- accessors for tests ---- */
int  or_pop_nc(const or_pop *p);
void or_pop_set_integ(or_pop *p, double integ);
void or_pop_set_gid_offset(or_pop *p, long off);
void or_pop_set_refresh_fn(or_pop *p, int on);
void or_pop_set_rho_floor(or_pop *p, double floor_);
void or_pop_get(const or_pop *p, int what, double *dst);
void or_pop_set(or_pop *p, int what, const double *src);
void or_pop_controller(const or_pop *p, double *others14, double *hwork8, int *flags3, double *sumdt);
void or_pop_stats(const or_pop *p, long *out5);
void or_pop_set_threads(or_pop *p, int nthreads);
int  or_pop_nnz(const or_pop *p);
void or_pop_csr(const or_pop *p, int *row, int *col, double *val);
void or_pop_enable_syn(or_pop *p, double e_syn);
void or_pop_poisson_step(or_pop *p, uint32_t step, uint32_t k0, uint32_t k1, double decay, double prob,
                         double g_event);

/* ---- test hooks (or_rkc.c): the RKC core on a population given by its matrix ---- */
or_pop *or_generic_create(int nc, const int *row, const int *col, const double *val, const double *cm,
                          const double *b, const double *v0, double atol, double rtol);
void or_rkc_step_test(or_pop *p, double h, int m);
void or_rkclow_test(or_pop *p);
void or_rkcrho_test(or_pop *p);
void or_rkcint_test(or_pop *p, double t, double *dst);
void or_pop_set_other(or_pop *p, int i, double v);
void or_pop_set_hwork(or_pop *p, int i, double v);

/* ---- synapses (or_syn.c), syn.cu ---- */
or_syn *or_mfgr_create(int n_gr, uint64_t seed);
or_syn *or_grgo_create(const or_pop *gr, const or_pop *go);
or_syn *or_gogr_create(const or_pop *go, const or_pop *gr);
or_syn *or_grpkj_create(const or_pop *gr, const or_pop *pkj, const double *x1600, uint64_t seed);
or_syn *or_mlipkj_create(int n_pkj, uint64_t seed);
void or_syn_free(or_syn *s);
void or_syn_update(or_syn *s, double t);
void or_net_update(or_syn *mfgr, or_syn *gogr, or_syn *grgo, or_syn *grpkj, double t);
int or_pop_attach_syn(or_pop *p, or_syn *s);
void or_syn_add_b(or_pop *p, const double *vv);
void or_syn_macro_start(or_pop *p);
void or_syn_ion_b(or_pop *p, const double *vv, double sum_dt);
void or_syn_cn(const or_pop *p, double *gamma2, double *omega2);
void or_syn_be(const or_pop *p, double *diag, double *b);
int or_syn_n(const or_syn *s);
void or_syn_get(const or_syn *s, int field, double *dst);
void or_syn_set(or_syn *s, int field, const double *src);
void or_syn_comp(const or_syn *s, int *pre, int *post);

/* ---- Philox4x32-10 (or_philox.c), Salmon et al., SC'11 ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- fixed in-degree network (or_net.c) ----
   kinds 0 mossy fibre -> granule (AMPA), 1 Golgi -> granule (GABA_A),
   2 granule -> Golgi (AMPA), 3 granule -> Purkinje (AMPA); populations
   0 granule, 1 Golgi, 2 Purkinje */
typedef struct or_net_spec {
  uint64_t seed;
  int deg[4];
  long long total;
  double w_min_nS, w_max_nS;
  double tau_ampa_ms, tau_gaba_ms, e_ampa_mV, e_gaba_mV;
  int delay;            /* intervals */
  double mf_rate_hz, dt_ms;
  int n[3], comp[3], tgt_lo[3], tgt_hi[3], spike_comp[3];
  double thr_mV[3];
} or_net_spec;
typedef struct or_net or_net;
long long or_net_size(const or_net_spec *s);
long long or_net_generate(const or_net_spec *s, int8_t *kind, int32_t *pre, int32_t *post, double *w_nS);
/* the populations (n and comp taken from them) take the network's sets;
   free the network before them */
or_net *or_net_create(or_pop *gr, or_pop *go, or_pop *pkj, const or_net_spec *spec);
void or_net_deliver(or_net *N);
void or_net_mark(or_net *N);
int or_net_run(or_net *N, int nint, unsigned char *spikes_out);
void or_net_conductance(const or_net *N, int pop, int rec, double *out);
void or_net_deliver_given(or_net *N, const unsigned char *arrive);
void or_net_free(or_net *N);

#ifdef __cplusplus
}
#endif
#endif
