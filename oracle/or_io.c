/*
 * Oracle: inferior-olive cell model on the RKC path (TEST INFRASTRUCTURE,
 * see oracle.h).
 *
 * Restates:
 *   constants               io.cuh:11-58
 *   initialize_host_io      io.cu:3-101   (IO.txt, conductances by compartment type)
 *   set_host_io_solve       io_solve.cu:2-150 (asymmetric 3-compartment coupling, CSR)
 *   rate functions          io_ion.cu:3-54
 *   io_initialize_ion       io_ion.cu:163-194
 *   io_update_ion_RKC       io_ion.cu:90-127
 *   io_update_matrix        solve_rkc.cu:412-434
 *   io_gap_initialize       gap.cu:4-133  (4-neighbour grid on compartment 0)
 *   io_gap_update           gap.cu:149-160, add_io_gap_val solve_rkc.cu:400-410
 *   io_set_current          io.cu:247-257
 *   io_update_ion_2nd       io_ion.cu:53-89 (CN path)
 *   CN conductance sums     solve_cnm.cu:741-821, gap term :729-739
 *
 * Reference quirks kept: the grid row of a cell is id / ny (gap.cu:11), so
 * non-square grids get the reference's neighbour sets; the dendrite's Na
 * inactivation has h_inf = tau_h = 1 (io_ion.cu:17-21, 37-41 evaluate only
 * compartment types 1 and 2); the gap current uses the voltage after the ion
 * update (elem[v] = vnew, io_ion.cu:125), and the Ca current of the update
 * uses the voltage and gate before it.
 *
 * Gap currents are added to b in neighbour order 0..3; the reference adds
 * them with atomicAdd in an unspecified order (parity unpinned at the last
 * bit for compartments with several neighbours).
 *
 * Pinning: rates against their closed-form limits and the alpha/(alpha+beta)
 * identities, the CN update against the (1,1) Pade form, the Ca RK4 against
 * the exact exponential, the coupling matrix against the published
 * two-parameter soma/dendrite/axon split (G_INT / p, G_INT / (1 - p)), the gap
 * current against its closed form and the reference's io_raster_by_RKC.csv
 * (tests/test_oracle_io.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IO_COMP 3
#define IO_DEND 0
#define IO_SOMA 1
#define IO_AXON 2
#define V_INIT_IO (-60.0)
#define V_LEAK_IO (10.0)
#define V_Na_IO (55.0)
#define V_Ca_IO (120.0)
#define V_K_IO (-75.0)
#define V_H_IO (-43.0)
#define G_LEAK_IO (0.016)
#define SOMA_G_CAL_IO (0.83)
#define SOMA_G_NA_IO (120.0)
#define SOMA_G_KDR_IO (9.0)
#define SOMA_G_K_IO (5.0)
#define DNDR_G_H_IO (0.15)
#define DNDR_G_CAH_IO (4.5)
#define DNDR_G_KCA_IO (35.0)
#define AXON_G_NA_IO (240.0)
#define AXON_G_K_IO (20.0)
#define G_INT_IO (0.13)
#define G_P_DS_IO (0.200)
#define G_P_SA_IO (0.869)
#define G_C_IO (0.04)
#define CA_PH_IO (3.0)
#define BETA_CA_IO (0.075)

/* ion order (io.cuh:52-53) */
enum { k_CaL, l_CaL, m_Na, h_Na, n_Kdr, p_Kdr, x_K, r_CaH, s_KCa, q_H, N_ION_IO };
/* conductance order (io.cuh:55) */
enum { g_leak, g_CaL, g_Na, g_Kdr, g_K, g_CaH, g_KCa, g_H, N_COND_IO };

/* ---------------- rate functions, io_ion.cu:3-54 ---------------- */
static double alpha_r(double v) { return (1.7) / (1.0 + exp((-(v - 5.0)) / 13.9)); }
static double alpha_s(double x) { return fmin(0.00002 * x, 0.01); }
static double alpha_x(double v) { return (0.13 * (v + 25.0)) / (1.0 - exp(-(v + 25.0) / 10.0)); }
static double beta_r(double v) { return (0.02 * (v + 8.5)) / (exp((v + 8.5) / 5.0) - 1.0); }
static double beta_s(void) { return 0.015; }
static double beta_x(double v) { return 1.69 * exp(-0.0125 * (v + 35.0)); }
static double h_infty(double v, int eval)
{
  double h = 1.0;
  if (eval == 1) h = 1.0 / (1.0 + exp((-v - 70.0) / (-5.8)));
  if (eval == 2) h = 1.0 / (1.0 + exp((-v - 60.0) / (-5.8)));
  return h;
}
static double k_infty(double v) { return 1.0 / (1.0 + exp((-v - 61.0) / 4.2)); }
static double l_infty(double v) { return 1.0 / (1.0 + exp((v + 85.5) / (8.5))); }
static double m_infty(double v) { return 1.0 / (1.0 + exp((-v - 30.0) / (5.5))); }
static double n_infty(double v) { return 1.0 / (1.0 + exp((-v - 3.0) / (10.0))); }
static double p_infty(double v) { return 1.0 / (1.0 + exp((-v - 51.0) / (-12.0))); }
static double q_infty(double v) { return 1.0 / (1.0 + exp((v + 80.0) / 4.0)); }
static double r_infty(double v) { return alpha_r(v) / (alpha_r(v) + beta_r(v)); }
static double s_infty(double x) { return alpha_s(x) / (alpha_s(x) + beta_s()); }
static double x_infty(double v) { return alpha_x(v) / (alpha_x(v) + beta_x(v)); }
static double tau_h(double v, int eval)
{
  double h = 1.0;
  if (eval == 1) h = 3.0 * exp((-v - 40.0) / 33.0);
  if (eval == 2) h = 1.5 * exp((-v - 40.0) / 33.0);
  return h;
}
static double tau_k(double v) { (void)v; return 1.0; }
static double tau_l(double v) { return ((20.0 * exp((v + 160.0) / 30.0)) / (1.0 + exp((v + 84.0) / 7.3))) + 35.0; }
static double tau_n(double v) { return 47.0 * exp((-(-v - 50.0)) / 900.0) + 5.0; }
static double tau_p(double v) { return tau_n(v); }
static double tau_q(double v) { return 1.0 / (exp(-0.086 * v - 14.6) + exp(0.07 * v - 1.87)); }
static double tau_r(double v) { return 5.0 / (alpha_r(v) + beta_r(v)); }
static double tau_s(double x) { return 1.0 / (alpha_s(x) + beta_s()); }
static double tau_x(double v) { return 1.0 / (alpha_x(v) + beta_x(v)); }

/* exported for the closed-form tests: which = 2*gate (+1 for tau); comp is
   the compartment type (h_Na), x is Ca for s_KCa */
double or_io_rate(int which, double v, int comp)
{
  int g = which / 2, t = which % 2;
  switch (g) {
    case k_CaL: return t ? tau_k(v) : k_infty(v);
    case l_CaL: return t ? tau_l(v) : l_infty(v);
    case m_Na: return t ? NAN : m_infty(v);
    case h_Na: return t ? tau_h(v, comp) : h_infty(v, comp);
    case n_Kdr: return t ? tau_n(v) : n_infty(v);
    case p_Kdr: return t ? tau_p(v) : p_infty(v);
    case x_K: return t ? tau_x(v) : x_infty(v);
    case r_CaH: return t ? tau_r(v) : r_infty(v);
    case s_KCa: return t ? tau_s(v) : s_infty(v);
    case q_H: return t ? tau_q(v) : q_infty(v);
  }
  return NAN;
}

static double *dal(int n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

/* neighbour list of cell id: gap.cu:4-133 (pre cell or -1, in slot order) */
static void io_neighbours(int id, int nx, int ny, int out[4])
{
  const int idx = id % nx, idy = id / ny;   /* the reference divides by ny */
  int l = id - 1, u = id - nx, r = id + 1, d = id + nx;
  if (idx == 0 && idy == 0) { out[0] = -1; out[1] = -1; out[2] = r; out[3] = d; }
  else if (idx == nx - 1 && idy == 0) { out[0] = l; out[1] = -1; out[2] = -1; out[3] = d; }
  else if (idx == 0 && idy == ny - 1) { out[0] = -1; out[1] = u; out[2] = r; out[3] = -1; }
  else if (idx == nx - 1 && idy == ny - 1) { out[0] = l; out[1] = u; out[2] = -1; out[3] = -1; }
  else if (idy == 0) { out[0] = l; out[1] = -1; out[2] = r; out[3] = d; }
  else if (idy == ny - 1) { out[0] = l; out[1] = u; out[2] = r; out[3] = -1; }
  else if (idx == 0) { out[0] = -1; out[1] = u; out[2] = r; out[3] = d; }
  else if (idx == nx - 1) { out[0] = l; out[1] = u; out[2] = -1; out[3] = d; }
  else { out[0] = l; out[1] = u; out[2] = r; out[3] = d; }
}

void or_io_neighbours(int id, int nx, int ny, int *out4) { io_neighbours(id, nx, ny, out4); }

or_pop *or_io_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand)
{
  FILE *f = fopen(param_file, "r");
  if (!f) return NULL;
  int connect[IO_COMP], compart[IO_COMP];
  double cmv[IO_COMP], rad[IO_COMP], len[IO_COMP];
  for (int k = 0; k < IO_COMP; k++) {
    int i1, i2, i3;
    double d1, d2, d4;
    if (fscanf(f, "%d %d %lf %lf %lf %d ", &i1, &i2, &d1, &d2, &d4, &i3) != 6 || i1 < 0 || i1 >= IO_COMP) {
      fclose(f);
      return NULL;
    }
    rad[i1] = 0.5 * d1 * 1e-4;
    len[i1] = d2 * 1e-4;
    connect[i1] = i2;
    compart[i1] = i3;
    cmv[i1] = d4;
  }
  fclose(f);

  const int n = nx * ny, nc = n * IO_COMP;
  or_pop *p = (or_pop *)calloc(1, sizeof(or_pop));
  p->type = OR_IO; p->n = n; p->nx = nx; p->ny = ny; p->comp = IO_COMP; p->nc = nc;
  p->ncond = N_COND_IO; p->nion = N_ION_IO;
  p->v = dal(nc); p->ca = dal(nc); p->cm = dal(nc); p->area = dal(nc);
  p->iext = dal(nc); p->shell = dal(nc); p->rev = dal(nc);
  p->compart = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < N_COND_IO; i++) p->cond[i] = dal(nc);
  for (int i = 0; i < N_ION_IO; i++) p->ion[i] = dal(nc);
  p->b = dal(nc); p->dammy = dal(nc);
  for (int i = 0; i < OR_NVEC; i++) p->vec[i] = dal(nc);
  p->hmax = 0.125;
  p->hmin = pow(2.0, -10.0);   /* H_MIN_IO, solve_rkc.cuh:27 */

  /* per-compartment parameters, io.cu:30-100 (densities; no area factor) */
  for (int c = 0; c < n; c++)
    for (int k = 0; k < IO_COMP; k++) {
      const int i = c * IO_COMP + k, t = compart[k];
      p->compart[i] = t;
      p->cm[i] = cmv[k];
      p->area[i] = 2.0 * M_PI * rad[k] * len[k];
      double g[N_COND_IO] = {0};
      g[g_leak] = G_LEAK_IO;
      if (t == IO_DEND) { g[g_CaH] = DNDR_G_CAH_IO; g[g_KCa] = DNDR_G_KCA_IO; g[g_H] = DNDR_G_H_IO; }
      else if (t == IO_SOMA) { g[g_CaL] = SOMA_G_CAL_IO; g[g_Na] = SOMA_G_NA_IO; g[g_Kdr] = SOMA_G_KDR_IO; g[g_K] = SOMA_G_K_IO; }
      else if (t == IO_AXON) { g[g_Na] = AXON_G_NA_IO; g[g_K] = AXON_G_K_IO; }
      for (int q = 0; q < N_COND_IO; q++) p->cond[q][i] = g[q];
    }

  /* coupling matrix, io_solve.cu:36-66, then CSR with the RKC sign (off-diagonal +g) */
  double mat[IO_COMP * IO_COMP];
  memset(mat, 0, sizeof(mat));
  for (int i = 0; i < IO_COMP; i++) {
    int d = connect[i], i3 = compart[i];
    double lg = 0.0, lg_rev = 0.0;
    if (d < 0) continue;
    if (i3 == 0 && d == 1) { lg = G_INT_IO / (1.0 - G_P_DS_IO); lg_rev = G_INT_IO / G_P_DS_IO; }
    else if (i3 == 1 && d == 2) { lg = G_INT_IO / G_P_SA_IO; lg_rev = G_INT_IO / (1.0 - G_P_SA_IO); }
    else if (i3 == 0 && d < 1) { lg = lg_rev = G_INT_IO / 0.5; }
    else { or_pop_free(p); return NULL; }
    mat[d + IO_COMP * i] = lg;
    mat[i + IO_COMP * d] = lg_rev;
  }
  int nnz1 = 0;
  for (int r = 0; r < IO_COMP; r++)
    for (int c = 0; c < IO_COMP; c++) nnz1 += (mat[c + IO_COMP * r] > 0.0 || r == c);
  p->nnz = nnz1 * n;
  p->row = (int *)malloc(sizeof(int) * (nc + 1));
  p->col = (int *)malloc(sizeof(int) * p->nnz);
  p->val = (double *)malloc(sizeof(double) * p->nnz);
  int k = 0;
  p->row[0] = 0;
  for (int cell = 0; cell < n; cell++)
    for (int r = 0; r < IO_COMP; r++) {
      double dsum = 0.0;
      for (int c = 0; c < IO_COMP; c++)
        if (c != r) dsum += mat[c + IO_COMP * r];
      for (int c = 0; c < IO_COMP; c++) {
        if (c == r) { p->col[k] = cell * IO_COMP + c; p->val[k] = -dsum; k++; }
        else if (mat[c + IO_COMP * r] > 0.0) { p->col[k] = cell * IO_COMP + c; p->val[k] = mat[c + IO_COMP * r]; k++; }
      }
      p->row[cell * IO_COMP + r + 1] = k;
    }

  /* initial state, io_ion.cu:163-194 */
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % IO_COMP == 0 && !v0_per_cell)
      init_v_rand = use_libc_rand ? ((double)rand() / RAND_MAX) - 0.5 : 0.0;
    double vv = v0_per_cell ? v0_per_cell[i / IO_COMP] : V_INIT_IO + 10.0 * init_v_rand;
    p->v[i] = vv;
    p->iext[i] = 0.0;
    int lc = p->compart[i];
    p->ion[k_CaL][i] = k_infty(vv);
    p->ion[l_CaL][i] = l_infty(vv);
    p->ion[n_Kdr][i] = n_infty(vv);
    p->ion[p_Kdr][i] = p_infty(vv);
    p->ion[x_K][i] = x_infty(vv);
    p->ion[r_CaH][i] = r_infty(vv);
    p->ion[q_H][i] = q_infty(vv);
    double I_Ca1 = p->cond[g_CaH][i] * p->ion[r_CaH][i] * p->ion[r_CaH][i] * (vv - V_Ca_IO);
    p->ca[i] = -I_Ca1 * CA_PH_IO / BETA_CA_IO;
    p->ion[s_KCa][i] = s_infty(p->ca[i]);
    p->ion[m_Na][i] = m_infty(vv);
    p->ion[h_Na][i] = h_infty(vv, lc);
  }
  p->stim_comp = 0;
  p->stim_t0 = 800.0;
  p->stim_t1 = 820.0;
  p->stim_amp = dal(n);
  for (int c = 0; c < n; c++) p->stim_amp[c] = 5.5;   /* io.cu:253 [uA/cm^2] */
  p->others[OR_ATOL] = atol;
  p->others[OR_RTOL] = rtol;
  if (or_rkc_init(p) != 0) { or_pop_free(p); return NULL; }
  return p;
}

/* io_update_matrix, solve_rkc.cu:412-434, then the gap currents of
   add_io_gap_val (solve_rkc.cu:400-410) when add_gap (the reference adds them
   only at ion updates, not at initialisation) */
void or_io_update_b(or_pop *p, const double *vv, int add_gap)
{
  double **c = p->cond, **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    double l_v = vv[id];
    p->b[id] = (-p->iext[id]
                + c[g_leak][id] * (l_v - V_LEAK_IO)
                + c[g_CaL][id] * (l_v - V_Ca_IO) * g[k_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[l_CaL][id]
                + c[g_Na][id] * (l_v - V_Na_IO) * g[m_Na][id] * g[m_Na][id] * g[m_Na][id] * g[h_Na][id]
                + c[g_Kdr][id] * (l_v - V_K_IO) * g[n_Kdr][id] * g[p_Kdr][id]
                + c[g_K][id] * (l_v - V_K_IO) * g[x_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id]
                + c[g_CaH][id] * (l_v - V_Ca_IO) * g[r_CaH][id] * g[r_CaH][id]
                + c[g_KCa][id] * (l_v - V_K_IO) * g[s_KCa][id]
                + c[g_H][id] * (l_v - V_H_IO) * g[q_H][id]);
  }
  if (add_gap && p->n > 1) {
    for (int cell = 0; cell < p->n; cell++) {
      int nb[4];
      io_neighbours(cell, p->nx, p->ny, nb);
      const int post = cell * IO_COMP;
      for (int s = 0; s < 4; s++) {
        /* partners outside the grid (non-square grids, where the reference
           reads past its arrays) are treated as absent */
        const int pre = (nb[s] > -1 && nb[s] < p->n) ? nb[s] : -1;
        double diff_v = (pre > -1) ? (p->v[post] - p->v[pre * IO_COMP]) : 0.0;
        p->b[post] += G_C_IO * (0.8 * exp(-0.01 * diff_v * diff_v) + 0.2) * (diff_v);
      }
    }
  }
}

#define CNU(gt, INF, TAU) \
  g[gt][id] = (2.0 * DT * (INF) + (2.0 * (TAU) - DT) * g[gt][id]) / (2.0 * (TAU) + DT)

/* Ca and gates of one compartment at the gate voltage v_val; the Ca current
   uses the stored voltage p->v (io_update_ion_RKC, io_ion.cu:90-127, and
   io_update_ion_2nd, :53-89, differ only in v_val) */
static void io_ion_at(or_pop *p, int id, double v_val, double DT)
{
  double **g = p->ion;
  int l_comp = p->compart[id];
  double Ca_val = p->ca[id];
  double I_Ca1 = p->cond[g_CaH][id] * g[r_CaH][id] * g[r_CaH][id] * (p->v[id] - V_Ca_IO);
  double k1 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val));
  double k2 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k1 / 2.0));
  double k3 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k2 / 2.0));
  double k4 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k3));
  p->ca[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  Ca_val = (Ca_val + p->ca[id]) / 2.0;
  CNU(k_CaL, k_infty(v_val), tau_k(v_val));
  CNU(l_CaL, l_infty(v_val), tau_l(v_val));
  CNU(n_Kdr, n_infty(v_val), tau_n(v_val));
  CNU(p_Kdr, p_infty(v_val), tau_p(v_val));
  CNU(x_K, x_infty(v_val), tau_x(v_val));
  CNU(r_CaH, r_infty(v_val), tau_r(v_val));
  CNU(q_H, q_infty(v_val), tau_q(v_val));
  CNU(s_KCa, s_infty(Ca_val), tau_s(Ca_val));
  g[m_Na][id] = m_infty(v_val);
  CNU(h_Na, h_infty(v_val, l_comp), tau_h(v_val, l_comp));
}

/* io_update_ion_RKC, io_ion.cu:90-127: gates at the mid-step voltage */
void or_io_update_ion(or_pop *p, const double *vnew, double DT)
{
  for (int id = 0; id < p->nc; id++) {
    io_ion_at(p, id, 0.5 * (vnew[id] + p->v[id]), DT);
    p->v[id] = vnew[id];
  }
}

/* io_update_ion_2nd, io_ion.cu:53-89 (CN path): gates at the voltage
   extrapolated from the last two steps */
void or_io_cn_ion(or_pop *p, double DT)
{
  for (int id = 0; id < p->nc; id++) {
    io_ion_at(p, id, 1.5 * p->v[id] - 0.5 * p->cn_vold[id], DT);
    p->cn_vold[id] = p->v[id];
  }
}

/* conductance and reversal sums of the CN path before halving
   (io_cnm_vec_initialize / io_update_matrix, solve_cnm.cu:751-768, 793-810) */
void or_io_cn_ge(const or_pop *p, int id, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = (c[g_leak][id]
           + c[g_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[l_CaL][id]
           + c[g_Na][id] * g[m_Na][id] * g[m_Na][id] * g[m_Na][id] * g[h_Na][id]
           + c[g_Kdr][id] * g[n_Kdr][id] * g[p_Kdr][id]
           + c[g_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id]
           + c[g_CaH][id] * g[r_CaH][id] * g[r_CaH][id]
           + c[g_KCa][id] * g[s_KCa][id]
           + c[g_H][id] * g[q_H][id]);
  *esum = (c[g_leak][id] * V_LEAK_IO + p->iext[id]
           + c[g_CaL][id] * V_Ca_IO * g[k_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[l_CaL][id]
           + c[g_Na][id] * V_Na_IO * g[m_Na][id] * g[m_Na][id] * g[m_Na][id] * g[h_Na][id]
           + c[g_Kdr][id] * V_K_IO * g[n_Kdr][id] * g[p_Kdr][id]
           + c[g_K][id] * V_K_IO * g[x_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id]
           + c[g_CaH][id] * V_Ca_IO * g[r_CaH][id] * g[r_CaH][id]
           + c[g_KCa][id] * V_K_IO * g[s_KCa][id]
           + c[g_H][id] * V_H_IO * g[q_H][id]);
}

/* base minus each gap current into the cell's dendrite from the stored
   voltages (io_gap_update, gap.cu:142-160), partners in slot order; the CN
   and BE paths subtract the currents on the right-hand side (add_io_gap_val,
   solve_cnm.cu:729-739, solve_bem.cu:463-473) */
double or_io_gap_add(const or_pop *p, int cell, double base)
{
  int nb[4];
  double s = base;
  io_neighbours(cell, p->nx, p->ny, nb);
  const int post = cell * IO_COMP;
  for (int k = 0; k < 4; k++) {
    const int pre = (nb[k] > -1 && nb[k] < p->n) ? nb[k] : -1;
    double diff_v = (pre > -1) ? (p->v[post] - p->v[pre * IO_COMP]) : 0.0;
    s += -1.0 * (G_C_IO * (0.8 * exp(-0.01 * diff_v * diff_v) + 0.2) * (diff_v));
  }
  return s;
}

double or_io_gap_into(const or_pop *p, int cell) { return or_io_gap_add(p, cell, 0.0); }

/* ---------------- the backward-Euler (BE) path ---------------- */
#define EEU(gt, INF, TAU) \
  do { const double inf_ = (INF); g[gt][id] = inf_ + (g[gt][id] - inf_) * exp(-DT / (TAU)); } while (0)

/* io_update_ion, io_ion.cu:128-161 (the BE path): every gate by exponential
   Euler at V (s_KCa at the present Ca, m_Na instantaneous), then the Ca RK4
   with the Ca current of the updated r_CaH gate */
void or_io_be_ion(or_pop *p, double DT)
{
  double **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const int l_comp = p->compart[id];
    const double v_val = p->v[id], Ca_val = p->ca[id];
    EEU(k_CaL, k_infty(v_val), tau_k(v_val));
    EEU(l_CaL, l_infty(v_val), tau_l(v_val));
    EEU(n_Kdr, n_infty(v_val), tau_n(v_val));
    EEU(p_Kdr, p_infty(v_val), tau_p(v_val));
    EEU(x_K, x_infty(v_val), tau_x(v_val));
    EEU(r_CaH, r_infty(v_val), tau_r(v_val));
    EEU(q_H, q_infty(v_val), tau_q(v_val));
    EEU(s_KCa, s_infty(Ca_val), tau_s(Ca_val));
    g[m_Na][id] = m_infty(v_val);
    EEU(h_Na, h_infty(v_val, l_comp), tau_h(v_val, l_comp));
    const double I_Ca1 = p->cond[g_CaH][id] * g[r_CaH][id] * g[r_CaH][id] * (p->v[id] - V_Ca_IO);
    const double k1 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val));
    const double k2 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k1 / 2.0));
    const double k3 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k2 / 2.0));
    const double k4 = DT * (-I_Ca1 * CA_PH_IO - BETA_CA_IO * (Ca_val + k3));
    p->ca[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  }
}

/* the BE system's diagonal addend and right-hand side, Cm/DT first, in the
   reference's term order (io_update_matrix, solve_bem.cu:478-508) */
void or_io_be_ge(const or_pop *p, int id, double DT, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = ((p->cm[id] / DT)
           + c[g_leak][id]
           + c[g_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[l_CaL][id]
           + c[g_Na][id] * g[m_Na][id] * g[m_Na][id] * g[m_Na][id] * g[h_Na][id]
           + c[g_Kdr][id] * g[n_Kdr][id] * g[p_Kdr][id]
           + c[g_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id]
           + c[g_CaH][id] * g[r_CaH][id] * g[r_CaH][id]
           + c[g_KCa][id] * g[s_KCa][id]
           + c[g_H][id] * g[q_H][id]);
  *esum = ((p->cm[id] / DT) * p->v[id]
           + c[g_leak][id] * V_LEAK_IO + p->iext[id]
           + c[g_CaL][id] * V_Ca_IO * g[k_CaL][id] * g[k_CaL][id] * g[k_CaL][id] * g[l_CaL][id]
           + c[g_Na][id] * V_Na_IO * g[m_Na][id] * g[m_Na][id] * g[m_Na][id] * g[h_Na][id]
           + c[g_Kdr][id] * V_K_IO * g[n_Kdr][id] * g[p_Kdr][id]
           + c[g_K][id] * V_K_IO * g[x_K][id] * g[x_K][id] * g[x_K][id] * g[x_K][id]
           + c[g_CaH][id] * V_Ca_IO * g[r_CaH][id] * g[r_CaH][id]
           + c[g_KCa][id] * V_K_IO * g[s_KCa][id]
           + c[g_H][id] * V_H_IO * g[q_H][id]);
}
