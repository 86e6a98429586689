/*
 * Oracle: granule cell model on the RKC path (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Restates:
 *   constants                  gr.cuh:11-40
 *   initialize_host_gr         gr.cu:3-82     (Gr.txt: 17 columns, densities x area, Cm = Cm' x area)
 *   set_host_gr_solve          gr_solve.cu:2-139 (cable conductances, CSR, RKC sign)
 *   rate functions             gr_ion.cu:3-43
 *   gr_Na_update_2order        gr_ion.cu:45-207 (12-state Na scheme on hillock, AIS, axon and
 *                                               parallel fibre; Crank-Nicolson, 12 x 12
 *                                               partial-pivot Gaussian elimination)
 *   gr_update_ion_RKC_exp_imp  gr_ion.cu:461-480 (CN gates of KV, KA, KIR, KM, KCa; Ca and the
 *                                               Ca gates are not advanced on the RKC path)
 *   gr_initialize_ion          gr_ion.cu:635-672
 *   gr_update_matrix           solve_rkc.cu:226-249
 *   gr_update_ion_and_matrix   solve_rkc.cu:1100-1119 (Na scheme, gates, then b)
 *   gr_set_current             gr.cu:245-254
 *   gr_update_ion_exp_imp      gr_ion.cu:438-459 (CN path)
 *   CN conductance sums        solve_cnm.cu:128-208
 *
 * Reference quirks kept: the Na scheme's explicit half uses the voltage of
 * the previous ion update and its implicit half the new one; Ca stays at its
 * resting value and ch_Ca / ci_Ca at their initial steady state on the RKC
 * path; the initial voltage draw is (rand() + 1) / (RAND_MAX + 2) - 0.5.
 * Without network synapses the reference's mossy-fibre and Golgi-cell
 * synaptic terms (solve_rkc.cu:1117-1118) add zero and are left out.
 *
 * Pinning (tests/test_oracle_gr.py): rates against their identities and the
 * guarded singular point of alpha_n_KV, the CN gates against the (1,1) Pade
 * form, the Na step against an independent dense solve of the Crank-Nicolson
 * system built from the scheme's state diagram (and its conservation of
 * total occupancy), the cable matrix against the series half-segment
 * conductances from Gr.txt geometry.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GR_COMP 578
#define V_INIT_GR (-80.0)
#define V_LEAK1_GR (-16.5)
#define V_LEAK2_GR (-65.0)
#define V_LEAK3_GR (-74.5)
#define V_Na_GR (87.39)
#define V_K_GR (-84.69)
#define V_Ca_GR (129.33)
#define Ca1_0_GR (1e-4)
#define GR_hill 5

/* ion order (gr.cuh:36-37) */
enum { n_KV, a_KA, b_KA, c_KCa, s_KM, ir_KIR, ch_Ca, ci_Ca, o_Na, c1_Na, c2_Na, c3_Na, c4_Na, c5_Na, i1_Na, i2_Na,
       i3_Na, i4_Na, i5_Na, i6_Na, N_ION_GR };
/* conductance order (gr.cuh:40) */
enum { g_leak1, g_leak2, g_leak3, g_Na, g_KV, g_KA, g_KCa, g_KM, g_KIR, g_Ca, N_COND_GR };

/* ---------------- rate functions, gr_ion.cu:3-43 ---------------- */
static double a_n_KV(double v)
{
  return (fabs(v + 25.0) > 1.0e-6) ? -0.017 * ((v + 25.0) / (exp(-(v + 25.0) / 10.0) - 1.0)) : -0.17 * (1.0 + (v + 25.0) / 20.0);
}
static double b_n_KV(double v) { return 0.2223 * exp(-(v + 35.0) / 80.0); }
static double a_a_KA(double v) { return 0.771 / (1.0 + exp(-(v + 9.17203) / 23.32708)); }
static double b_a_KA(double v) { return 0.1567 / exp((v + 18.27914) / 19.47175); }
static double a_b_KA(double v) { return 0.0348 / (1.0 + exp((v + 111.33208) / 12.8433)); }
static double b_b_KA(double v) { return 0.0327 / (1.0 + exp(-(v + 49.9537) / 8.90123)); }
static double a_ir_KIR(double v) { return 0.2302 * exp(-(v + 83.94) / 24.3902); }
static double b_ir_KIR(double v) { return 0.2943 * exp((v + 83.94) / 35.714); }
static double a_s_KM(double v) { return 0.0046 * exp((v + 30.0) / 40.0); }
static double b_s_KM(double v) { return 0.0046 * exp(-(v + 30.0) / 20.0); }
static double a_c_KCa(double v, double ca) { return 1.4433 / (1.0 + (1.5e-3 * exp(-v / 11.765)) / ca); }
static double b_c_KCa(double v, double ca) { return 0.8660 / (1.0 + ca / (1.5e-4 * exp(-v / 11.765))); }
static double a_ch_Ca(double v) { return 0.0856 * exp((v + 29.6) / 15.873); }
static double b_ch_Ca(double v) { return 0.1437 * exp(-(v + 18.66) / 25.641); }
static double a_ci_Ca(double v) { return 0.0023 * exp(-(v + 48.0) / 18.183); }
static double b_ci_Ca(double v) { return 0.0025 * exp((v + 48.0) / 83.33); }

static double inf_n_KV(double v) { return (a_n_KV(v) / (a_n_KV(v) + b_n_KV(v))); }
static double tau_n_KV(double v) { return (1.0 / (a_n_KV(v) + b_n_KV(v))); }
static double inf_a_KA(double v) { return 1.0 / (1.0 + exp(-(v + 38.0) / 17.0)); }
static double tau_a_KA(double v) { return (1.0 / (a_a_KA(v) + b_a_KA(v))); }
static double inf_b_KA(double v) { return 1.0 / (1.0 + exp((v + 78.8) / 8.4)); }
static double tau_b_KA(double v) { return (1.0 / (a_b_KA(v) + b_b_KA(v))); }
static double inf_ir_KIR(double v) { return (a_ir_KIR(v) / (a_ir_KIR(v) + b_ir_KIR(v))); }
static double tau_ir_KIR(double v) { return (1.0 / (a_ir_KIR(v) + b_ir_KIR(v))); }
static double inf_s_KM(double v) { return 1.0 / (1.0 + exp(-(v + 35.0) / 6.0)); }
static double tau_s_KM(double v) { return (1.0 / (a_s_KM(v) + b_s_KM(v))); }
static double inf_c_KCa(double v, double ca) { return a_c_KCa(v, ca) / (a_c_KCa(v, ca) + b_c_KCa(v, ca)); }
static double tau_c_KCa(double v, double ca) { return (1.0 / (a_c_KCa(v, ca) + b_c_KCa(v, ca))); }
static double inf_ch_Ca(double v) { return (a_ch_Ca(v) / (a_ch_Ca(v) + b_ch_Ca(v))); }
static double tau_ch_Ca(double v) { return (1.0 / (a_ch_Ca(v) + b_ch_Ca(v))); }
static double inf_ci_Ca(double v) { return (a_ci_Ca(v) / (a_ci_Ca(v) + b_ci_Ca(v))); }
static double tau_ci_Ca(double v) { return (1.0 / (a_ci_Ca(v) + b_ci_Ca(v))); }

/* exported for the tests: which = 2 * gate (+1 for tau) for the gates
   n_KV .. ci_Ca; 1000 + 2 * gate (+1) for alpha / beta */
double or_gr_rate(int which, double v, double ca)
{
  if (which >= 1000) {
    const int k = which - 1000, g = k / 2, b = k % 2;
    switch (g) {
      case n_KV: return b ? b_n_KV(v) : a_n_KV(v);
      case a_KA: return b ? b_a_KA(v) : a_a_KA(v);
      case b_KA: return b ? b_b_KA(v) : a_b_KA(v);
      case c_KCa: return b ? b_c_KCa(v, ca) : a_c_KCa(v, ca);
      case s_KM: return b ? b_s_KM(v) : a_s_KM(v);
      case ir_KIR: return b ? b_ir_KIR(v) : a_ir_KIR(v);
      case ch_Ca: return b ? b_ch_Ca(v) : a_ch_Ca(v);
      case ci_Ca: return b ? b_ci_Ca(v) : a_ci_Ca(v);
    }
    return NAN;
  }
  const int g = which / 2, t = which % 2;
  switch (g) {
    case n_KV: return t ? tau_n_KV(v) : inf_n_KV(v);
    case a_KA: return t ? tau_a_KA(v) : inf_a_KA(v);
    case b_KA: return t ? tau_b_KA(v) : inf_b_KA(v);
    case c_KCa: return t ? tau_c_KCa(v, ca) : inf_c_KCa(v, ca);
    case s_KM: return t ? tau_s_KM(v) : inf_s_KM(v);
    case ir_KIR: return t ? tau_ir_KIR(v) : inf_ir_KIR(v);
    case ch_Ca: return t ? tau_ch_Ca(v) : inf_ch_Ca(v);
    case ci_Ca: return t ? tau_ci_Ca(v) : inf_ci_Ca(v);
  }
  return NAN;
}

/* Na scheme parameters by compartment type (gr_ion.cu:57-69; only types
   GR_hill and above use them) */
static const double Na_n1[] = {1.0, 1.0, 1.0, 1.0, 1.0, 5.422, 5.422, 5.422, 5.422};
static const double Na_n2[] = {1.0, 1.0, 1.0, 1.0, 1.0, 3.279, 3.279, 3.279, 3.279};
static const double Na_n3[] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.83, 1.83, 1.83, 1.83};
static const double Na_n4[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.738, 0.738, 0.738, 0.738};
static const double Na_Aalpha[] = {1.0, 1.0, 1.0, 1.0, 1.0, 8.85, 8.85, 8.85, 8.85};
static const double Na_Abeta[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.0318, 0.0318, 0.0318, 0.0318};
static const double Na_Vshift[] = {1.0, 1.0, 1.0, 1.0, 1.0, -22.0, -22.0, -25.0, -25.0};
static const double Na_Con[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.003464, 0.003464, 0.09353, 0.09353};
static const double Na_Coff[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.4157, 0.4157, 0.04676, 0.04676};
static const double Na_Oon[] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.8186, 1.8186, 2.5981, 2.5981};
static const double Na_Ooff[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.006928, 0.006928, 0.000173, 0.000173};

/* the scheme's matrix at (alpha, beta) plus s * (2/DT) on the diagonal
   (gr_ion.cu:93-106 with s = +1 after the sign flip of the explicit half,
   :121-134 with s = -1); states c1..c5, i1..i6, o */
static void na_matrix(double M[12][12], int t, double alpha, double beta, double s2dt)
{
  const double n1 = Na_n1[t], n2 = Na_n2[t], n3 = Na_n3[t], n4 = Na_n4[t];
  const double Con = Na_Con[t], Coff = Na_Coff[t], Oon = Na_Oon[t], Ooff = Na_Ooff[t];
  const double a = pow((Na_Oon[t] / Na_Con[t]), 0.25);
  const double b = pow((Na_Ooff[t] / Na_Coff[t]), 0.25);
  const double a4 = pow(a, 4.0), b4 = pow(b, 4.0);
  memset(M, 0, sizeof(double) * 144);
  /* diagonal entries are -(sum of outflow rates) -/+ 2/DT, written as the
     reference writes them: -(rates - 2/DT) for the explicit half and
     -(rates + 2/DT) for the implicit one */
  M[0][0] = -(n1 * alpha + Con + s2dt);
  M[0][1] = n4 * beta; M[0][5] = Coff;
  M[1][0] = n1 * alpha; M[1][1] = -(n4 * beta + n2 * alpha + Con * a + s2dt); M[1][2] = n3 * beta; M[1][6] = Coff * b;
  M[2][1] = n2 * alpha; M[2][2] = -(n3 * beta + n3 * alpha + Con * a * a + s2dt); M[2][3] = n2 * beta; M[2][7] = Coff * b * b;
  M[3][2] = n3 * alpha; M[3][3] = -(n2 * beta + n4 * alpha + Con * a * a * a + s2dt); M[3][4] = n1 * beta;
  M[3][8] = Coff * b * b * b;
  M[4][3] = n4 * alpha; M[4][4] = -(n1 * beta + 259.8 + Con * a4 + s2dt); M[4][9] = Coff * b4; M[4][11] = 69.3;
  M[5][0] = Con; M[5][5] = -(Coff + n1 * alpha * a + s2dt); M[5][6] = n4 * beta * b;
  M[6][1] = Con * a; M[6][5] = n1 * alpha * a; M[6][6] = -(Coff * b + n4 * beta * b + n2 * alpha * a + s2dt);
  M[6][7] = n3 * beta * b;
  M[7][2] = Con * a * a; M[7][6] = n2 * alpha * a; M[7][7] = -(Coff * b * b + n3 * beta * b + n3 * alpha * a + s2dt);
  M[7][8] = n2 * beta * b;
  M[8][3] = Con * a * a * a; M[8][7] = n3 * alpha * a;
  M[8][8] = -(Coff * b * b * b + n2 * beta * b + n4 * alpha * a + s2dt); M[8][9] = n1 * beta * b;
  M[9][4] = Con * a4; M[9][8] = n4 * alpha * a; M[9][9] = -(Coff * b4 + n1 * beta * b + 259.8 + s2dt); M[9][10] = 69.3;
  M[10][9] = 259.8; M[10][10] = -(Ooff + 69.3 + s2dt); M[10][11] = Oon;
  M[11][4] = 259.8; M[11][10] = Ooff; M[11][11] = -(69.3 + Oon + s2dt);
}

/* Gaussian elimination with partial pivoting on the 12 x 12 system M r' = r,
   as gr_ion.cu:150-199 and :276-330 write it (strictly larger |pivot|
   wins, rows swapped from the pivot column on); returns 1 and leaves r
   alone on a pivot below 1e-12 (the reference returns there). */
static int solve12(double M[12][12], double *x, const double *r_in)
{
  double r[12];
  memcpy(r, r_in, sizeof(r));
  for (int i = 0; i < 11; i++) {
    int piv = i;
    double pmax = fabs(M[i][i]);
    for (int j = i + 1; j < 12; j++)
      if (fabs(M[j][i]) > pmax) { piv = j; pmax = fabs(M[j][i]); }
    if (fabs(pmax) < 1.0e-12) return 1;
    if (piv != i) {
      for (int j = i; j < 12; j++) { double tmp = M[i][j]; M[i][j] = M[piv][j]; M[piv][j] = tmp; }
      double tmp = r[i]; r[i] = r[piv]; r[piv] = tmp;
    }
    for (int j = i + 1; j < 12; j++) {
      double w = M[j][i] / M[i][i];
      M[j][i] = 0.0;
      for (int k = i + 1; k < 12; k++) M[j][k] = M[j][k] - M[i][k] * w;
      r[j] = r[j] - r[i] * w;
    }
  }
  for (int i = 11; i >= 0; i--) {
    for (int j = i + 1; j < 12; j++) r[i] = r[i] - M[i][j] * r[j];
    r[i] = r[i] / M[i][i];
  }
  memcpy(x, r, sizeof(r));
  return 0;
}

/* One Na step, gr_ion.cu:86-202.  x = (c1..c5, i1..i6, o); returns 1 and
   leaves x alone on a pivot below 1e-12 (the reference returns there). */
static int na_step(int t, double v_old, double v_new, double DT, double *x)
{
  double M[12][12], r[12];
  double alpha = Na_Aalpha[t] * exp((v_old - Na_Vshift[t]) / 10.0);
  double beta = Na_Abeta[t] * exp(-(v_old - Na_Vshift[t]) / 10.0);
  na_matrix(M, t, alpha, beta, -2.0 / DT);
  for (int i = 0; i < 12; i++) {
    r[i] = 0.0;
    for (int j = 0; j < 12; j++) r[i] -= x[j] * M[i][j];
  }
  alpha = Na_Aalpha[t] * exp((v_new - Na_Vshift[t]) / 10.0);
  beta = Na_Abeta[t] * exp(-(v_new - Na_Vshift[t]) / 10.0);
  na_matrix(M, t, alpha, beta, 2.0 / DT);
  return solve12(M, x, r);
}

/* One implicit-Euler Na step at voltage v (gr_Na_update, gr_ion.cu:209-357,
   the BE path): (I - DT A(v)) x' = x, the matrix formed as the reference
   forms it (every entry times -DT, then 1 added to the diagonal). */
static int na_be(int t, double v, double DT, double *x)
{
  double M[12][12];
  const double alpha = Na_Aalpha[t] * exp((v - Na_Vshift[t]) / 10.0);
  const double beta = Na_Abeta[t] * exp(-(v - Na_Vshift[t]) / 10.0);
  na_matrix(M, t, alpha, beta, 0.0);
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) M[i][j] *= -DT;
  for (int i = 0; i < 12; i++) M[i][i] += 1.0;
  return solve12(M, x, x);
}

/* test hook: x12 = (c1..c5, i1..i6, o) in and out */
int or_gr_na_step(int type, double v_old, double v_new, double dt, double *x12)
{
  if (type < GR_hill || type > 8) return 2;
  return na_step(type, v_old, v_new, dt, x12);
}

/* test hook: one implicit-Euler Na step at v */
int or_gr_na_be(int type, double v, double dt, double *x12)
{
  if (type < GR_hill || type > 8) return 2;
  return na_be(type, v, dt, x12);
}

static double *dal(int n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

/* A granule population whose cells have `comp` compartments (the reference's
 * GR_COMP = 578 in or_gr_create; fewer for synthetic chains written in the
 * Gr.txt format, one block of comp rows shared by all cells). */
or_pop *or_gr_create_comp(const char *param_file, int comp, int nx, int ny, const double *v0_per_cell, double atol,
                          double rtol, int use_libc_rand)
{
  if (comp < 1 || comp > GR_COMP) return NULL;
  FILE *f = fopen(param_file, "r");
  if (!f) return NULL;
  double *rad = dal(comp), *len = dal(comp), *Ra = dal(comp), *cmd = dal(comp);
  double (*dens)[N_COND_GR] = (double (*)[N_COND_GR])calloc((size_t)comp, sizeof(double[N_COND_GR]));
  int *l_connect = (int *)calloc((size_t)comp, sizeof(int)), *compart = (int *)calloc((size_t)comp, sizeof(int));
#define GR_CREATE_FREE() do { free(rad); free(len); free(Ra); free(cmd); free(dens); free(l_connect); free(compart); } while (0)
  for (int i = 0; i < comp; i++) {
    int i1, i2, i3;
    double d1, d2, d3, d4, x[N_COND_GR];
    if (fscanf(f, "%d %d %lf %lf %lf %lf %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf ", &i1, &i2, &d1, &d2, &d3, &d4,
               &i3, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6], &x[7], &x[8], &x[9]) != 17 || i1 < 0 ||
        i1 >= comp || i3 < 0 || i3 > 8) {
      fclose(f);
      GR_CREATE_FREE();
      return NULL;
    }
    rad[i1] = 0.5 * d1 * 1e-4;
    len[i1] = d2 * 1e-4;
    Ra[i1] = d3 * 1e-3;
    cmd[i1] = d4;
    compart[i1] = i3;
    for (int q = 0; q < N_COND_GR; q++) dens[i1][q] = x[q];
    l_connect[i] = i2;
  }
  fclose(f);

  const int n = nx * ny, nc = n * comp;
  or_pop *p = (or_pop *)calloc(1, sizeof(or_pop));
  p->type = OR_GR; p->n = n; p->nx = nx; p->ny = ny; p->comp = comp; p->nc = nc;
  p->ncond = N_COND_GR; p->nion = N_ION_GR;
  p->v = dal(nc); p->ca = dal(nc); p->cm = dal(nc); p->area = dal(nc);
  p->iext = dal(nc); p->shell = dal(nc); p->rev = dal(nc);
  p->compart = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < N_COND_GR; i++) p->cond[i] = dal(nc);
  for (int i = 0; i < N_ION_GR; i++) p->ion[i] = dal(nc);
  p->b = dal(nc); p->dammy = dal(nc);
  for (int i = 0; i < OR_NVEC; i++) p->vec[i] = dal(nc);
  p->hmax = 0.125;
  p->hmin = pow(2.0, -12.0);   /* H_MIN_GR, solve_rkc.cuh:24 */

  /* gr.cu:39-57 */
  for (int c = 0; c < n; c++)
    for (int k = 0; k < comp; k++) {
      const int i = c * comp + k;
      const double a = 2.0 * M_PI * rad[k] * len[k];
      p->area[i] = a;
      p->cm[i] = cmd[k] * a;
      p->compart[i] = compart[k];
      for (int q = 0; q < N_COND_GR; q++) p->cond[q][i] = dens[k][q] * a;
    }

  /* gr_solve.cu:37-43, then CSR (:73-133) with the RKC sign */
  double *mat = (double *)calloc((size_t)comp * comp, sizeof(double));
  for (int i = 0; i < comp; i++) {
    int d = l_connect[i];
    if (d >= 0) {
      mat[d + comp * i] = (2.0 / ((Ra[i] * len[i]) / (rad[i] * rad[i] * M_PI) + (Ra[d] * len[d]) / (rad[d] * rad[d] * M_PI)));
      mat[i + comp * d] = mat[d + comp * i];
    }
  }
  int nnz1 = 0;
  for (int r = 0; r < comp; r++)
    for (int c = 0; c < comp; c++) nnz1 += (mat[c + (size_t)comp * r] > 0.0 || r == c);
  p->nnz = nnz1 * n;
  p->row = (int *)malloc(sizeof(int) * (nc + 1));
  p->col = (int *)malloc(sizeof(int) * (p->nnz > 0 ? p->nnz : 1));
  p->val = (double *)malloc(sizeof(double) * (p->nnz > 0 ? p->nnz : 1));
  int k = 0;
  p->row[0] = 0;
  for (int cell = 0; cell < n; cell++)
    for (int r = 0; r < comp; r++) {
      double dsum = 0.0;
      for (int c = 0; c < comp; c++)
        if (c != r) dsum += mat[c + (size_t)comp * r];
      for (int c = 0; c < comp; c++) {
        const double m = mat[c + (size_t)comp * r];
        if (m > 0.0) { p->col[k] = cell * comp + c; p->val[k] = m; k++; }
        else if (c == r) { p->col[k] = cell * comp + c; p->val[k] = -dsum; k++; }
      }
      p->row[cell * comp + r + 1] = k;
    }
  free(mat);

  /* initial state, gr_ion.cu:635-672 */
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % comp == 0 && !v0_per_cell)
      init_v_rand = use_libc_rand ? (((double)rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 0.5 : 0.0;
    const double vv = v0_per_cell ? v0_per_cell[i / comp] : V_INIT_GR + 2.0 * init_v_rand;
    double **g = p->ion;
    p->v[i] = vv;
    p->ca[i] = Ca1_0_GR;
    p->iext[i] = 0.0;
    g[n_KV][i] = inf_n_KV(vv);
    g[a_KA][i] = inf_a_KA(vv);
    g[b_KA][i] = inf_b_KA(vv);
    g[ir_KIR][i] = inf_ir_KIR(vv);
    g[s_KM][i] = inf_s_KM(vv);
    g[c_KCa][i] = inf_c_KCa(vv, p->ca[i]);
    g[ch_Ca][i] = inf_ch_Ca(vv);
    g[ci_Ca][i] = inf_ci_Ca(vv);
    for (int q = o_Na; q <= i6_Na; q++) g[q][i] = 0.0;
    g[c1_Na][i] = 1.0;
  }
  /* gr_set_current (gr.cu:245-254): 0.01 nA into compartment 16 over [50, 100) ms */
  p->stim_comp = comp > 16 ? 16 : 0;
  p->stim_t0 = 50.0;
  p->stim_t1 = 100.0;
  p->stim_amp = dal(n);
  for (int c = 0; c < n; c++) p->stim_amp[c] = 0.01e-3;
  p->others[OR_ATOL] = atol;
  p->others[OR_RTOL] = rtol;
  GR_CREATE_FREE();
#undef GR_CREATE_FREE
  if (or_rkc_init(p) != 0) { or_pop_free(p); return NULL; }
  return p;
}

or_pop *or_gr_create(const char *param_file, int nx, int ny, const double *v0_per_cell, double atol, double rtol,
                     int use_libc_rand)
{
  return or_gr_create_comp(param_file, GR_COMP, nx, ny, v0_per_cell, atol, rtol, use_libc_rand);
}

/* gr_update_matrix, solve_rkc.cu:226-249 (terms in the reference's order) */
void or_gr_update_b(or_pop *p, const double *vv)
{
  double **c = p->cond, **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const double l_v = vv[id];
    p->b[id] = (+c[g_leak1][id] * (l_v - V_LEAK1_GR)
                + c[g_leak2][id] * (l_v - V_LEAK2_GR)
                + c[g_leak3][id] * (l_v - V_LEAK3_GR)
                - p->iext[id]
                + c[g_Na][id] * (l_v - V_Na_GR) * g[o_Na][id]
                + c[g_Ca][id] * (l_v - V_Ca_GR) * g[ch_Ca][id] * g[ch_Ca][id] * g[ci_Ca][id]
                + c[g_KV][id] * (l_v - V_K_GR) * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
                + c[g_KIR][id] * (l_v - V_K_GR) * g[ir_KIR][id]
                + c[g_KA][id] * (l_v - V_K_GR) * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
                + c[g_KCa][id] * (l_v - V_K_GR) * g[c_KCa][id]
                + c[g_KM][id] * (l_v - V_K_GR) * g[s_KM][id]);
  }
}

#define CNU(gt, INF, TAU) \
  g[gt][id] = (2.0 * DT * (INF) + (2.0 * (TAU) - DT) * g[gt][id]) / (2.0 * (TAU) + DT)

/* Na scheme from v_na_old to v_na_new, then the gates at (p->v + v_other) / 2
   (gr_update_ion_and_matrix, solve_rkc.cu:1111-1115, with v_na_old = elem[v],
   v_na_new = v_other = vnew; gr_solve_by_cnm, solve_cnm.cu:221-226, with
   v_na_old = v_other = v_old, v_na_new = elem[v]) */
static void gr_ion_all(or_pop *p, const double *v_na_old, const double *v_na_new, const double *v_other, double DT)
{
  double **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const int t = p->compart[id];
    if (t < GR_hill) continue;
    double x[12] = {g[c1_Na][id], g[c2_Na][id], g[c3_Na][id], g[c4_Na][id], g[c5_Na][id], g[i1_Na][id],
                    g[i2_Na][id], g[i3_Na][id], g[i4_Na][id], g[i5_Na][id], g[i6_Na][id], g[o_Na][id]};
    if (na_step(t, v_na_old[id], v_na_new[id], DT, x) != 0) continue;
    g[c1_Na][id] = x[0]; g[c2_Na][id] = x[1]; g[c3_Na][id] = x[2]; g[c4_Na][id] = x[3]; g[c5_Na][id] = x[4];
    g[i1_Na][id] = x[5]; g[i2_Na][id] = x[6]; g[i3_Na][id] = x[7]; g[i4_Na][id] = x[8]; g[i5_Na][id] = x[9];
    g[i6_Na][id] = x[10]; g[o_Na][id] = x[11];
  }
  for (int id = 0; id < p->nc; id++) {
    const double v_val = (p->v[id] + v_other[id]) / 2.0;
    const double Ca_val = p->ca[id];
    CNU(n_KV, inf_n_KV(v_val), tau_n_KV(v_val));
    CNU(a_KA, inf_a_KA(v_val), tau_a_KA(v_val));
    CNU(b_KA, inf_b_KA(v_val), tau_b_KA(v_val));
    CNU(ir_KIR, inf_ir_KIR(v_val), tau_ir_KIR(v_val));
    CNU(s_KM, inf_s_KM(v_val), tau_s_KM(v_val));
    CNU(c_KCa, inf_c_KCa(v_val, Ca_val), tau_c_KCa(v_val, Ca_val));
  }
}

/* gr_Na_update_2order (gr_ion.cu:45-207) then gr_update_ion_RKC_exp_imp
   (gr_ion.cu:461-480), as gr_update_ion_and_matrix runs them
   (solve_rkc.cu:1111-1115) */
void or_gr_update_ion(or_pop *p, const double *vnew, double DT)
{
  gr_ion_all(p, p->v, vnew, vnew, DT);
  for (int id = 0; id < p->nc; id++) p->v[id] = vnew[id];
}

/* the CN path: gr_Na_update_2order from v_old to v, then gr_update_ion_exp_imp
   (gr_ion.cu:438-459) */
void or_gr_cn_ion(or_pop *p, double DT)
{
  gr_ion_all(p, p->cn_vold, p->v, p->cn_vold, DT);
  for (int id = 0; id < p->nc; id++) p->cn_vold[id] = p->v[id];
}

/* conductance and reversal sums of the CN path before halving
   (gr_cnm_vec_initialize / gr_update_matrix, solve_cnm.cu:138-161, 179-200) */
void or_gr_cn_ge(const or_pop *p, int id, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = (c[g_leak1][id] + c[g_leak2][id] + c[g_leak3][id]
           + c[g_Na][id] * g[o_Na][id]
           + c[g_Ca][id] * g[ch_Ca][id] * g[ch_Ca][id] * g[ci_Ca][id]
           + c[g_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KIR][id] * g[ir_KIR][id]
           + c[g_KA][id] * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KCa][id] * g[c_KCa][id]
           + c[g_KM][id] * g[s_KM][id]);
  *esum = (c[g_leak1][id] * V_LEAK1_GR + c[g_leak2][id] * V_LEAK2_GR + c[g_leak3][id] * V_LEAK3_GR + p->iext[id]
           + c[g_Na][id] * V_Na_GR * g[o_Na][id]
           + c[g_Ca][id] * V_Ca_GR * g[ch_Ca][id] * g[ch_Ca][id] * g[ci_Ca][id]
           + c[g_KV][id] * V_K_GR * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KIR][id] * V_K_GR * g[ir_KIR][id]
           + c[g_KA][id] * V_K_GR * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KCa][id] * V_K_GR * g[c_KCa][id]
           + c[g_KM][id] * V_K_GR * g[s_KM][id]);
}

/* ---------------- the backward-Euler (BE) path ---------------- */
#define EEU(gt, INF, TAU) \
  do { const double inf_ = (INF); g[gt][id] = inf_ + (g[gt][id] - inf_) * exp(-DT / (TAU)); } while (0)

/* gr_Na_update (gr_ion.cu:209-357) then gr_update_ion (gr_ion.cu:579-600),
   both at the step's starting voltage, as gr_solve_by_bem runs them
   (solve_bem.cu:298-304); the Ca gates and Ca stay as they are */
void or_gr_be_ion(or_pop *p, double DT)
{
  double **g = p->ion;
  for (int id = 0; id < p->nc; id++) {
    const int t = p->compart[id];
    if (t < GR_hill) continue;
    double x[12] = {g[c1_Na][id], g[c2_Na][id], g[c3_Na][id], g[c4_Na][id], g[c5_Na][id], g[i1_Na][id],
                    g[i2_Na][id], g[i3_Na][id], g[i4_Na][id], g[i5_Na][id], g[i6_Na][id], g[o_Na][id]};
    if (na_be(t, p->v[id], DT, x) != 0) continue;
    g[c1_Na][id] = x[0]; g[c2_Na][id] = x[1]; g[c3_Na][id] = x[2]; g[c4_Na][id] = x[3]; g[c5_Na][id] = x[4];
    g[i1_Na][id] = x[5]; g[i2_Na][id] = x[6]; g[i3_Na][id] = x[7]; g[i4_Na][id] = x[8]; g[i5_Na][id] = x[9];
    g[i6_Na][id] = x[10]; g[o_Na][id] = x[11];
  }
  for (int id = 0; id < p->nc; id++) {
    const double v_val = p->v[id], Ca_val = p->ca[id];
    EEU(n_KV, inf_n_KV(v_val), tau_n_KV(v_val));
    EEU(a_KA, inf_a_KA(v_val), tau_a_KA(v_val));
    EEU(b_KA, inf_b_KA(v_val), tau_b_KA(v_val));
    EEU(ir_KIR, inf_ir_KIR(v_val), tau_ir_KIR(v_val));
    EEU(s_KM, inf_s_KM(v_val), tau_s_KM(v_val));
    EEU(c_KCa, inf_c_KCa(v_val, Ca_val), tau_c_KCa(v_val, Ca_val));
  }
}

/* the BE system's diagonal addend and right-hand side, Cm/DT first, in the
   reference's term order (gr_update_matrix, solve_bem.cu:258-291) */
void or_gr_be_ge(const or_pop *p, int id, double DT, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  *gsum = ((p->cm[id] / DT)
           + c[g_leak1][id] + c[g_leak2][id] + c[g_leak3][id]
           + c[g_Na][id] * g[o_Na][id]
           + c[g_Ca][id] * g[ch_Ca][id] * g[ch_Ca][id] * g[ci_Ca][id]
           + c[g_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KIR][id] * g[ir_KIR][id]
           + c[g_KA][id] * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KCa][id] * g[c_KCa][id]
           + c[g_KM][id] * g[s_KM][id]);
  *esum = ((p->cm[id] / DT) * p->v[id]
           + c[g_leak1][id] * V_LEAK1_GR + c[g_leak2][id] * V_LEAK2_GR + c[g_leak3][id] * V_LEAK3_GR + p->iext[id]
           + c[g_Na][id] * V_Na_GR * g[o_Na][id]
           + c[g_Ca][id] * V_Ca_GR * g[ch_Ca][id] * g[ch_Ca][id] * g[ci_Ca][id]
           + c[g_KV][id] * V_K_GR * g[n_KV][id] * g[n_KV][id] * g[n_KV][id] * g[n_KV][id]
           + c[g_KIR][id] * V_K_GR * g[ir_KIR][id]
           + c[g_KA][id] * V_K_GR * g[a_KA][id] * g[a_KA][id] * g[a_KA][id] * g[b_KA][id]
           + c[g_KCa][id] * V_K_GR * g[c_KCa][id]
           + c[g_KM][id] * V_K_GR * g[s_KM][id]);
}
