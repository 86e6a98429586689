/*
 * Oracle: Purkinje cell model on the RKC path (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Restates:
 *   constants               pkj.cuh:11-82
 *   initialize_host_pkj     pkj.cu:3-169   (areas, Cm, shell, conductances)
 *   set_host_pkj_solve      pkj_solve.cu:3-254 (cable CSR, gap junctions)
 *   rate functions          pkj_ion.cu:4-91
 *   pkj_initialize_ion      pkj_ion.cu:360-395
 *   pkj_update_ion_RKC      pkj_ion.cu:304-358
 *   pkj_update_matrix       solve_rkc.cu:369-395
 *   pkj_update_ion_2nd      pkj_ion.cu:246-302 (CN path)
 *   CN conductance sums     solve_cnm.cu:420-512
 *
 * Reference quirks kept: the fixed somatic Ca reversal is applied where the
 * GLOBAL compartment index is a multiple of 1599 (pkj_ion.cu:329-331), not on
 * the soma; the Ca used by the KC/K2 gates is quantised
 * (pkj_ion.cu:335).
 *
 * Morphology file: the Pkj.txt format (idx parent x y z diam type), one block
 * of `comp` lines per cell.  A file holding exactly one block is replicated to
 * every cell, as the reference does (pkj.cu:143-168); a file with n blocks
 * gives each cell its own tree.
 *
 * Pinning: rate functions against alpha/(alpha+beta) identities, singular
 * points and hand-evaluated values; the CN gate update against the (1,1) Pade
 * form of exp(-dt/tau); the Ca RK4 against the exact exponential solution;
 * the cable matrix against a dense brute force and charge conservation
 * (tests/test_oracle_pkj.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PKJ_SOMA 0
#define PKJ_MAIN 1
#define PKJ_SMOOTH 2
#define PKJ_SPINY 3

#define V_INIT_PKJ (-68.0)
#define V_LEAK_PKJ (-80.0)
#define V_Na_PKJ (45.0)
#define V_Ca_PKJ (135.0)
#define V_KH_PKJ (-30.0)
#define V_K_PKJ (-85.0)
#define G_GJ_PKJ (2.5e-6)
#define GJ_EACH_PKJ 26

#define PKJ_SPINE_DENS (13.0e4)
#define PKJ_SPINE_AREA (1.33e-8)
#define PKJ_CM (1.64)
#define PKJ_RM_S (10.0)
#define PKJ_RM_D (30.0)
#define PKJ_RI (0.25)
#define B_Ca1_PKJ (10.0)
#define PKJ_SHELLD (0.2e-4)
#define Ca1OUT_PKJ (2.4)
#define Ca1_0_PKJ (4.0e-5)
#define F_PKJ (9.6485e4)

/* gate order, pkj.cuh:86-87 */
enum { m_NaF, h_NaF, m_NaP, m_CaP, h_CaP, m_CaT, h_CaT, m_KA, h_KA, m_KC, m_Kh1, m_Kh2,
       m_Kdr, h_Kdr, m_KM, z_KC, m_K2, z_K2, N_ION };
/* conductance order, pkj.cuh:90-91 */
enum { g_leak, g_NaF, g_NaP, g_CaP, g_CaT, g_Kh, g_Kdr, g_KM, g_KA, g_KC, g_K2, N_COND };

/* densities [mS/cm^2] per compartment type, pkj.cuh:28-62; order NaF NaP CaP CaT KH KDR KM KA KC K2 */
static const double DENS[3][10] = {
  {7500.0, 1.0, 0.0, 0.5, 0.3, 600.0, 0.040, 15.0, 0.0, 0.0},     /* soma */
  {0.0, 0.0, 4.5, 0.5, 0.0, 60.0, 0.01, 2.0, 80.0, 0.39},        /* main dendrite */
  {0.0, 0.0, 4.5, 0.5, 0.0, 0.0, 0.013, 0.0, 80.0, 0.39},        /* rest of dendrite */
};

/* ---------------- rate functions, pkj_ion.cu:4-91 ---------------- */
static double lf(double v, double A, double B, double C, double D) { return A / (B + exp((v + C) / D)); }
static double a_m_NaF(double v) { return lf(v, 35.0, 0.0, 5.0, -10.0); }
static double a_h_NaF(double v) { return lf(v, 0.225, 1.0, 80.0, 10.0); }
static double a_m_NaP(double v) { return lf(v, 200.0, 1.0, -18.0, -16.0); }
static double a_m_CaP(double v) { return lf(v, 8.5, 1.0, -8.0, -12.5); }
static double a_h_CaP(double v) { return lf(v, 0.0015, 1.0, 29.0, 8.0); }
static double a_m_CaT(double v) { return lf(v, 2.60, 1.0, 21.0, -8.0); }
static double a_h_CaT(double v) { return lf(v, 0.0025, 1.0, 40.0, 8.0); }
static double a_m_KA(double v) { return lf(v, 1.40, 1.0, 27.0, -12.0); }
static double a_h_KA(double v) { return lf(v, 0.0175, 1.0, 50.0, 8.0); }
static double a_m_KC(double v) { (void)v; return 7.5; }
static double a_m_K2(double v) { (void)v; return 25.0; }
static double b_m_NaF(double v) { return lf(v, 7.0, 0.0, 65.0, 20.0); }
static double b_h_NaF(double v) { return lf(v, 7.5, 0.0, -3.0, -18.0); }
static double b_m_NaP(double v) { return lf(v, 25.0, 1.0, 58.0, 8.0); }
static double b_m_CaP(double v) { return lf(v, 35.0, 1.0, 74.0, 14.5); }
static double b_h_CaP(double v) { return lf(v, 0.0055, 1.0, 23.0, -8.0); }
static double b_m_CaT(double v) { return lf(v, 0.180, 1.0, 40.0, 4.0); }
static double b_h_CaT(double v) { return lf(v, 0.190, 1.0, 50.0, -10.0); }
static double b_m_KA(double v) { return lf(v, 0.490, 1.0, 30.0, 4.0); }
static double b_h_KA(double v) { return lf(v, 1.30, 1.0, 13.0, -10.0); }
static double b_m_KC(double v) { return lf(v, 0.110, 0.0, -35.0, 14.9); }
static double b_m_K2(double v) { return lf(v, 0.075, 0.0, 25.0, 6.0); }

#define TAU_INF(name)                                                                        \
  static double tau_##name(double v) { return 1.0 / (a_##name(v) + b_##name(v)); }          \
  static double inf_##name(double v) { return a_##name(v) / (a_##name(v) + b_##name(v)); }
TAU_INF(m_NaF) TAU_INF(h_NaF) TAU_INF(m_NaP) TAU_INF(m_CaP) TAU_INF(h_CaP) TAU_INF(m_CaT)
TAU_INF(h_CaT) TAU_INF(m_KA) TAU_INF(h_KA) TAU_INF(m_KC) TAU_INF(m_K2)

static double tau_m_Kh1(double v) { (void)v; return 7.6; }
static double tau_m_Kh2(double v) { (void)v; return 36.8; }
static double tau_m_Kdr(double v)
{
  double alpha;
  (fabs(v + 12.0) > 1e-6) ? (alpha = -0.0235 * (v + 12.0) / (exp(-(v + 12.0) / 12.0) - 1.0))
                          : (alpha = 0.0235 * 12.0);
  double beta = 5.0 * exp(-(v + 147.0) / 30.0);
  return 1.0 / (alpha + beta);
}
static double tau_h_Kdr(double v) { return (v < -25.0) ? 1200.0 : 10.0; }
static double tau_m_KM(double v) { return 200.0 / (3.3 * (exp((v + 35.0) / 20.0)) + exp(-(v + 35.0) / 20.0)); }
static double tau_z_KC(double v) { (void)v; return 10.0; }
static double tau_z_K2(double v) { (void)v; return 10.0; }
static double inf_m_Kh1(double v) { return 0.8 / (1.0 + exp((v + 82.0) / 7.0)); }
static double inf_m_Kh2(double v) { return 0.2 / (1.0 + exp((v + 82.0) / 7.0)); }
static double inf_m_Kdr(double v)
{
  double alpha;
  (fabs(v - 8.0) > 1e-6) ? (alpha = -0.0235 * (v - 8.0) / (exp(-(v - 8.0) / 12.0) - 1.0))
                         : (alpha = 0.0235 * 12.0);
  double beta = 5.0 * exp(-(v + 127.0) / 30.0);
  return alpha / (alpha + beta);
}
static double inf_h_Kdr(double v) { return 1.0 / (1.0 + exp((v + 25.0) / 4.0)); }
static double inf_m_KM(double v) { return 1.0 / (1.0 + exp(-(v + 35.0) / 10.0)); }
static double inf_z_KC(double v, double Ca) { (void)v; return 1.0 / (1.0 + 4.00 / Ca); }
static double inf_z_K2(double v, double Ca) { (void)v; return 1.0 / (1.0 + 0.20 / Ca); }

/* exported for the closed-form tests: which = 2*gate (+1 for tau) */
double or_pkj_rate(int which, double v, double ca)
{
  int g = which / 2, t = which % 2;
  switch (g) {
    case m_NaF: return t ? tau_m_NaF(v) : inf_m_NaF(v);
    case h_NaF: return t ? tau_h_NaF(v) : inf_h_NaF(v);
    case m_NaP: return t ? tau_m_NaP(v) : inf_m_NaP(v);
    case m_CaP: return t ? tau_m_CaP(v) : inf_m_CaP(v);
    case h_CaP: return t ? tau_h_CaP(v) : inf_h_CaP(v);
    case m_CaT: return t ? tau_m_CaT(v) : inf_m_CaT(v);
    case h_CaT: return t ? tau_h_CaT(v) : inf_h_CaT(v);
    case m_KA: return t ? tau_m_KA(v) : inf_m_KA(v);
    case h_KA: return t ? tau_h_KA(v) : inf_h_KA(v);
    case m_KC: return t ? tau_m_KC(v) : inf_m_KC(v);
    case m_Kh1: return t ? tau_m_Kh1(v) : inf_m_Kh1(v);
    case m_Kh2: return t ? tau_m_Kh2(v) : inf_m_Kh2(v);
    case m_Kdr: return t ? tau_m_Kdr(v) : inf_m_Kdr(v);
    case h_Kdr: return t ? tau_h_Kdr(v) : inf_h_Kdr(v);
    case m_KM: return t ? tau_m_KM(v) : inf_m_KM(v);
    case z_KC: return t ? tau_z_KC(v) : inf_z_KC(v, ca);
    case m_K2: return t ? tau_m_K2(v) : inf_m_K2(v);
    case z_K2: return t ? tau_z_K2(v) : inf_z_K2(v, ca);
  }
  return NAN;
}

/* ---------------- morphology reading ---------------- */
typedef struct { int idx, parent, type; double x, y, z, d; } mrow;

static mrow *read_morph(const char *path, int comp, int *nblocks)
{
  FILE *f = fopen(path, "r");
  if (!f) { fprintf(stderr, "oracle: no such file %s\n", path); return NULL; }
  int cap = comp, n = 0;
  mrow *r = (mrow *)malloc(sizeof(mrow) * cap);
  while (1) {
    mrow q;
    if (fscanf(f, "%d %d %lf %lf %lf %lf %d ", &q.idx, &q.parent, &q.x, &q.y, &q.z, &q.d, &q.type) != 7) break;
    if (n == cap) { cap *= 2; r = (mrow *)realloc(r, sizeof(mrow) * cap); }
    r[n++] = q;
  }
  fclose(f);
  if (n == 0 || n % comp != 0) { free(r); return NULL; }
  *nblocks = n / comp;
  return r;
}

/* ---------------- creation ---------------- */
static double *dalloc(int n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

or_pop *or_pkj_create(const char *morph_file, const char *gj_file, int nx, int ny,
                      const double *v0_per_cell, double atol, double rtol, int use_libc_rand)
{
  const int COMP = 1600;
  int nblocks = 0;
  mrow *mr = read_morph(morph_file, COMP, &nblocks);
  if (!mr) return NULL;
  const int n = nx * ny, nc = n * COMP;
  if (nblocks != 1 && nblocks != n) { free(mr); return NULL; }

  or_pop *p = (or_pop *)calloc(1, sizeof(or_pop));
  p->type = OR_PKJ; p->n = n; p->nx = nx; p->ny = ny; p->comp = COMP; p->nc = nc;
  p->ncond = N_COND; p->nion = N_ION;
  p->v = dalloc(nc); p->ca = dalloc(nc); p->cm = dalloc(nc); p->area = dalloc(nc);
  p->iext = dalloc(nc); p->shell = dalloc(nc); p->rev = dalloc(nc);
  p->compart = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < N_COND; i++) p->cond[i] = dalloc(nc);
  for (int i = 0; i < N_ION; i++) p->ion[i] = dalloc(nc);
  p->b = dalloc(nc); p->dammy = dalloc(nc);
  for (int i = 0; i < OR_NVEC; i++) p->vec[i] = dalloc(nc);
  p->hmax = 0.125;
  p->hmin = pow(2.0, -10.0);

  /* per-cell geometry and parameters, pkj.cu:30-140 */
  int *parent = (int *)malloc(sizeof(int) * nc);
  double *rad = dalloc(nc), *len = dalloc(nc);
  for (int c = 0; c < n; c++) {
    const mrow *blk = mr + (nblocks == 1 ? 0 : c * COMP);
    for (int k = 0; k < COMP; k++) {
      const mrow *q = &blk[k];
      int i1 = q->idx, i3 = q->type;
      int gi = c * COMP + i1;
      double d = q->d, area_val;
      parent[gi] = q->parent;
      if (i3 == PKJ_SOMA) {
        rad[gi] = 0.5 * d * 1e-4;
        len[gi] = d * 1e-4;
        p->compart[gi] = i3;
        p->area[gi] = 4.0 * M_PI * rad[gi] * rad[gi];
        area_val = p->area[gi];
        p->cm[gi] = PKJ_CM * area_val;
        p->shell[gi] = (PKJ_SHELLD - 2 * pow(PKJ_SHELLD, 2) / (d * 1e-4)
                        + 4 * pow(PKJ_SHELLD, 3) / (3 * pow(d * 1e-4, 2))) * p->area[gi];
      } else {
        rad[gi] = 0.5 * d * 1e-4;
        len[gi] = sqrt(pow(q->x, 2) + pow(q->y, 2) + pow(q->z, 2)) * 1e-4;
        p->compart[gi] = i3;
        p->area[gi] = 2.0 * M_PI * rad[gi] * len[gi];
        area_val = p->area[gi];
        p->shell[gi] = (PKJ_SHELLD - pow(PKJ_SHELLD, 2) / (d * 1e-4)) * p->area[gi];
        if (i3 == PKJ_SPINY) {
          double other_spines = area_val + len[gi] * PKJ_SPINE_AREA * PKJ_SPINE_DENS;
          p->cm[gi] = PKJ_CM * other_spines;
        } else {
          p->cm[gi] = PKJ_CM * area_val;
        }
      }
      p->iext[gi] = 0.0;
      int row = (i3 == PKJ_SOMA) ? 0 : (i3 == PKJ_MAIN) ? 1 : 2;
      double leak_area = area_val;
      if (i3 == PKJ_SPINY) leak_area = area_val + len[gi] * PKJ_SPINE_AREA * PKJ_SPINE_DENS;
      p->cond[g_leak][gi] = 1.0 / ((i3 == PKJ_SOMA) ? PKJ_RM_S : PKJ_RM_D) * leak_area;
      for (int g = 0; g < 10; g++) p->cond[g_NaF + g][gi] = DENS[row][g] * area_val;
    }
  }

  /* cable matrix, pkj_solve.cu:68-74 (symmetric g between child and parent) */
  /* adjacency per compartment as (col, g) lists, then CSR sorted by column */
  int n_gj = 0;
  int *gj_pre = NULL, *gj_post = NULL;
  if (n > 1 && gj_file) {
    FILE *fg = fopen(gj_file, "r");
    if (!fg) { fprintf(stderr, "oracle: no such file %s\n", gj_file); }
    else {
      int lp[GJ_EACH_PKJ], lq[GJ_EACH_PKJ], ok = 1;
      for (int i = 0; i < GJ_EACH_PKJ; i++) {
        if (fscanf(fg, "%d, %d", &lp[i], &lq[i]) != 2) { ok = 0; break; }
      }
      fclose(fg);
      if (ok) {
        int npairs = GJ_EACH_PKJ * (nx - 1) * ny;   /* pkj_solve.cu:81, 100-116 */
        n_gj = 2 * npairs;
        gj_pre = (int *)malloc(sizeof(int) * n_gj);
        gj_post = (int *)malloc(sizeof(int) * n_gj);
        int c_gj = 0;
        for (int y = 0; y < ny; y++)
          for (int x = 0; x < nx - 1; x++)
            for (int i = 0; i < GJ_EACH_PKJ; i++) {
              int a = lp[i] + ((x) + y * nx) * COMP;
              int b = lq[i] + ((x + 1) + y * nx) * COMP;
              gj_pre[2 * c_gj] = a; gj_pre[2 * c_gj + 1] = b;
              gj_post[2 * c_gj] = b; gj_post[2 * c_gj + 1] = a;
              c_gj++;
            }
      }
    }
  }

  int *deg = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < nc; i++) {
    int c = i / COMP, pl = parent[i];
    if (pl >= 0) { deg[i]++; deg[c * COMP + pl]++; }
  }
  for (int i = 0; i < n_gj; i++) deg[gj_pre[i]]++;
  p->row = (int *)malloc(sizeof(int) * (nc + 1));
  p->row[0] = 0;
  for (int i = 0; i < nc; i++) p->row[i + 1] = p->row[i] + deg[i] + 1;  /* +1: stored diagonal */
  p->nnz = p->row[nc];
  p->col = (int *)malloc(sizeof(int) * p->nnz);
  p->val = (double *)malloc(sizeof(double) * p->nnz);
  int *fill = (int *)calloc(nc, sizeof(int));
  for (int i = 0; i < nc; i++) {
    int c = i / COMP, pl = parent[i];
    if (pl >= 0) {
      int d = c * COMP + pl;
      /* pkj_solve.cu:71: g = 2 / (Ra_i len_i/(r_i^2 pi) + Ra_d len_d/(r_d^2 pi)) [mS] */
      double g = (2.0 / ((PKJ_RI * len[i]) / (rad[i] * rad[i] * M_PI) + (PKJ_RI * len[d]) / (rad[d] * rad[d] * M_PI)));
      int k = p->row[i] + fill[i]++;
      p->col[k] = d; p->val[k] = g;
      k = p->row[d] + fill[d]++;
      p->col[k] = i; p->val[k] = g;
    }
  }
  for (int i = 0; i < n_gj; i++) {
    int a = gj_pre[i];
    int k = p->row[a] + fill[a]++;
    p->col[k] = gj_post[i]; p->val[k] = G_GJ_PKJ;
  }
  for (int i = 0; i < nc; i++) {
    int k = p->row[i] + fill[i]++;
    p->col[k] = i; p->val[k] = 0.0;
    /* sort the row by column (insertion sort; rows are short) */
    int s = p->row[i], e = p->row[i + 1];
    for (int a = s + 1; a < e; a++) {
      int cc = p->col[a]; double vv = p->val[a]; int b = a - 1;
      while (b >= s && p->col[b] > cc) { p->col[b + 1] = p->col[b]; p->val[b + 1] = p->val[b]; b--; }
      p->col[b + 1] = cc; p->val[b + 1] = vv;
    }
    /* diagonal = -(sum of off-diagonal conductances), after the RKC sign flip (pkj_solve.cu:159-181, 238-242).
       Summed in column order, as the dense-row scan does. */
    double dsum = 0.0; int kd = -1, ngj_here = 0;
    for (int k = s; k < e; k++) {
      if (p->col[k] == i) kd = k;
      else if (p->col[k] / COMP != i / COMP) ngj_here++;
      else dsum += p->val[k];
    }
    for (int k = 0; k < ngj_here; k++) dsum += G_GJ_PKJ;
    p->val[kd] = -dsum;
  }
  free(fill); free(deg); free(gj_pre); free(gj_post);
  free(parent); free(rad); free(len); free(mr);

  /* initial state, pkj_ion.cu:360-395 */
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % COMP == 0 && !v0_per_cell) {
      if (use_libc_rand) init_v_rand = (((double)rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 0.5;
      else init_v_rand = 0.0;
    }
    p->v[i] = v0_per_cell ? v0_per_cell[i / COMP] : V_INIT_PKJ + 2.0 * init_v_rand;
    p->ca[i] = Ca1_0_PKJ;
    p->rev[i] = V_Ca_PKJ;
    p->iext[i] = 0.0;
    double vv = p->v[i], cv = p->ca[i];
    p->ion[m_NaF][i] = inf_m_NaF(vv);
    p->ion[h_NaF][i] = inf_h_NaF(vv);
    p->ion[m_NaP][i] = inf_m_NaP(vv);
    p->ion[m_CaP][i] = inf_m_CaP(vv);
    p->ion[h_CaP][i] = inf_h_CaP(vv);
    p->ion[m_CaT][i] = inf_m_CaT(vv);
    p->ion[h_CaT][i] = inf_h_CaT(vv);
    p->ion[m_Kh1][i] = inf_m_Kh1(vv);
    p->ion[m_Kh2][i] = inf_m_Kh2(vv);
    p->ion[m_Kdr][i] = inf_m_Kdr(vv);
    p->ion[h_Kdr][i] = inf_h_Kdr(vv);
    p->ion[m_KM][i] = inf_m_KM(vv);
    p->ion[m_KA][i] = inf_m_KA(vv);
    p->ion[h_KA][i] = inf_h_KA(vv);
    p->ion[m_KC][i] = inf_m_KC(vv);
    p->ion[z_KC][i] = inf_z_KC(vv, cv);
    p->ion[m_K2][i] = inf_m_K2(vv);
    p->ion[z_K2][i] = inf_z_K2(vv, cv);
  }
  p->others[OR_ATOL] = atol;
  p->others[OR_RTOL] = rtol;
  p->stim_comp = 1599;
  if (or_rkc_init(p) != 0) { or_pop_free(p); return NULL; }
  return p;
}

/* ---------------- pkj_update_matrix, solve_rkc.cu:369-395 ---------------- */
void or_pkj_update_b(or_pop *p, const double *vv)
{
  double **cond = p->cond, **ion = p->ion;
#pragma omp parallel for schedule(static)
  for (int id = 0; id < p->nc; id++) {
    double v_ca2_pkj = p->rev[id];
    double l_v = vv[id];
    p->b[id] = (-p->iext[id]
      + cond[g_leak][id] * (l_v - V_LEAK_PKJ)
      + cond[g_NaF][id] * ion[m_NaF][id] * ion[m_NaF][id] * ion[m_NaF][id] * ion[h_NaF][id] * (l_v - V_Na_PKJ)
      + cond[g_NaP][id] * ion[m_NaP][id] * ion[m_NaP][id] * ion[m_NaP][id] * (l_v - V_Na_PKJ)
      + cond[g_CaP][id] * ion[m_CaP][id] * ion[h_CaP][id] * (l_v - v_ca2_pkj)
      + cond[g_CaT][id] * ion[m_CaT][id] * ion[h_CaT][id] * (l_v - v_ca2_pkj)
      + cond[g_Kh][id] * ion[m_Kh1][id] * (l_v - V_KH_PKJ)
      + cond[g_Kh][id] * ion[m_Kh2][id] * (l_v - V_KH_PKJ)
      + cond[g_Kdr][id] * ion[m_Kdr][id] * ion[m_Kdr][id] * ion[h_Kdr][id] * (l_v - V_K_PKJ)
      + cond[g_KM][id] * ion[m_KM][id] * (l_v - V_K_PKJ)
      + cond[g_KA][id] * ion[m_KA][id] * ion[m_KA][id] * ion[m_KA][id] * ion[m_KA][id] * ion[h_KA][id] * (l_v - V_K_PKJ)
      + cond[g_KC][id] * ion[m_KC][id] * ion[z_KC][id] * ion[z_KC][id] * (l_v - V_K_PKJ)
      + cond[g_K2][id] * ion[m_K2][id] * ion[z_K2][id] * ion[z_K2][id] * (l_v - V_K_PKJ));
  }
}

/* ---------------- pkj_update_ion_RKC, pkj_ion.cu:304-358 ---------------- */
#define CNU(g, INF, TAU) \
  ion[g][id] = (2.0 * DT * (INF) + (2.0 * (TAU) - DT) * ion[g][id]) / (2.0 * (TAU) + DT)

/* Ca shell and gates of one compartment at the gate voltage v_val; the Ca
   current uses the stored voltage p->v (pkj_ion.cu:255-297 and :312-353
   differ only in v_val) */
static void pkj_ion_at(or_pop *p, int id, double v_val, double DT)
{
  double **ion = p->ion;
  double Ca_val = p->ca[id];
  double I_Ca1 = 1e-3 * (p->v[id] - p->rev[id])
                 * (p->cond[g_CaP][id] * ion[m_CaP][id] * ion[h_CaP][id]
                    + p->cond[g_CaT][id] * ion[m_CaT][id] * ion[h_CaT][id]);
  (I_Ca1 > 0.0) ? (I_Ca1 = 0.0) : (I_Ca1 = -I_Ca1 / (2.0 * F_PKJ * p->shell[id]));
  double k1 = DT * (I_Ca1 - B_Ca1_PKJ * (Ca_val - Ca1_0_PKJ));
  double k2 = DT * (I_Ca1 - B_Ca1_PKJ * (Ca_val + k1 / 2.0 - Ca1_0_PKJ));
  double k3 = DT * (I_Ca1 - B_Ca1_PKJ * (Ca_val + k2 / 2.0 - Ca1_0_PKJ));
  double k4 = DT * (I_Ca1 - B_Ca1_PKJ * (Ca_val + k3 - Ca1_0_PKJ));
  k1 = Ca_val + (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
  p->ca[id] = k1;
  Ca_val = 0.5 * (Ca_val + k1);
  ((id + p->gid_offset) % 1599 == 0) ? (p->rev[id] = 12.5 * log(Ca1OUT_PKJ / Ca1_0_PKJ))
                   : (p->rev[id] = 13.361624877 * log(Ca1OUT_PKJ / Ca_val));
  Ca_val = floor((Ca_val * 1000 - 0.04) / 0.099986667) * 0.099986667 + 0.04;
  CNU(m_NaF, inf_m_NaF(v_val), tau_m_NaF(v_val));
  CNU(h_NaF, inf_h_NaF(v_val), tau_h_NaF(v_val));
  CNU(m_NaP, inf_m_NaP(v_val), tau_m_NaP(v_val));
  CNU(m_CaP, inf_m_CaP(v_val), tau_m_CaP(v_val));
  CNU(h_CaP, inf_h_CaP(v_val), tau_h_CaP(v_val));
  CNU(m_CaT, inf_m_CaT(v_val), tau_m_CaT(v_val));
  CNU(h_CaT, inf_h_CaT(v_val), tau_h_CaT(v_val));
  CNU(m_Kh1, inf_m_Kh1(v_val), tau_m_Kh1(v_val));
  CNU(m_Kh2, inf_m_Kh2(v_val), tau_m_Kh2(v_val));
  CNU(m_Kdr, inf_m_Kdr(v_val), tau_m_Kdr(v_val));
  CNU(h_Kdr, inf_h_Kdr(v_val), tau_h_Kdr(v_val));
  CNU(m_KM, inf_m_KM(v_val), tau_m_KM(v_val));
  CNU(m_KA, inf_m_KA(v_val), tau_m_KA(v_val));
  CNU(h_KA, inf_h_KA(v_val), tau_h_KA(v_val));
  CNU(m_KC, inf_m_KC(v_val), tau_m_KC(v_val));
  CNU(z_KC, inf_z_KC(v_val, Ca_val), tau_z_KC(v_val));
  CNU(m_K2, inf_m_K2(v_val), tau_m_K2(v_val));
  CNU(z_K2, inf_z_K2(v_val, Ca_val), tau_z_K2(v_val));
}

/* pkj_update_ion_RKC, pkj_ion.cu:304-358: gates at the mid-step voltage */
void or_pkj_update_ion(or_pop *p, const double *elem_vnew, double DT)
{
#pragma omp parallel for schedule(static)
  for (int id = 0; id < p->nc; id++) {
    pkj_ion_at(p, id, 0.5 * (elem_vnew[id] + p->v[id]), DT);
    p->v[id] = elem_vnew[id];
  }
}

/* pkj_update_ion_2nd, pkj_ion.cu:246-302 (CN path): gates at the voltage
   extrapolated from the last two steps */
void or_pkj_cn_ion(or_pop *p, double DT)
{
#pragma omp parallel for schedule(static)
  for (int id = 0; id < p->nc; id++) {
    pkj_ion_at(p, id, (1.5 * p->v[id] - 0.5 * p->cn_vold[id]), DT);
    p->cn_vold[id] = p->v[id];
  }
}

/* the ionic conductance sum and the conductance-weighted reversal sum plus
   the injected current, before halving (pkj_cnm_vec_initialize /
   pkj_update_matrix, solve_cnm.cu:429-456, 480-509) */
void or_pkj_cn_ge(const or_pop *p, int id, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  const double l_v_Ca = p->rev[id];
  *gsum = (+c[g_leak][id]
           + c[g_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[h_NaF][id]
           + c[g_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * g[m_NaP][id]
           + c[g_CaP][id] * g[m_CaP][id] * g[h_CaP][id]
           + c[g_CaT][id] * g[m_CaT][id] * g[h_CaT][id]
           + c[g_Kh][id] * g[m_Kh1][id]
           + c[g_Kh][id] * g[m_Kh2][id]
           + c[g_Kdr][id] * g[m_Kdr][id] * g[m_Kdr][id] * g[h_Kdr][id]
           + c[g_KM][id] * g[m_KM][id]
           + c[g_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[h_KA][id]
           + c[g_KC][id] * g[m_KC][id] * g[z_KC][id] * g[z_KC][id]
           + c[g_K2][id] * g[m_K2][id] * g[z_K2][id] * g[z_K2][id]);
  *esum = (+c[g_leak][id] * (V_LEAK_PKJ) + p->iext[id]
           + c[g_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[h_NaF][id] * (V_Na_PKJ)
           + c[g_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * (V_Na_PKJ)
           + c[g_CaP][id] * g[m_CaP][id] * g[h_CaP][id] * (l_v_Ca)
           + c[g_CaT][id] * g[m_CaT][id] * g[h_CaT][id] * (l_v_Ca)
           + c[g_Kh][id] * g[m_Kh1][id] * (V_KH_PKJ)
           + c[g_Kh][id] * g[m_Kh2][id] * (V_KH_PKJ)
           + c[g_Kdr][id] * g[m_Kdr][id] * g[m_Kdr][id] * g[h_Kdr][id] * (V_K_PKJ)
           + c[g_KM][id] * g[m_KM][id] * (V_K_PKJ)
           + c[g_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[h_KA][id] * (V_K_PKJ)
           + c[g_KC][id] * g[m_KC][id] * g[z_KC][id] * g[z_KC][id] * (V_K_PKJ)
           + c[g_K2][id] * g[m_K2][id] * g[z_K2][id] * g[z_K2][id] * (V_K_PKJ));
}

/* ---------------- the backward-Euler (BE) path ---------------- */
#define EEU(gt, INF, TAU) \
  do { const double inf_ = (INF); ion[gt][id] = inf_ + (ion[gt][id] - inf_) * exp(-DT / (TAU)); } while (0)

/* pkj_update_ion, pkj_ion.cu:94-151 (the BE path): every gate by exponential
   Euler at V, KC / K2 at the quantised present Ca; the Ca shell by RK4 with
   the present gates' Ca current, clamped at zero from above and divided by
   shell x area (the stored shell volume already carries the area, so the
   area enters twice, as in the reference); the reversal from the present Ca,
   fixed on compartments of the soma type. */
void or_pkj_be_ion(or_pop *p, double DT)
{
#pragma omp parallel for schedule(static)
  for (int id = 0; id < p->nc; id++) {
    double **ion = p->ion, **cond = p->cond;
    const double v_val = p->v[id];
    const double Ca_val = p->ca[id];
    const double l_v_Ca = p->rev[id];
    const double cafloor = floor((Ca_val * 1000 - 0.04) / 0.099986667) * 0.099986667 + 0.04;
    double I_Ca1 = 1e-3 * (v_val - l_v_Ca)
                   * (cond[g_CaP][id] * ion[m_CaP][id] * ion[h_CaP][id]
                      + cond[g_CaT][id] * ion[m_CaT][id] * ion[h_CaT][id]);
    if (I_Ca1 > 0.0) I_Ca1 = 0.0;
    EEU(m_NaF, inf_m_NaF(v_val), tau_m_NaF(v_val));
    EEU(h_NaF, inf_h_NaF(v_val), tau_h_NaF(v_val));
    EEU(m_NaP, inf_m_NaP(v_val), tau_m_NaP(v_val));
    EEU(m_CaP, inf_m_CaP(v_val), tau_m_CaP(v_val));
    EEU(h_CaP, inf_h_CaP(v_val), tau_h_CaP(v_val));
    EEU(m_CaT, inf_m_CaT(v_val), tau_m_CaT(v_val));
    EEU(h_CaT, inf_h_CaT(v_val), tau_h_CaT(v_val));
    EEU(m_Kh1, inf_m_Kh1(v_val), tau_m_Kh1(v_val));
    EEU(m_Kh2, inf_m_Kh2(v_val), tau_m_Kh2(v_val));
    EEU(m_Kdr, inf_m_Kdr(v_val), tau_m_Kdr(v_val));
    EEU(h_Kdr, inf_h_Kdr(v_val), tau_h_Kdr(v_val));
    EEU(m_KM, inf_m_KM(v_val), tau_m_KM(v_val));
    EEU(m_KA, inf_m_KA(v_val), tau_m_KA(v_val));
    EEU(h_KA, inf_h_KA(v_val), tau_h_KA(v_val));
    EEU(m_KC, inf_m_KC(v_val), tau_m_KC(v_val));
    EEU(z_KC, inf_z_KC(v_val, cafloor), tau_z_KC(v_val));
    EEU(m_K2, inf_m_K2(v_val), tau_m_K2(v_val));
    EEU(z_K2, inf_z_K2(v_val, cafloor), tau_z_K2(v_val));
    const double l_shell = p->shell[id] * p->area[id];
    const double k1 = DT * (-I_Ca1 / (2.0 * F_PKJ * l_shell) - B_Ca1_PKJ * (Ca_val - Ca1_0_PKJ));
    const double k2 = DT * (-I_Ca1 / (2.0 * F_PKJ * l_shell) - B_Ca1_PKJ * (Ca_val + k1 / 2.0 - Ca1_0_PKJ));
    const double k3 = DT * (-I_Ca1 / (2.0 * F_PKJ * l_shell) - B_Ca1_PKJ * (Ca_val + k2 / 2.0 - Ca1_0_PKJ));
    const double k4 = DT * (-I_Ca1 / (2.0 * F_PKJ * l_shell) - B_Ca1_PKJ * (Ca_val + k3 - Ca1_0_PKJ));
    p->ca[id] += (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
    p->rev[id] = (p->compart[id] == PKJ_SOMA) ? 12.5 * log(Ca1OUT_PKJ / Ca1_0_PKJ)
                                              : 13.361624877 * log(Ca1OUT_PKJ / Ca_val);
  }
}

/* the BE system's diagonal addend and right-hand side, Cm/DT first, in the
   reference's term order (pkj_update_matrix, solve_bem.cu:407-445) */
void or_pkj_be_ge(const or_pop *p, int id, double DT, double *gsum, double *esum)
{
  double *const *c = p->cond, *const *g = p->ion;
  const double l_v_Ca = p->rev[id];
  *gsum = ((p->cm[id] / DT)
           + c[g_leak][id]
           + c[g_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[h_NaF][id]
           + c[g_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * g[m_NaP][id]
           + c[g_CaP][id] * g[m_CaP][id] * g[h_CaP][id]
           + c[g_CaT][id] * g[m_CaT][id] * g[h_CaT][id]
           + c[g_Kh][id] * g[m_Kh1][id]
           + c[g_Kh][id] * g[m_Kh2][id]
           + c[g_Kdr][id] * g[m_Kdr][id] * g[m_Kdr][id] * g[h_Kdr][id]
           + c[g_KM][id] * g[m_KM][id]
           + c[g_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[h_KA][id]
           + c[g_KC][id] * g[m_KC][id] * g[z_KC][id] * g[z_KC][id]
           + c[g_K2][id] * g[m_K2][id] * g[z_K2][id] * g[z_K2][id]);
  *esum = ((p->cm[id] / DT) * p->v[id] + p->iext[id]
           + c[g_leak][id] * (V_LEAK_PKJ)
           + c[g_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[m_NaF][id] * g[h_NaF][id] * (V_Na_PKJ)
           + c[g_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * g[m_NaP][id] * (V_Na_PKJ)
           + c[g_CaP][id] * g[m_CaP][id] * g[h_CaP][id] * (l_v_Ca)
           + c[g_CaT][id] * g[m_CaT][id] * g[h_CaT][id] * (l_v_Ca)
           + c[g_Kh][id] * g[m_Kh1][id] * (V_KH_PKJ)
           + c[g_Kh][id] * g[m_Kh2][id] * (V_KH_PKJ)
           + c[g_Kdr][id] * g[m_Kdr][id] * g[m_Kdr][id] * g[h_Kdr][id] * (V_K_PKJ)
           + c[g_KM][id] * g[m_KM][id] * (V_K_PKJ)
           + c[g_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[m_KA][id] * g[h_KA][id] * (V_K_PKJ)
           + c[g_KC][id] * g[m_KC][id] * g[z_KC][id] * g[z_KC][id] * (V_K_PKJ)
           + c[g_K2][id] * g[m_K2][id] * g[z_K2][id] * g[z_K2][id] * (V_K_PKJ));
}
