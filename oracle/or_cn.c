/*
 * Oracle: the Crank-Nicolson (CN) solver path (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Restates:
 *   xx_cnm_vec_initialize   solve_cnm.cu:128-162, 259-303, 420-465, 741-775
 *                           (gamma1 / omega1 = conductance and reversal sums / 2, v_old = v)
 *   xx_update_matrix        solve_cnm.cu:164-208, 305-362, 466-525, 777-821
 *                           (diagonal Cm/DT + gamma2, right-hand side from the halved cable
 *                           matrix, the previous half-step sums and the new ones)
 *   xx_solve_by_cnm         solve_cnm.cu:210-243, 364-392, 527-557, 823-852
 *                           (ion update, reset, IO gap term, matrix, linear solve)
 *   cg_cusparse_crs         solve_cnm.cu:20-98   (CG from x = 0, at most 100 iterations,
 *                                                 stop when |r| <= 1e-15 |b|)
 *   bicg_cusparse_crs       solve_cnm.cu:617-727 (BiCG with A^T, same start and stop; IO)
 *   the CN sub-steps        main.cu:193-200 etc. (l_dt += CN_DT while l_dt < H_MAX)
 *
 * and the backward-Euler (BE) path of solve_bem.cu, which shares the solvers:
 *   xx_solve_by_bem         solve_bem.cu:293-305, 350-360, 448-461, 510-531
 *                           (ion update at the present voltage, val = val_ori,
 *                           b = 0, xx_update_matrix, the IO gap term, solve)
 *   xx_update_matrix        solve_bem.cu:258-291, 323-347, 407-445, 476-508
 *                           (diagonal += Cm/DT + conductances, b += Cm/DT v +
 *                           reversal-weighted conductances + i_ext)
 *   cg_cusparse_crs /       solve_bem.cu:7-77, 122-227 (the same CG and BiCG)
 *   bicg_cusparse_crs
 *
 * The CN cable matrix is the reference's unflipped matrix halved
 * (pkj_solve.cu:243 and the other xx_solve.cu files): off-diagonal -g/2,
 * diagonal +sum(g)/2; here it is -val/2 of the RKC-signed matrix, which is the
 * same value bit for bit.  Reductions (the CG / BiCG dot products) use
 * or_det_sum, the fixed 256-wide tree order the HIP kernels use; the
 * reference's cuBLAS order is not reproducible (parity unpinned at the last
 * bit).  Products of the sparse matrix with a vector sum each row's entries in
 * column order.  Without network synapses the reference's synaptic terms add
 * zero and are left out.
 *
 * The BE matrix is the unflipped, unhalved one (-val of the RKC-signed
 * matrix, pkj_solve.cu:238-245 with neither the RKC flip nor the CN halving).
 *
 * Pinning (tests/test_oracle_cn.py, tests/test_oracle_be.py): on a passive
 * cable the CN step equals the dense solve of
 * (C/DT + L/2 + G/2) v' = (C/DT - L/2 - G/2) v + I and the BE step the dense
 * solve of (C/DT + L + G) v' = (C/DT) v + G E + I, to the CG tolerance when CG
 * converges; CG and BiCG against numpy's dense solve on
 * symmetric and unsymmetric systems and their 100-iteration truncation on an
 * ill-conditioned one; the time step's second-order convergence on a
 * linear cable with a known solution.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CN_DT_ (0.025)
#define BE_DT_ (0.025)
#define H_MAX_ (0.125)

static void cn_ge(const or_pop *p, int id, double *g, double *e)
{
  switch (p->type) {
    case OR_PKJ: or_pkj_cn_ge(p, id, g, e); break;
    case OR_IO: or_io_cn_ge(p, id, g, e); break;
    case OR_GO: or_go_cn_ge(p, id, g, e); break;
    case OR_GR: or_gr_cn_ge(p, id, g, e); break;
    default: *g = 0.0; *e = p->iext[id]; break;
  }
}

static void cn_ion(or_pop *p, double DT)
{
  switch (p->type) {
    case OR_PKJ: or_pkj_cn_ion(p, DT); break;
    case OR_IO: or_io_cn_ion(p, DT); break;
    case OR_GO: or_go_cn_ion(p, DT); break;
    case OR_GR: or_gr_cn_ion(p, DT); break;
    default:
      for (int i = 0; i < p->nc; i++) p->cn_vold[i] = p->v[i];
      break;
  }
}

static void be_ge(const or_pop *p, int id, double DT, double *g, double *e)
{
  switch (p->type) {
    case OR_PKJ: or_pkj_be_ge(p, id, DT, g, e); break;
    case OR_IO: or_io_be_ge(p, id, DT, g, e); break;
    case OR_GO: or_go_be_ge(p, id, DT, g, e); break;
    case OR_GR: or_gr_be_ge(p, id, DT, g, e); break;
    default: *g = p->cm[id] / DT; *e = (p->cm[id] / DT) * p->v[id] + p->iext[id]; break;
  }
}

static void be_ion(or_pop *p, double DT)
{
  switch (p->type) {
    case OR_PKJ: or_pkj_be_ion(p, DT); break;
    case OR_IO: or_io_be_ion(p, DT); break;
    case OR_GO: or_go_be_ion(p, DT); break;
    case OR_GR: or_gr_be_ion(p, DT); break;
    default: break;
  }
}

static double *zal(long n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }

/* A x with the step's diagonal in place of the stored one */
static void spmv(const or_pop *p, const double *diag, const double *x, double *y)
{
  for (int i = 0; i < p->nc; i++) {
    double s = 0.0;
    for (int j = p->row[i]; j < p->row[i + 1]; j++) s += ((j == p->cn_dpos[i]) ? diag[i] : p->cn_val[j]) * x[p->col[j]];
    y[i] = s;
  }
}

/* A^T x: rows of the transpose in column order (the reference's
   create_transposed_matrix, solve_cnm.cu:566-615, lists them that way) */
static void spmv_t(const or_pop *p, const double *diag, const double *x, double *y)
{
  for (int i = 0; i < p->nc; i++) {
    double s = 0.0;
    for (int t = p->cn_trow[i]; t < p->cn_trow[i + 1]; t++) {
      const int k = p->cn_tsrc[t], j = p->cn_tcol[t];
      s += ((k == p->cn_dpos[j]) ? diag[j] : p->cn_val[k]) * x[j];
    }
    y[i] = s;
  }
}

static double dot(const double *a, const double *b, double *tmp, int n)
{
  for (int i = 0; i < n; i++) tmp[i] = a[i] * b[i];
  return or_det_sum(tmp, n);
}

/* cg_cusparse_crs, solve_cnm.cu:20-98; returns the iterations run */
static int cg(or_pop *p, const double *diag, double *x, const double *b)
{
  const int n = p->nc;
  double *r = zal(n), *pv = zal(n), *ap = zal(n), *tmp = zal(n);
  for (int i = 0; i < n; i++) { x[i] = 0.0 * x[i]; r[i] = b[i]; pv[i] = r[i]; }
  const double bnorm = dot(b, b, tmp, n);
  int k;
  for (k = 0; k < 100; k++) {
    spmv(p, diag, pv, ap);
    const double rk = dot(r, r, tmp, n);
    const double pap = dot(pv, ap, tmp, n);
    double alpha = rk / pap;
    for (int i = 0; i < n; i++) x[i] = alpha * pv[i] + x[i];
    alpha = -1.0 * alpha;
    for (int i = 0; i < n; i++) r[i] = alpha * ap[i] + r[i];
    const double rk1 = dot(r, r, tmp, n);
    if (sqrt(rk1) <= 1.0e-15 * sqrt(bnorm)) { k++; break; }
    const double beta = rk1 / rk;
    for (int i = 0; i < n; i++) pv[i] = beta * pv[i];
    for (int i = 0; i < n; i++) pv[i] = 1.0 * r[i] + pv[i];
  }
  free(r); free(pv); free(ap); free(tmp);
  return k;
}

/* bicg_cusparse_crs, solve_cnm.cu:617-727 */
static int bicg(or_pop *p, const double *diag, double *x, const double *b)
{
  const int n = p->nc;
  double *r = zal(n), *rs = zal(n), *pv = zal(n), *ps = zal(n), *ap = zal(n), *atps = zal(n), *tmp = zal(n);
  for (int i = 0; i < n; i++) { x[i] = 0.0 * x[i]; r[i] = b[i]; rs[i] = r[i]; pv[i] = r[i]; ps[i] = rs[i]; }
  const double bnorm = dot(b, b, tmp, n);
  int k;
  for (k = 0; k < 100; k++) {
    spmv(p, diag, pv, ap);
    spmv_t(p, diag, ps, atps);
    const double rk = dot(r, rs, tmp, n);
    const double pap = dot(ps, ap, tmp, n);
    double alpha = rk / pap;
    for (int i = 0; i < n; i++) x[i] = alpha * pv[i] + x[i];
    alpha = -1.0 * alpha;
    for (int i = 0; i < n; i++) r[i] = alpha * ap[i] + r[i];
    for (int i = 0; i < n; i++) rs[i] = alpha * atps[i] + rs[i];
    const double rk1 = dot(r, rs, tmp, n);
    if (sqrt(rk1) <= 1.0e-15 * sqrt(bnorm)) { k++; break; }
    const double beta = rk1 / rk;
    for (int i = 0; i < n; i++) pv[i] = beta * pv[i];
    for (int i = 0; i < n; i++) pv[i] = 1.0 * r[i] + pv[i];
    for (int i = 0; i < n; i++) ps[i] = beta * ps[i];
    for (int i = 0; i < n; i++) ps[i] = 1.0 * rs[i] + ps[i];
  }
  free(r); free(rs); free(pv); free(ps); free(ap); free(atps); free(tmp);
  return k;
}

/* The exact solve of A x = b for a matrix whose graph is a forest (every
   cell a tree, no intercell entries), by Gaussian elimination in leaf order:
   leaves are eliminated into their one remaining neighbour (ascending index
   first, then each node as its degree drops to one), which creates no
   fill-in, and the solution follows in the reverse order.  This is the
   direct counterpart of the reference's truncated CG / BiCG (the CN / BE
   systems are diagonally dominant, so no pivoting is needed); the GPU's
   per-cell solver eliminates level by level from the tree centre, a
   different order, so the two agree to rounding.  Returns -1 when the graph
   has a cycle or is not structurally symmetric. */
static double entry(const or_pop *p, const double *diag, int i, int j)
{
  for (int k = p->row[i]; k < p->row[i + 1]; k++)
    if (p->col[k] == j) return (k == p->cn_dpos[i]) ? diag[i] : p->cn_val[k];
  return NAN;
}

static int direct(or_pop *p, const double *diag, double *x, const double *b)
{
  const int n = p->nc;
  int *deg = (int *)calloc(n + 1, sizeof(int)), *gone = (int *)calloc(n + 1, sizeof(int));
  int *order = (int *)malloc(sizeof(int) * (n + 1)), *nbr = (int *)malloc(sizeof(int) * (n + 1));
  int *queue = (int *)malloc(sizeof(int) * (n + 1));
  double *d = zal(n), *r = zal(n);
  int bad = 0, nq = 0, head = 0, ne = 0;
  for (int i = 0; i < n; i++) {
    d[i] = diag[i];
    r[i] = b[i];
    for (int k = p->row[i]; k < p->row[i + 1]; k++)
      if (p->col[k] != i) {
        deg[i]++;
        if (isnan(entry(p, diag, p->col[k], i))) bad = 1;
      }
  }
  for (int i = 0; i < n && !bad; i++)
    if (deg[i] <= 1) queue[nq++] = i;
  while (!bad && head < nq) {
    const int i = queue[head++];
    int q = -1;
    for (int k = p->row[i]; k < p->row[i + 1]; k++)
      if (p->col[k] != i && !gone[p->col[k]]) q = p->col[k];
    gone[i] = 1;
    order[ne] = i;
    nbr[ne++] = q;
    if (q < 0) continue;               /* the last node of its tree */
    const double f = entry(p, diag, q, i) / d[i];
    d[q] -= f * entry(p, diag, i, q);
    r[q] -= f * r[i];
    if (--deg[q] == 1) queue[nq++] = q;
  }
  if (ne < n) bad = 1;                 /* nodes left on a cycle */
  for (int e = ne - 1; e >= 0 && !bad; e--) {
    const int i = order[e], q = nbr[e];
    x[i] = (q < 0) ? r[i] / d[i] : (r[i] - entry(p, diag, i, q) * x[q]) / d[i];
  }
  free(deg); free(gone); free(order); free(nbr); free(queue); free(d); free(r);
  return bad ? -1 : 0;
}

/* the path's linear solve: CG / BiCG as the reference, or the exact one */
static int solve(or_pop *p, const double *diag, double *x, const double *b)
{
  if (p->cn_direct) return direct(p, diag, x, b) == 0 ? 0 : -1;
  return (p->type == OR_IO) ? bicg(p, diag, x, b) : cg(p, diag, x, b);
}

/* switches the CN / BE paths to the exact solve (on = 1); 1 when the
   population's matrix is not a forest (gap junctions closing cycles) */
int or_cn_set_direct(or_pop *p, int on)
{
  if (on) {
    double *dg = zal(p->nc), *bb = zal(p->nc), *xx = zal(p->nc);
    for (int i = 0; i < p->nc; i++) { dg[i] = 1.0 + fabs(p->cn_val[p->cn_dpos[i]]) * 4.0; bb[i] = 1.0; }
    const int r = direct(p, dg, xx, bb);
    free(dg); free(bb); free(xx);
    if (r != 0) return 1;
  }
  p->cn_direct = on ? 1 : 0;
  return 0;
}

/* test hook: the exact solve for the given diagonal; 0, or -1 as above */
int or_cn_direct_test(or_pop *p, const double *diag, const double *b, double *x)
{
  return direct(p, diag, x, b);
}

/* the matrix (CN: the unflipped matrix halved, BE: unhalved), the diagonal
   positions and the transpose; 1 when a row lacks its stored diagonal */
static int sys_init(or_pop *p, int be)
{
  const int n = p->nc;
  free(p->cn_vold); free(p->cn_gamma1); free(p->cn_omega1); free(p->cn_val); free(p->cn_dpos);
  free(p->cn_trow); free(p->cn_tcol); free(p->cn_tsrc);
  p->cn_vold = zal(n); p->cn_gamma1 = zal(n); p->cn_omega1 = zal(n);
  p->cn_val = zal(p->nnz);
  p->cn_dpos = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int j = 0; j < p->nnz; j++) p->cn_val[j] = be ? -p->val[j] : -p->val[j] / 2.0;
  for (int i = 0; i < n; i++) {
    p->cn_dpos[i] = -1;
    for (int j = p->row[i]; j < p->row[i + 1]; j++)
      if (p->col[j] == i) p->cn_dpos[i] = j;
    if (p->cn_dpos[i] < 0) return 1;   /* both paths need the stored diagonal */
  }
  /* transpose (structure, and where each entry comes from) for BiCG */
  p->cn_trow = (int *)calloc(n + 1, sizeof(int));
  p->cn_tcol = (int *)malloc(sizeof(int) * (p->nnz > 0 ? p->nnz : 1));
  p->cn_tsrc = (int *)malloc(sizeof(int) * (p->nnz > 0 ? p->nnz : 1));
  for (int j = 0; j < p->nnz; j++) p->cn_trow[p->col[j] + 1]++;
  for (int i = 0; i < n; i++) p->cn_trow[i + 1] += p->cn_trow[i];
  int *fill = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  for (int i = 0; i < n; i++)             /* rows of A in order: transposed rows come out in column order */
    for (int j = p->row[i]; j < p->row[i + 1]; j++) {
      const int c = p->col[j], t = p->cn_trow[c] + fill[c]++;
      p->cn_tcol[t] = i;
      p->cn_tsrc[t] = j;
    }
  free(fill);
  p->n_steps = p->n_stages = 0;
  return 0;
}

int or_cn_init(or_pop *p)
{
  if (sys_init(p, 0) != 0) return 1;
  /* xx_cnm_vec_initialize */
  for (int i = 0; i < p->nc; i++) {
    double g, e;
    cn_ge(p, i, &g, &e);
    p->cn_gamma1[i] = g / 2.0;
    p->cn_omega1[i] = e / 2.0;
    p->cn_vold[i] = p->v[i];
  }
  return 0;
}

/* the step of the path: the reference's DT, or the one set with or_cn_set_dt */
static double step_dt(const or_pop *p, int be) { return p->cn_dt > 0.0 ? p->cn_dt : (be ? BE_DT_ : CN_DT_); }

/* sets the CN / BE step [ms] (dt <= 0: the reference's 0.025); 1 unless
   H_MAX / dt is a whole number of steps */
int or_cn_set_dt(or_pop *p, double dt)
{
  if (dt <= 0.0) { p->cn_dt = 0.0; return 0; }
  const double k = H_MAX_ / dt;
  if (k < 1.0 || fabs(k - floor(k + 0.5)) > 1e-9) return 1;
  p->cn_dt = dt;
  return 0;
}

/* one CN step of CN_DT (xx_solve_by_cnm) */
int or_cn_step(or_pop *p)
{
  const int n = p->nc;
  const double DT = step_dt(p, 0);
  cn_ion(p, DT);
  double *gamma2 = zal(n), *omega2 = zal(n), *diag = zal(n), *b = zal(n);
  if (p->type == OR_IO && p->n > 1)
    for (int c = 0; c < p->n; c++) omega2[c * 3] = or_io_gap_into(p, c);
  if (p->n_syn_in) or_syn_cn(p, gamma2, omega2);
  for (int i = 0; i < n; i++) {
    double g, e;
    cn_ge(p, i, &g, &e);
    gamma2[i] += g / 2.0;
    omega2[i] += e / 2.0;
    diag[i] = p->cn_val[p->cn_dpos[i]] + ((p->cm[i] / DT) + gamma2[i]);
    double bb = 0.0;
    for (int j = p->row[i]; j < p->row[i + 1]; j++) bb -= p->v[p->col[j]] * p->cn_val[j];
    bb += (p->cm[i] / DT - p->cn_gamma1[i]) * p->v[i] + p->cn_omega1[i] + omega2[i];
    b[i] = bb;
  }
  for (int i = 0; i < n; i++) { p->cn_omega1[i] = omega2[i]; p->cn_gamma1[i] = gamma2[i]; }
  const int it = solve(p, diag, p->v, b);
  p->n_steps++;
  p->n_stages += it;
  free(gamma2); free(omega2); free(diag); free(b);
  return 0;
}

/* the reference's sub-step loop for one H_MAX output interval from time t */
int or_cn_advance(or_pop *p, double t)
{
  if (p->cn_dt > 0.0) {                 /* a set step: exactly H_MAX / dt steps */
    const int nsub = (int)floor(H_MAX_ / p->cn_dt + 0.5);
    for (int j = 0; j < nsub; j++) {
      or_set_current(p, t + j * p->cn_dt);
      or_cn_step(p);
    }
    return 0;
  }
  double l_dt = 0.0;
  while (l_dt < H_MAX_) {
    or_set_current(p, t + l_dt);
    or_cn_step(p);
    l_dt += step_dt(p, 0);
  }
  return 0;
}

/* test hook: solve A x = b by the CN path's CG (sym = 1) or BiCG (sym = 0)
   for the population's current diagonal (diag given) and stored
   off-diagonals; returns the iterations run */
int or_cn_solve_test(or_pop *p, const double *diag, const double *b, double *x, int sym)
{
  return sym ? cg(p, diag, x, b) : bicg(p, diag, x, b);
}

/* ---------------- the backward-Euler path ---------------- */

int or_be_init(or_pop *p) { return sys_init(p, 1); }

/* the BE system of the present state (xx_update_matrix after val = val_ori,
   b = 0, then the IO gap term): the diagonal and the right-hand side */
static void be_assemble(const or_pop *p, double DT, double *diag, double *b)
{
  for (int i = 0; i < p->nc; i++) {
    diag[i] = p->cn_val[p->cn_dpos[i]];
    b[i] = 0.0;
  }
  if (p->n_syn_in) or_syn_be(p, diag, b);   /* add_*_val before xx_update_matrix */
  for (int i = 0; i < p->nc; i++) {
    double g, e;
    be_ge(p, i, DT, &g, &e);
    diag[i] = diag[i] + g;
    b[i] = b[i] + e;
  }
  if (p->type == OR_IO && p->n > 1)
    for (int c = 0; c < p->n; c++) b[c * 3] = or_io_gap_add(p, c, b[c * 3]);
}

/* test hook: the BE system of the present state, without the ion update */
int or_be_assemble_test(or_pop *p, double *diag, double *b)
{
  be_assemble(p, step_dt(p, 1), diag, b);
  return 0;
}

/* one BE step of BE_DT (xx_solve_by_bem) */
int or_be_step(or_pop *p)
{
  const int n = p->nc;
  const double DT = step_dt(p, 1);
  be_ion(p, DT);
  double *diag = zal(n), *b = zal(n);
  be_assemble(p, DT, diag, b);
  const int it = solve(p, diag, p->v, b);
  p->n_steps++;
  p->n_stages += it;
  free(diag); free(b);
  return 0;
}

/* the sub-step loop for one H_MAX output interval from time t */
int or_be_advance(or_pop *p, double t)
{
  if (p->cn_dt > 0.0) {                 /* a set step: exactly H_MAX / dt steps */
    const int nsub = (int)floor(H_MAX_ / p->cn_dt + 0.5);
    for (int j = 0; j < nsub; j++) {
      or_set_current(p, t + j * p->cn_dt);
      or_be_step(p);
    }
    return 0;
  }
  double l_dt = 0.0;
  while (l_dt < H_MAX_) {
    or_set_current(p, t + l_dt);
    or_be_step(p);
    l_dt += step_dt(p, 1);
  }
  return 0;
}
