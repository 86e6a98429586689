/*
 * Oracle: synapses between the populations (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Restates syn.cu / syn.cuh:
 *   constants, element layout   syn.cuh:13-73
 *   mfgr_create / _update        syn.cu:42-102   (4 mossy-fibre inputs per granule cell,
 *                                                 compartments 0..3; Poisson 5 Hz, 60 Hz in
 *                                                 [S_Stimuli, E_Stimuli))
 *   grgo_create / _update        syn.cu:144-289  (granule axon -> Golgi compartment 0)
 *   gogr_create / _update        syn.cu:292-437  (Golgi compartment 30 -> granule compartment 16)
 *   grpkj_create / _update       syn.cu:440-663  (25 synapses per granule-Purkinje pair on the
 *                                                 dendritic compartment nearest in x)
 *   mlipkj_create                syn.cu:666-759  (one inert synapse per Purkinje cell on 1599)
 *   the drivers                  syn.cu:821-917, called after every macro step (main.cu:302-306)
 * and their coupling into the solvers:
 *   RKC  GR: b += val (y - E) at each ion update (add_mfgr_val / add_gogr_val,
 *            solve_rkc.cu:200-224, 1117-1118);
 *        GO: s = sum val (y - E) at the macro step's start, b += s at each ion update
 *            (add_grgo_val_new / go_update_matrix_new, solve_rkc.cu:266-287, 1181, 1199-1200);
 *        PKJ: the same s, b += s exp(-dt_ion / 0.6) (solve_rkc.cu:344-367, 1245-1249, 1267-1269)
 *   CN   gamma2 += val / 2, omega2 += val E / 2 before the ion sums (solve_cnm.cu:100-128,
 *        245-257, 393-418); the inhibitory input to Purkinje cells is not coupled there
 *   BE   diagonal += val, b += val E before the ion sums (solve_bem.cu:229-255, 306-320);
 *        Purkinje cells are not coupled on the BE path (pkj_solve_by_bem has no synapses)
 *
 * Every atomicAdd of the reference is restated as a sum in synapse order
 * (synapse types in the order the reference launches them); the reference's
 * order is nondeterministic, so parity at the last bit is unpinned there.
 * The reference draws its random numbers from cuRAND XORWOW streams seeded
 * with rand(); here they come from Philox4x32-10 (or_philox.c) keyed by the
 * seed given, counter (synapse slot, macro step, stream, 0), as a uniform in
 * (0, 1] like curand_uniform: the distribution is the reference's, the
 * stream is not (parity unpinned for the random draws).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GR_COMP_ 578
#define GO_COMP_ 131
#define GO_COMP_DEND_ 10
#define PKJ_COMP_ 1600
#define RKC_DT_ 0.125
#define S_STIMULI_ 900.0
#define E_STIMULI_ 1000.0

#define G_MFGR_AMPA (0.18e-9)
#define G_MFGR_NMDA (0.025e-9)
#define E_MFGR (0.0)
#define G_GRGO_AMPA (45.5e-9)
#define G_GRGO_NMDA (30.0e-9)
#define E_GRGO (0.0)
#define G_GOGR_GABA (0.028e-9)
#define E_GOGR (-82.0)
#define G_GRPKJ_AMPA (0.7e-9)
#define E_GRPKJ (0.0)
#define G_MLIPKJ_GABA (1.0e-9)
#define E_MLIPKJ (-75.0)
#define N_SYN_PER_GRPKJ 25

/* Philox stream tags */
enum { ST_MF = 0, ST_GRPKJ_JITTER = 1, ST_MLI = 2 };

static double *dz(long n) { return (double *)calloc(n > 0 ? n : 1, sizeof(double)); }
static int *iz(long n) { return (int *)calloc(n > 0 ? n : 1, sizeof(int)); }

/* uniform in (0, 1] from the first Philox word */
static double philox_u(const or_syn *s, uint32_t slot, uint32_t step, uint32_t stream)
{
  const uint32_t ctr[4] = {slot, step, stream, 0u};
  uint32_t out[4];
  or_philox4x32_10(ctr, s->key, out);
  return ((double)out[0] + 1.0) * (1.0 / 4294967296.0);
}

static or_syn *syn_alloc(int kind, int n, double E, uint64_t seed)
{
  or_syn *s = (or_syn *)calloc(1, sizeof(or_syn));
  s->kind = kind;
  s->n = n;
  s->E = E;
  s->pre = iz(n);
  s->post = iz(n);
  for (int k = 0; k < 3; k++) s->x[k] = dz(n);
  s->weight = dz(n);
  s->val = dz(n);
  s->old_v = dz(n);
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  return s;
}

void or_syn_free(or_syn *s)
{
  if (!s) return;
  free(s->pre); free(s->post);
  for (int k = 0; k < 3; k++) free(s->x[k]);
  free(s->weight); free(s->val); free(s->old_v);
  free(s);
}

/* mfgr_create, syn.cu:104-128 (with mfgr_initialize, :42-61) */
or_syn *or_mfgr_create(int n_gr, uint64_t seed)
{
  or_syn *s = syn_alloc(OR_SYN_MFGR, n_gr * 4, E_MFGR, seed);
  for (int id = 0; id < n_gr; id++)
    for (int j = 0; j < 4; j++) {
      const int k = id * 4 + j;
      s->pre[k] = id;
      s->post[k] = j + id * GR_COMP_;
      s->weight[k] = 40000.0 / 4.0;
    }
  return s;
}

/* grgo_create, syn.cu:181-261 (with grgo_initialize, :144-179): a granule
   cell contacts every Golgi cell within 16 um in x and 220 um in y of its
   soma grid point, from the axonal compartment at the Golgi soma's height */
or_syn *or_grgo_create(const or_pop *gr, const or_pop *go)
{
  const int nx_gr = gr->nx, ny_gr = gr->ny, nx_go = go->nx, ny_go = go->ny;
  const int n_gr = nx_gr * ny_gr, n_go = nx_go * ny_go;
  int *label = (int *)malloc(sizeof(int) * (size_t)(ny_go > 0 ? ny_go : 1) * (n_gr > 0 ? n_gr : 1));
  int *num_gr = iz(n_gr), *num_go = iz(n_go);
  int total = 0;
  for (int i_gr = 0; i_gr < n_gr; i_gr++) {
    const double lx_gr = (int)(i_gr % nx_gr) * 16.0, ly_gr = (int)(i_gr / nx_gr) * 16.0;
    int l_count = 0;
    for (int i_go = 0; i_go < n_go; i_go++) {
      const double lx_go = (int)(i_go % nx_go) * 32.0 + 8.0, ly_go = (int)(i_go / nx_go) * 32.0 + 8.0;
      if (fabs(lx_go - lx_gr) < 16.0 && fabs(ly_go - ly_gr) < 220.0) {
        label[i_gr * ny_go + l_count] = i_go;
        l_count++;
        num_go[i_go]++;
      }
    }
    num_gr[i_gr] = l_count;
    total += l_count;
  }
  or_syn *s = syn_alloc(OR_SYN_GRGO, total, E_GRGO, 0);
  s->pre_pop = gr;
  int k = 0;
  for (int l_n_gr = 0; l_n_gr < n_gr; l_n_gr++)
    for (int c = 0; c < num_gr[l_n_gr]; c++, k++) {
      const int l_n_go = label[l_n_gr * ny_go + c];
      const int iy_gr = l_n_gr / nx_gr, iy_go = l_n_go / nx_go;
      const float diff_grgo = 8.0 + iy_go * 32.0 - iy_gr * 16.0;
      const int gr_ax_comp = 250 + 77 - (int)(diff_grgo / 10.0);
      s->pre[k] = gr_ax_comp + GR_COMP_ * l_n_gr;
      s->post[k] = 0 + GO_COMP_ * l_n_go;
      s->weight[k] = 64.0 / 4.0 / (num_go[l_n_go] * 1.0 / 4.0);
      s->old_v[k] = 1000.0;
    }
  free(label); free(num_gr); free(num_go);
  return s;
}

/* gogr_create, syn.cu:339-410 (with gogr_initialize3, :292-337): a Golgi cell
   contacts every granule cell within 32 um in x and in y, from its
   compartment 30 onto the granule cell's compartment 16 */
or_syn *or_gogr_create(const or_pop *go, const or_pop *gr)
{
  const int nx_go = go->nx, ny_go = go->ny, nx_gr = gr->nx, ny_gr = gr->ny;
  const int n_go = nx_go * ny_go, n_gr = nx_gr * ny_gr;
  int *label = (int *)malloc(sizeof(int) * (size_t)(n_go > 0 ? n_go : 1) * 16);
  int *num_go = iz(n_go), *num_gr = iz(n_gr);
  int total = 0;
  for (int i_go = 0; i_go < n_go; i_go++) {
    int l_count = 0;
    const double lx_go = (int)(i_go % nx_go) * 32.0 + 8.0, ly_go = (int)(i_go / nx_go) * 32.0 + 8.0;
    for (int i_gr = 0; i_gr < n_gr; i_gr++) {
      const double lx_gr = (int)(i_gr % nx_gr) * 16.0, ly_gr = (int)(i_gr / nx_gr) * 16.0;
      if (fabs(lx_go - lx_gr) < 32.0 && fabs(ly_go - ly_gr) < 32.0) {
        label[i_go * 4 * 4 + l_count] = i_gr;
        l_count++;
        num_gr[i_gr]++;
      }
    }
    num_go[i_go] = l_count;
    total += l_count;
  }
  or_syn *s = syn_alloc(OR_SYN_GOGR, total, E_GOGR, 0);
  s->pre_pop = go;
  int k = 0;
  for (int l_n_go = 0; l_n_go < n_go; l_n_go++)
    for (int c = 0; c < num_go[l_n_go]; c++, k++) {
      const int l_n_gr = label[l_n_go * 4 * 4 + c];
      s->pre[k] = (GO_COMP_DEND_ * 3) + GO_COMP_ * l_n_go;
      s->post[k] = 16 + GR_COMP_ * l_n_gr;
      s->weight[k] = 160000.0 * 4.0 / (num_gr[l_n_gr] * 4.0);
      s->old_v[k] = 1000.0;
    }
  free(label); free(num_go); free(num_gr);
  return s;
}

/* grpkj_create, syn.cu:519-644 (with grpkj_initialize, :440-517): a granule
   cell contacts every Purkinje cell within 120 um in x and 220 um in y with
   25 synapses, each on the dendritic compartment whose x (from
   Pkj_location_info.csv) is nearest the granule cell's offset jittered by
   10 (U - 0.5) um; the running minimum is not reset between a pair's 25
   synapses, as in the reference.  x1600: the x column of the location file. */
or_syn *or_grpkj_create(const or_pop *gr, const or_pop *pkj, const double *x1600, uint64_t seed)
{
  const int nx_gr = gr->nx, ny_gr = gr->ny, nx_pkj = pkj->nx, ny_pkj = pkj->ny;
  const int n_gr = nx_gr * ny_gr, n_pkj = nx_pkj * ny_pkj;
  int *label = (int *)malloc(sizeof(int) * (size_t)(n_pkj > 0 ? n_pkj : 1) * (n_gr > 0 ? n_gr : 1));
  int *num_gr = iz(n_gr), *num_pkj = iz(n_pkj);
  int total = 0;
  for (int i_gr = 0; i_gr < n_gr; i_gr++) {
    const double lx_gr = (int)(i_gr % nx_gr) * 16.0, ly_gr = (int)(i_gr / nx_gr) * 16.0;
    int l_count = 0;
    for (int i_pkj = 0; i_pkj < n_pkj; i_pkj++) {
      const double lx_pkj = (int)(i_pkj % nx_pkj) * 32.0 + 8.0, ly_pkj = (int)(i_pkj / nx_pkj) * 32.0 + 8.0;
      if (fabs(lx_pkj - lx_gr) < 120.0 && fabs(ly_pkj - ly_gr) < 220.0) {
        label[i_gr * n_pkj + l_count] = i_pkj;
        l_count++;
        num_pkj[i_pkj]++;
      }
    }
    num_gr[i_gr] = l_count;
    total += l_count * N_SYN_PER_GRPKJ;
  }
  or_syn *s = syn_alloc(OR_SYN_GRPKJ, total, E_GRPKJ, seed);
  s->pre_pop = gr;
  int pair = 0;
  for (int l_n_gr = 0; l_n_gr < n_gr; l_n_gr++)
    for (int c = 0; c < num_gr[l_n_gr]; c++, pair++) {
      const int l_n_pkj = label[l_n_gr * n_pkj + c];
      const int ix_gr = l_n_gr % nx_gr, ix_pkj = l_n_pkj % nx_pkj;
      const int iy_gr = l_n_gr / nx_gr, iy_pkj = l_n_pkj / nx_pkj;
      const double diff_x = ix_gr * 16.0 - (8.0 + ix_pkj * 32.0);
      const double diff_y = iy_gr * 16.0 - (8.0 + iy_pkj * 32.0);
      const int gr_ax_comp = 250 + 77 + (int)(diff_y / 10.0);
      int pkj_d_comp = -1;
      double min_x = 1000000.0;
      const int guarantee = 1;
      for (int i_cp = 0; i_cp < N_SYN_PER_GRPKJ; i_cp++) {
        const int k = pair * N_SYN_PER_GRPKJ + i_cp;
        const double f = philox_u(s, (uint32_t)k, 0u, ST_GRPKJ_JITTER) - 0.5;
        for (int i_comp = 0; i_comp < PKJ_COMP_; i_comp++)
          if (fabs(x1600[i_comp] - diff_x + 10.0 * f) < min_x) {
            min_x = fabs(x1600[i_comp] - diff_x + 10.0 * f);
            pkj_d_comp = i_comp;
          }
        s->pre[k] = gr_ax_comp + GR_COMP_ * l_n_gr;
        s->post[k] = pkj_d_comp + PKJ_COMP_ * l_n_pkj;
        s->weight[k] = 1000000.0 * 0.3 / (512.0 * num_pkj[l_n_pkj] * 1.0 * guarantee / 10.0);
        s->old_v[k] = 1000.0;
      }
    }
  free(label); free(num_gr); free(num_pkj);
  return s;
}

/* mlipkj_create, syn.cu:690-738 (with mlipkj_initialize, :666-688): one
   synapse per Purkinje cell on compartment 1599, weight W_MLIPKJ = 0; the
   reference's drivers never update it and no solver couples it */
or_syn *or_mlipkj_create(int n_pkj, uint64_t seed)
{
  or_syn *s = syn_alloc(OR_SYN_MLIPKJ, n_pkj, E_MLIPKJ, seed);
  for (int id = 0; id < n_pkj; id++) {
    s->pre[id] = id;
    s->post[id] = 1599 + id * PKJ_COMP_;
    s->weight[id] = 0.0 / (1.0);
  }
  return s;
}

/* one update after a macro step starting at t (xx_update kernels) */
void or_syn_update(or_syn *s, double t)
{
  const double *pv = s->pre_pop ? s->pre_pop->v : NULL;
  for (int id = 0; id < s->n; id++) {
    double *x0 = s->x[0], *x1 = s->x[1], *x2 = s->x[2];
    switch (s->kind) {
      case OR_SYN_MFGR: {   /* mfgr_update, syn.cu:63-102 */
        double fr_mf = 5.0;
        (S_STIMULI_ <= t && t < E_STIMULI_) ? (fr_mf = 60.0) : (fr_mf = 5.0);
        const double f = philox_u(s, (uint32_t)id, (uint32_t)s->step, ST_MF);
        const int firing_flag = (fr_mf * RKC_DT_ * 0.001 > f) ? 1 : 0;
        x0[id] = x0[id] * exp(-RKC_DT_ / 1.2) + firing_flag;
        /* the NMDA trace is fed from the just-updated AMPA trace (syn.cu:96) */
        x1[id] = x0[id] * exp(-RKC_DT_ / 52.0) + firing_flag;
        s->val[id] = s->weight[id] * (G_MFGR_AMPA * x0[id] + G_MFGR_NMDA * x1[id]);
        break;
      }
      case OR_SYN_GRGO: {   /* grgo_update, syn.cu:263-288 */
        const double pre_comp_v = pv[s->pre[id]];
        const int firing_flag = (pre_comp_v > 0.0 && s->old_v[id] < 0.0) ? 1 : 0;
        x0[id] = x0[id] * exp(-RKC_DT_ / 1.5) + firing_flag;
        x1[id] = x1[id] * exp(-RKC_DT_ / 31.0) + firing_flag;
        x2[id] = x2[id] * exp(-RKC_DT_ / 170.0) + firing_flag;
        s->val[id] = s->weight[id] * (G_GRGO_AMPA * x0[id] + G_GRGO_NMDA * (0.33 * x1[id] + 0.67 * x2[id]));
        s->old_v[id] = pre_comp_v;
        break;
      }
      case OR_SYN_GOGR: {   /* gogr_update, syn.cu:412-437 */
        const double pre_comp_v = pv[s->pre[id]];
        const int firing_flag = (pre_comp_v > 0.0 && s->old_v[id] < 0.0) ? 1 : 0;
        x0[id] = x0[id] * exp(-RKC_DT_ / 7.0) + firing_flag;
        x1[id] = x1[id] * exp(-RKC_DT_ / 59.0) + firing_flag;
        s->val[id] = s->weight[id] * G_GOGR_GABA * (0.43 * x0[id] + 0.57 * x1[id]);
        s->old_v[id] = pre_comp_v;
        break;
      }
      case OR_SYN_GRPKJ: {  /* grpkj_update, syn.cu:646-668 */
        const double pre_comp_v = pv[s->pre[id]];
        const int firing_flag = (pre_comp_v > 0.0 && s->old_v[id] < 0.0) ? 1 : 0;
        x0[id] = x0[id] * exp(-RKC_DT_ / 6.0) + firing_flag;
        s->val[id] = x0[id] * s->weight[id] * G_GRPKJ_AMPA;
        s->old_v[id] = pre_comp_v;
        break;
      }
      case OR_SYN_MLIPKJ: { /* mlipkj_update, syn.cu:761-788 (not called by the drivers) */
        const double fr_mli = 30.0;
        const double f = philox_u(s, (uint32_t)id, (uint32_t)s->step, ST_MLI);
        const int firing_flag = (fr_mli * 0.125 * 0.001 > f) ? 1 : 0;
        x0[id] = x0[id] * 0.98757780049 + firing_flag;
        s->val[id] = s->weight[id] * (G_MLIPKJ_GABA * x0[id]);
        break;
      }
    }
  }
  s->step++;
}

/* gr_synapse_update / go_synapse_update / pkj_synapse_update (syn.cu:821-917):
   the same for every solver; the inhibitory input to Purkinje cells is not
   updated there */
void or_net_update(or_syn *mfgr, or_syn *gogr, or_syn *grgo, or_syn *grpkj, double t)
{
  if (mfgr && mfgr->n > 0) or_syn_update(mfgr, t);
  if (gogr && gogr->n > 0) or_syn_update(gogr, t);
  if (grgo && grgo->n > 0) or_syn_update(grgo, t);
  if (grpkj && grpkj->n > 0) or_syn_update(grpkj, t);
}

int or_pop_attach_syn(or_pop *p, or_syn *s)
{
  if (!s || s->n == 0) return 0;
  if (p->n_syn_in >= 2) return 1;
  p->syn_in[p->n_syn_in++] = s;
  if ((p->type == OR_GO || p->type == OR_PKJ) && !p->syn_s) p->syn_s = dz(p->nc);
  return 0;
}

/* RKC, granule cells: b += val (y - E) in synapse order (add_mfgr_val, add_gogr_val) */
void or_syn_add_b(or_pop *p, const double *vv)
{
  for (int a = 0; a < p->n_syn_in; a++) {
    const or_syn *s = p->syn_in[a];
    for (int k = 0; k < s->n; k++) {
      const int post_num = s->post[k];
      p->b[post_num] += s->val[k] * (vv[post_num] - s->E);
    }
  }
}

/* RKC, Golgi and Purkinje cells: the macro step's synaptic current from its
   starting y (reset_b + add_grgo_val_new / add_grpkj_val_new) */
void or_syn_macro_start(or_pop *p)
{
  if (!p->syn_s) return;
  const double *y = p->vec[OR_Y];
  for (int i = 0; i < p->nc; i++) p->syn_s[i] = 0.0;
  for (int a = 0; a < p->n_syn_in; a++) {
    const or_syn *s = p->syn_in[a];
    if (s->kind == OR_SYN_MLIPKJ) continue;
    for (int k = 0; k < s->n; k++) {
      const int post_num = s->post[k];
      p->syn_s[post_num] += s->val[k] * (y[post_num] - s->E);
    }
  }
}

/* RKC ion update: go_update_matrix_new (b += s) / pkj_update_matrix_new
   (b += s exp(-dt / 0.6), only with granule input present) */
void or_syn_ion_b(or_pop *p, const double *vv, double sum_dt)
{
  if (p->type == OR_GR) {
    or_syn_add_b(p, vv);
  } else if (p->type == OR_GO && p->syn_s) {
    for (int i = 0; i < p->nc; i++) p->b[i] += p->syn_s[i];
  } else if (p->type == OR_PKJ && p->syn_s) {
    int has_grpkj = 0;
    for (int a = 0; a < p->n_syn_in; a++) has_grpkj |= (p->syn_in[a]->kind == OR_SYN_GRPKJ);
    if (!has_grpkj) return;
    const double sum_decay = exp(-sum_dt / 0.6);
    for (int i = 0; i < p->nc; i++) p->b[i] += p->syn_s[i] * sum_decay;
  }
}

/* CN: gamma2 += val / 2, omega2 += val E / 2 (GR, GO, PKJ; not the
   inhibitory Purkinje input, whose call is commented out in the reference) */
void or_syn_cn(const or_pop *p, double *gamma2, double *omega2)
{
  for (int a = 0; a < p->n_syn_in; a++) {
    const or_syn *s = p->syn_in[a];
    if (s->kind == OR_SYN_MLIPKJ) continue;
    for (int k = 0; k < s->n; k++) {
      const int post_num = s->post[k];
      gamma2[post_num] += 0.5 * s->val[k];
      omega2[post_num] += 0.5 * s->val[k] * s->E;
    }
  }
}

/* BE: diagonal += val, b += val E (GR, GO; a model-free population as
   well, for tests; not Purkinje cells) */
void or_syn_be(const or_pop *p, double *diag, double *b)
{
  if (p->type == OR_PKJ || p->type == OR_IO) return;
  for (int a = 0; a < p->n_syn_in; a++) {
    const or_syn *s = p->syn_in[a];
    for (int k = 0; k < s->n; k++) {
      const int post_num = s->post[k];
      diag[post_num] += s->val[k];
      b[post_num] += s->val[k] * s->E;
    }
  }
}

/* test access: fields 0..2 traces, 3 weight, 4 val, 5 old_v */
int or_syn_n(const or_syn *s) { return s->n; }
void or_syn_get(const or_syn *s, int field, double *dst)
{
  const double *src = field < 3 ? s->x[field] : field == 3 ? s->weight : field == 4 ? s->val : s->old_v;
  memcpy(dst, src, sizeof(double) * (size_t)s->n);
}
void or_syn_set(or_syn *s, int field, const double *src)
{
  double *dst = field < 3 ? s->x[field] : field == 3 ? s->weight : field == 4 ? s->val : s->old_v;
  memcpy(dst, src, sizeof(double) * (size_t)s->n);
}
void or_syn_comp(const or_syn *s, int *pre, int *post)
{
  memcpy(pre, s->pre, sizeof(int) * (size_t)s->n);
  memcpy(post, s->post, sizeof(int) * (size_t)s->n);
}
