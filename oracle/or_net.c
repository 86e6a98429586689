/*
 * Oracle: fixed in-degree network between granule, Golgi and Purkinje
 * populations (TEST INFRASTRUCTURE; never linked into the product).
 *
 * Restates, in plain sequential C, the network the product builds over its
 * populations: the synapse generator, the per-(compartment, receptor)
 * conductances, the spike ring with a whole-interval delay and the mossy-fibre
 * draws.  The reference's own network (syn.cu) connects by geometric rules;
 * what is kept from it is the per-macro-step order of main.cu:182-311 with
 * network_sim on (advance every population, then update the synapses from the
 * presynaptic voltages), the trace update s <- s exp(-dt / tau) + spike with
 * val = w s (syn.cu:87-111, 264-288, 413-437, 647-668), the spike test
 * pre V > 0 && old_v < 0 (as an upward crossing of a threshold between
 * interval ends) and the mossy-fibre firing test rate x dt > u
 * (syn.cu:87-102).  The coupling into the populations is the reference's,
 * through or_syn.c: each (population, receptor) becomes one synapse set whose
 * entries are that receptor's conductance slots.
 *
 * Pinned by tests/test_oracle_net.py: generator properties (in-degrees,
 * ranges, exact total, determinism, a brute-force restatement in Python of
 * the draw of single synapses), delivery against a hand-built case, and the
 * product's generator and delivery against this one.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "oracle.h"

#define NET_KEY 0x6e657477u  /* "netw" */
#define MF_KEY 0x6d666472u   /* "mfdr" */

static int post_pop(int kind) { return kind <= 1 ? 0 : kind == 2 ? 1 : 2; }
static int pre_pop(int kind) { return kind == 0 ? -1 : kind == 1 ? 1 : 0; }
static int receptor(int kind) { return kind == 1 ? 1 : 0; }

/* 53-bit uniform in [0, 1) from two 32-bit words: (a >> 5) 2^26 + (b >> 6)
   over 2^53 (the res53 construction of Matsumoto and Nishimura's MT19937
   reference code) */
static double u53(uint32_t a, uint32_t b)
{
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

static long long extra_of(const or_net_spec *s)
{
  long long base = 0;
  for (int k = 0; k < 4; k++) base += (long long)s->deg[k] * s->n[post_pop(k)];
  return s->total > 0 ? s->total - base : 0;
}

long long or_net_size(const or_net_spec *s)
{
  long long size = extra_of(s);
  for (int k = 0; k < 4; k++) size += (long long)s->deg[k] * s->n[post_pop(k)];
  return size;
}

long long or_net_generate(const or_net_spec *s, int8_t *kind, int32_t *pre, int32_t *post, double *w_nS)
{
  const long long extra = extra_of(s);
  const uint32_t key[2] = {(uint32_t)s->seed ^ NET_KEY, (uint32_t)(s->seed >> 32)};
  long long j = 0;
  int fibre = 0;
  for (int k = 0; k < 4; k++) {
    const int po = post_pop(k), pr = pre_pop(k);
    for (int c = 0; c < s->n[po]; c++) {
      const int m = s->deg[k] + ((k == 3 && c < extra) ? 1 : 0);
      for (int q = 0; q < m; q++) {
        const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)q, (uint32_t)k, 0u};
        uint32_t r[4];
        or_philox4x32_10(ctr, key, r);
        int src;
        if (pr < 0) {
          src = fibre++;
        } else {
          src = (int)(((uint64_t)r[0] * (uint64_t)s->n[pr]) >> 32);
        }
        const uint64_t span = (uint64_t)(s->tgt_hi[po] - s->tgt_lo[po]);
        kind[j] = (int8_t)k;
        pre[j] = src;
        post[j] = c * s->comp[po] + s->tgt_lo[po] + (int)(((uint64_t)r[3] * span) >> 32);
        w_nS[j] = s->w_min_nS + (s->w_max_nS - s->w_min_nS) * u53(r[1], r[2]);
        j++;
      }
    }
  }
  return j;
}

struct or_net {
  or_net_spec s;
  or_pop *pop[3];
  long long nsyn;
  int8_t *kind;
  int32_t *pre, *post;
  double *w;
  /* per postsynaptic population: slots (compartment, receptor), their
     synapses in generation order, the conductance */
  int nslot[3];
  int *slot_comp[3], *slot_rec[3], *row[3], *syn[3];
  double *G[3];
  or_syn *set[3][2];      /* per receptor: entries = that receptor's slots */
  int *set_slot[3][2];    /* slot index of each entry */
  /* spikes: ring[(k mod nring)][pop][cell] */
  int nring;
  unsigned char *ring;
  int ring_stride, ring_off[3];
  double *prev[3];
  long k;
};

static or_syn *pseudo_set(int kind, int n, double E)
{
  or_syn *s = (or_syn *)calloc(1, sizeof(or_syn));
  s->kind = kind;
  s->n = n;
  s->E = E;
  s->pre = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  s->post = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  for (int k = 0; k < 3; k++) s->x[k] = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  s->weight = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  s->val = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  s->old_v = (double *)calloc(n > 0 ? n : 1, sizeof(double));
  return s;
}

or_net *or_net_create(or_pop *gr, or_pop *go, or_pop *pkj, const or_net_spec *spec)
{
  or_net *N = (or_net *)calloc(1, sizeof(or_net));
  N->s = *spec;
  N->pop[0] = gr;
  N->pop[1] = go;
  N->pop[2] = pkj;
  for (int p = 0; p < 3; p++) {
    N->s.n[p] = N->pop[p]->n;
    N->s.comp[p] = N->pop[p]->comp;
  }
  N->nsyn = or_net_size(&N->s);
  const long long m = N->nsyn > 0 ? N->nsyn : 1;
  N->kind = (int8_t *)malloc(m);
  N->pre = (int32_t *)malloc(m * sizeof(int32_t));
  N->post = (int32_t *)malloc(m * sizeof(int32_t));
  N->w = (double *)malloc(m * sizeof(double));
  or_net_generate(&N->s, N->kind, N->pre, N->post, N->w);
  /* slots of each population: compartment order, AMPA before GABA_A */
  for (int p = 0; p < 3; p++) {
    const int nc = N->pop[p]->nc;
    int *cnt = (int *)calloc((size_t)nc * 2, sizeof(int));
    for (long long j = 0; j < N->nsyn; j++)
      if (post_pop(N->kind[j]) == p) cnt[(size_t)N->post[j] * 2 + receptor(N->kind[j])]++;
    int ns = 0;
    for (long long key = 0; key < (long long)nc * 2; key++) ns += cnt[key] > 0;
    N->nslot[p] = ns;
    N->slot_comp[p] = (int *)calloc(ns + 1, sizeof(int));
    N->slot_rec[p] = (int *)calloc(ns + 1, sizeof(int));
    N->row[p] = (int *)calloc(ns + 1, sizeof(int));
    N->G[p] = (double *)calloc(ns + 1, sizeof(double));
    int *slot_of = (int *)malloc(sizeof(int) * ((size_t)nc * 2 + 1));
    int s = 0;
    for (long long key = 0; key < (long long)nc * 2; key++) {
      slot_of[key] = -1;
      if (cnt[key] > 0) {
        slot_of[key] = s;
        N->slot_comp[p][s] = (int)(key / 2);
        N->slot_rec[p][s] = (int)(key % 2);
        N->row[p][s + 1] = N->row[p][s] + cnt[key];
        s++;
      }
    }
    N->syn[p] = (int *)calloc(N->row[p][ns] + 1, sizeof(int));
    int *fill = (int *)calloc(ns + 1, sizeof(int));
    for (long long j = 0; j < N->nsyn; j++)
      if (post_pop(N->kind[j]) == p) {
        const int sl = slot_of[(size_t)N->post[j] * 2 + receptor(N->kind[j])];
        N->syn[p][N->row[p][sl] + fill[sl]++] = (int)j;
      }
    free(fill);
    free(slot_of);
    free(cnt);
    /* one set per receptor present, attached in receptor order */
    for (int r = 0; r < 2; r++) {
      int n = 0;
      for (int sl = 0; sl < ns; sl++) n += N->slot_rec[p][sl] == r;
      if (n == 0) continue;
      const int kind = p == 0 ? (r == 0 ? OR_SYN_MFGR : OR_SYN_GOGR) : p == 1 ? OR_SYN_GRGO : OR_SYN_GRPKJ;
      or_syn *set = pseudo_set(kind, n, r == 0 ? N->s.e_ampa_mV : N->s.e_gaba_mV);
      N->set_slot[p][r] = (int *)calloc(n, sizeof(int));
      int e = 0;
      for (int sl = 0; sl < ns; sl++)
        if (N->slot_rec[p][sl] == r) {
          set->post[e] = N->slot_comp[p][sl];
          N->set_slot[p][r][e] = sl;
          e++;
        }
      N->set[p][r] = set;
      or_pop_attach_syn(N->pop[p], set);
    }
  }
  N->nring = N->s.delay + 1;
  int off = 0;
  for (int p = 0; p < 3; p++) {
    N->ring_off[p] = off;
    off += N->s.n[p];
    N->prev[p] = (double *)malloc(sizeof(double) * (N->s.n[p] > 0 ? N->s.n[p] : 1));
    for (int c = 0; c < N->s.n[p]; c++) N->prev[p][c] = 100.0;
  }
  N->ring_stride = off;
  N->ring = (unsigned char *)calloc((size_t)N->nring * (off > 0 ? off : 1), 1);
  return N;
}

/* the conductances of interval k: G <- G exp(-dt / tau) + the weights of the
   synapses whose presynaptic spike of interval k - 1 - delay arrives, or whose
   mossy fibre fires in interval k, in synapse order */
void or_net_deliver(or_net *N)
{
  const long k = N->k;
  const long j = k - 1 - N->s.delay;
  const unsigned char *arrive = j >= 0 ? N->ring + (size_t)(j % N->nring) * N->ring_stride : NULL;
  const double decay[2] = {exp(-N->s.dt_ms / N->s.tau_ampa_ms), exp(-N->s.dt_ms / N->s.tau_gaba_ms)};
  const double p_mf = N->s.mf_rate_hz * N->s.dt_ms * 0.001;
  const uint32_t key[2] = {(uint32_t)N->s.seed ^ MF_KEY, (uint32_t)(N->s.seed >> 32)};
  for (int p = 0; p < 3; p++) {
    for (int sl = 0; sl < N->nslot[p]; sl++) {
      double add = 0.0;
      for (int e = N->row[p][sl]; e < N->row[p][sl + 1]; e++) {
        const int s = N->syn[p][e];
        const int pr = pre_pop(N->kind[s]);
        int hit;
        if (pr < 0) {
          const uint32_t ctr[4] = {(uint32_t)N->pre[s], (uint32_t)k, 0u, 0u};
          uint32_t r[4];
          or_philox4x32_10(ctr, key, r);
          hit = p_mf > ((double)r[0] + 1.0) * (1.0 / 4294967296.0);
        } else {
          hit = arrive != NULL && arrive[N->ring_off[pr] + N->pre[s]];
        }
        if (hit) add += N->w[s] * 1e-6;
      }
      N->G[p][sl] = N->G[p][sl] * decay[N->slot_rec[p][sl]] + add;
    }
    for (int r = 0; r < 2; r++)
      if (N->set[p][r])
        for (int e = 0; e < N->set[p][r]->n; e++) N->set[p][r]->val[e] = N->G[p][N->set_slot[p][r][e]];
  }
}

/* after every population advanced interval k: its spikes into ring slot k */
void or_net_mark(or_net *N)
{
  unsigned char *slot = N->ring + (size_t)(N->k % N->nring) * N->ring_stride;
  for (int p = 0; p < 3; p++) {
    const or_pop *P = N->pop[p];
    for (int c = 0; c < N->s.n[p]; c++) {
      const double x = P->vec[OR_VNEW][(size_t)c * P->comp + N->s.spike_comp[p]];
      slot[N->ring_off[p] + c] = (x > N->s.thr_mV[p] && N->prev[p][c] <= N->s.thr_mV[p]) ? 1 : 0;
      N->prev[p][c] = x;
    }
  }
  N->k++;
}

/* nint intervals (deliver, set_current + advance of every population, mark);
   spikes_out, when given, receives [nint][granule, Golgi, Purkinje cells]
   bytes; returns 0 or the first failing population's RKC status */
int or_net_run(or_net *N, int nint, unsigned char *spikes_out)
{
  for (int i = 0; i < nint; i++) {
    or_net_deliver(N);
    const double t = (double)N->k * N->s.dt_ms;
    for (int p = 0; p < 3; p++) {
      or_set_current(N->pop[p], t);
      const int r = or_rkc_advance(N->pop[p]);
      if (r != 0) return r;
    }
    const long k = N->k;
    or_net_mark(N);
    if (spikes_out)
      memcpy(spikes_out + (size_t)i * N->ring_stride, N->ring + (size_t)(k % N->nring) * N->ring_stride,
             N->ring_stride);
  }
  return 0;
}

/* the conductance [mS] of every compartment of a population for a receptor */
void or_net_conductance(const or_net *N, int pop, int rec, double *out)
{
  memset(out, 0, sizeof(double) * (size_t)N->pop[pop]->nc);
  for (int sl = 0; sl < N->nslot[pop]; sl++)
    if (N->slot_rec[pop][sl] == rec) out[N->slot_comp[pop][sl]] = N->G[pop][sl];
}

/* the delivery alone on given spikes (tests): arrive[k] = the spikes of
   interval k - 1 - delay as bytes per (granule, Golgi, Purkinje) cell, or a
   null row; advances the interval counter */
void or_net_deliver_given(or_net *N, const unsigned char *arrive)
{
  if (arrive) {
    const long j = N->k - 1 - N->s.delay;
    if (j >= 0) memcpy(N->ring + (size_t)(j % N->nring) * N->ring_stride, arrive, N->ring_stride);
  }
  or_net_deliver(N);
  N->k++;
}

void or_net_free(or_net *N)
{
  if (!N) return;
  for (int p = 0; p < 3; p++) N->pop[p]->n_syn_in = 0;   /* detach the sets */
  for (int p = 0; p < 3; p++) {
    free(N->slot_comp[p]); free(N->slot_rec[p]); free(N->row[p]); free(N->syn[p]); free(N->G[p]);
    free(N->prev[p]);
    for (int r = 0; r < 2; r++) {
      if (N->set[p][r]) or_syn_free(N->set[p][r]);
      free(N->set_slot[p][r]);
    }
  }
  free(N->kind); free(N->pre); free(N->post); free(N->w); free(N->ring);
  free(N);
}
