/* Test infrastructure: a short run of every oracle path (RKC, CN, BE and the
 * exact tree solve for the Purkinje, olive, Golgi and granule models, plus the
 * Purkinje cell's synaptic Poisson drive and the fixed in-degree network) as a
 * standalone program, so the
 * oracle can be built and run under AddressSanitizer / UBSan.
 * usage: oracle_asan <directory holding Pkj.txt Pkj_GJ.csv IO.txt Go_131.txt Gr.txt>
 * Exits nonzero when a run reports an error or produces a non-finite value. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static char dir[4096];
static const char *path(const char *f) {
  static char buf[8][8192];
  static int k = 0;
  k = (k + 1) & 7;
  snprintf(buf[k], sizeof buf[k], "%s/%s", dir, f);
  return buf[k];
}

static int finite_state(or_pop *p, const char *what) {
  const int nc = or_pop_nc(p);
  double *v = malloc(sizeof(double) * (size_t)nc);
  or_pop_get(p, 0, v);
  for (int i = 0; i < nc; i++)
    if (!isfinite(v[i])) {
      fprintf(stderr, "%s: non-finite V at compartment %d\n", what, i);
      free(v);
      return 1;
    }
  free(v);
  printf("%s: ok (%d compartments)\n", what, nc);
  return 0;
}

/* mode 0: RKC, 1: CN Krylov, 2: BE Krylov, 3: CN tree solve */
static int run(or_pop *p, const char *what, int mode, int steps, int stim_comp, double amp) {
  if (!p) {
    fprintf(stderr, "%s: create failed\n", what);
    return 1;
  }
  or_pop_set_threads(p, 1);
  double a[64];
  for (int c = 0; c < 64; c++) a[c] = amp;
  or_set_stim(p, stim_comp, 0.0, 1e300, a);
  int bad = 0;
  if (mode == 1 || mode == 3) bad |= or_cn_init(p) != 0;
  if (mode == 2) bad |= or_be_init(p) != 0;
  if (mode == 3) bad |= or_cn_set_direct(p, 1) != 0;
  for (int k = 0; k < steps && !bad; k++) {
    const double t = k * 0.125;
    if (mode == 0) {
      or_set_current(p, t);
      bad |= or_rkc_advance(p) != 0;
    } else if (mode == 2) {
      bad |= or_be_advance(p, t) != 0;
    } else {
      bad |= or_cn_advance(p, t) != 0;
    }
  }
  if (bad) {
    fprintf(stderr, "%s: the oracle reported an error\n", what);
    or_pop_free(p);
    return 1;
  }
  const int r = finite_state(p, what);
  or_pop_free(p);
  return r;
}

/* the fixed in-degree network (or_net.c) over 2 granule cells, a 2 x 1
   Golgi grid and 1 Purkinje cell: generation, slots, deliveries from the
   ring and from given spikes (every cell firing), mossy-fibre draws, four
   coupled intervals, conductance read-back */
static int run_network(const double *v0) {
  or_pop *gr = or_gr_create(path("Gr.txt"), 2, 1, v0, 1e-9, 1e-9, 0);
  or_pop *go = or_go_create(path("Go_131.txt"), 2, 1, v0, 1e-6, 1e-6, 0);
  or_pop *pk = or_pkj_create(path("Pkj.txt"), NULL, 1, 1, v0, 1e-6, 1e-6, 0);
  if (!gr || !go || !pk) {
    fprintf(stderr, "network: create failed\n");
    return 1;
  }
  double zero[4] = {0.0, 0.0, 0.0, 0.0};
  or_set_stim(gr, 0, 0.0, 0.0, zero);
  or_set_stim(go, 0, 0.0, 0.0, zero);
  or_set_stim(pk, 1599, 0.0, 0.0, zero);
  or_net_spec s;
  memset(&s, 0, sizeof s);
  s.seed = 9;
  s.deg[0] = 4; s.deg[1] = 2; s.deg[2] = 3; s.deg[3] = 5;
  s.total = 4 * 2 + 2 * 2 + 3 * 2 + 5 * 1 + 1;   /* one extra granule input for the Purkinje cell */
  s.w_min_nS = 0.5; s.w_max_nS = 1.5;
  s.tau_ampa_ms = 1.5; s.tau_gaba_ms = 6.0; s.e_ampa_mV = 0.0; s.e_gaba_mV = -82.0;
  s.delay = 2; s.mf_rate_hz = 2000.0; s.dt_ms = 0.125;
  s.tgt_lo[0] = 0; s.tgt_hi[0] = 1; s.tgt_lo[1] = 31; s.tgt_hi[1] = 131; s.tgt_lo[2] = 0; s.tgt_hi[2] = 1599;
  s.spike_comp[0] = 77; s.spike_comp[1] = 30; s.spike_comp[2] = 1599;
  s.thr_mV[0] = 0.0; s.thr_mV[1] = 0.0; s.thr_mV[2] = -10.0;
  or_net *N = or_net_create(gr, go, pk, &s);
  unsigned char spikes[4 * 5], all[5] = {1, 1, 1, 1, 1};
  int bad = or_net_run(N, 4, spikes) != 0;
  for (int k = 0; k < 4; k++) or_net_deliver_given(N, all);
  for (int p = 0; p < 3 && !bad; p++)
    for (int r = 0; r < 2; r++) {
      or_pop *P = p == 0 ? gr : p == 1 ? go : pk;
      double *out = malloc(sizeof(double) * (size_t)or_pop_nc(P));
      or_net_conductance(N, p, r, out);
      for (int i = 0; i < or_pop_nc(P); i++) bad |= !isfinite(out[i]) || out[i] < 0.0;
      free(out);
    }
  or_net_free(N);
  if (bad) fprintf(stderr, "network: the oracle reported an error\n");
  else printf("network: ok (%lld synapses)\n", or_net_size(&s));
  const int r = bad | finite_state(gr, "network granule") | finite_state(go, "network Golgi") |
                finite_state(pk, "network Purkinje");
  or_pop_free(gr);
  or_pop_free(go);
  or_pop_free(pk);
  return r;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s DIR\n", argv[0]);
    return 2;
  }
  snprintf(dir, sizeof dir, "%s", argv[1]);
  double v0[64];
  for (int c = 0; c < 64; c++) v0[c] = -65.0 + 0.25 * c;
  int fail = 0;
  for (int mode = 0; mode < 4; mode++) {
    char name[64];
    const char *m[4] = {"RKC", "CN", "BE", "CN tree"};
    snprintf(name, sizeof name, "PKJ 1x1 %s", m[mode]);
    fail |= run(or_pkj_create(path("Pkj.txt"), NULL, 1, 1, v0, 1e-6, 1e-6, 0), name, mode, 4, 1599, 1e-3);
    snprintf(name, sizeof name, "IO 2x2 %s", m[mode]);
    fail |= run(or_io_create(path("IO.txt"), 2, 2, v0, 1e-6, 1e-6, 0), name, mode, 8, 0, 5.0);
    snprintf(name, sizeof name, "GR 1x1 %s", m[mode]);
    fail |= run(or_gr_create(path("Gr.txt"), 1, 1, v0, 1e-9, 1e-9, 0), name, mode, 2, 1, 0.01e-3);
    if (mode != 3) {   /* a Golgi grid's junctions are not a tree */
      snprintf(name, sizeof name, "GO 2x2 %s", m[mode]);
      fail |= run(or_go_create(path("Go_131.txt"), 2, 2, v0, 1e-6, 1e-6, 0), name, mode, 4, 30, 0.2e-3);
    }
  }
  /* Purkinje grid with its gap-junction pairs */
  fail |= run(or_pkj_create(path("Pkj.txt"), path("Pkj_GJ.csv"), 2, 1, v0, 1e-6, 1e-6, 0), "PKJ 2x1 GJ RKC", 0, 2,
              1599, 1e-3);
  fail |= run_network(v0);
  return fail;
}
