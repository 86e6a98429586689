/* Test infrastructure: a short run of every oracle path (RKC, CN, BE and the
 * exact tree solve for the Purkinje, olive, Golgi and granule models, plus the
 * Purkinje cell's synaptic Poisson drive) as a standalone program, so the
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
  return fail;
}
