// Host-side population builder for the Purkinje RKC path.
//   morphology parsing      pkj.cu:27-38, pkj_solve.cu:43-66
//   areas, Cm, shell, g     pkj.cu:40-140
//   cable conductance       pkj_solve.cu:68-74
//   gap junction grid       pkj_solve.cu:79-120, 200-226
//   initial state           pkj_ion.cu:360-395
#include "morph.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <thread>

#include "philox.hpp"
#include "pkj_model.hpp"

namespace cbrkc {

bool read_pkj_morph(const std::string &path, int comp, std::vector<MorphRow> &rows, std::string &err) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) {
    err = "no such file " + path;
    return false;
  }
  rows.clear();
  MorphRow r;
  while (std::fscanf(f, "%d %d %lf %lf %lf %lf %d ", &r.idx, &r.parent, &r.x, &r.y, &r.z, &r.d, &r.type) == 7)
    rows.push_back(r);
  std::fclose(f);
  if (rows.empty() || rows.size() % (size_t)comp != 0) {
    err = "morphology " + path + " has " + std::to_string(rows.size()) + " rows, not a multiple of " +
          std::to_string(comp);
    return false;
  }
  for (size_t b = 0; b < rows.size(); b += comp)
    for (int k = 0; k < comp; k++) {
      const MorphRow &q = rows[b + k];
      if (q.idx < 0 || q.idx >= comp || q.parent >= comp || q.parent < -1) {
        err = "morphology " + path + ": bad index on row " + std::to_string(b + k);
        return false;
      }
    }
  return true;
}

bool read_pkj_gj(const std::string &path, std::vector<int> &pre, std::vector<int> &post, std::string &err) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) {
    err = "no such file " + path;
    return false;
  }
  pre.clear();
  post.clear();
  for (int i = 0; i < pkj::GJ_EACH; i++) {
    int a, b;
    if (std::fscanf(f, "%d, %d", &a, &b) != 2) {
      std::fclose(f);
      err = "gap junction file " + path + " is short";
      return false;
    }
    pre.push_back(a);
    post.push_back(b);
  }
  std::fclose(f);
  return true;
}

bool write_pkj_morph(const std::string &path, const std::vector<MorphRow> &rows, std::string &err) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot write " + path;
    return false;
  }
  for (const MorphRow &r : rows)
    std::fprintf(f, "%d\t%d\t%.6f\t%.6f\t%.6f\t%.6f\t%d\n", r.idx, r.parent, r.x, r.y, r.z, r.d, r.type);
  std::fclose(f);
  return true;
}

void synth_pkj_tree(uint64_t seed, int comp, double soma_r_um, double lmin, double lmax, double d0, double d1,
                    double dscale, std::vector<MorphRow> &out) {
  // Growth process in creation order (creation index 0 = soma):
  //   pick a growing tip uniformly; with probability 0.25 it bifurcates into
  //   two children, otherwise it extends by one compartment.
  PhiloxStream rng(seed, 0x7472u);  // stream "tr"
  const int ndend = comp - 1;
  std::vector<int> par(comp, -1), depth(comp, 0);
  std::vector<double> len(comp, 0.0);
  std::vector<int> tips;
  int made = 1;
  par[1] = 0;
  depth[1] = 1;
  made = 2;
  tips.push_back(1);
  while (made < comp) {
    int ti = (int)(rng.uniform() * tips.size());
    if (ti >= (int)tips.size()) ti = (int)tips.size() - 1;
    int tip = tips[ti];
    bool branch = rng.uniform() < 0.25 && made + 2 <= comp;
    int nchild = branch ? 2 : 1;
    tips[ti] = tips.back();
    tips.pop_back();
    for (int c = 0; c < nchild; c++) {
      par[made] = tip;
      depth[made] = depth[tip] + 1;
      tips.push_back(made);
      made++;
    }
  }
  int maxd = 1;
  for (int i = 1; i < comp; i++) maxd = std::max(maxd, depth[i]);
  out.assign(comp, MorphRow{});
  for (int i = 0; i < comp; i++) {
    int idx = comp - 1 - i;  // parents get higher indices, soma is last
    MorphRow &r = out[idx];
    r.idx = idx;
    r.parent = (par[i] < 0) ? -1 : comp - 1 - par[i];
    if (i == 0) {
      r.x = r.y = r.z = 0.0;
      r.d = 2.0 * soma_r_um;
      r.type = pkj::SOMA;
      continue;
    }
    double L = lmin + (lmax - lmin) * rng.uniform();
    double cz = 2.0 * rng.uniform() - 1.0, ph = 2.0 * M_PI * rng.uniform();
    double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
    r.x = L * sz * std::cos(ph);
    r.y = L * sz * std::sin(ph);
    r.z = L * cz;
    double frac = (maxd > 1) ? (double)(depth[i] - 1) / (double)(maxd - 1) : 0.0;
    r.d = (d0 + (d1 - d0) * frac) * dscale;
    if (depth[i] <= 2)
      r.type = pkj::MAIN;
    else if (d0 + (d1 - d0) * frac >= 2.0)
      r.type = pkj::SMOOTH;
    else
      r.type = pkj::SPINY;
  }
  (void)ndend;
}

bool build_pkj_host(const std::vector<MorphRow> &rows, int nx, int ny, const std::vector<int> *gj_pre,
                    const std::vector<int> *gj_post, const double *v0, PkjHost &P, std::string &err) {
  using namespace pkj;
  const int C = COMP;
  const int n = nx * ny;
  const int nblocks = (int)(rows.size() / C);
  if (nblocks != 1 && nblocks != n) {
    err = "morphology holds " + std::to_string(nblocks) + " cells but the population has " + std::to_string(n);
    return false;
  }
  P.n = n;
  P.nx = nx;
  P.ny = ny;
  P.comp = C;
  const int nc = n * C;
  P.nc = nc;
  P.cm.assign(nc, 0.0);
  P.area.assign(nc, 0.0);
  P.shell.assign(nc, 0.0);
  P.iext.assign(nc, 0.0);
  P.v.assign(nc, 0.0);
  P.ca.assign(nc, 0.0);
  P.rev.assign(nc, 0.0);
  P.compart.assign(nc, 0);
  P.cond.assign((size_t)NCOND * nc, 0.0);
  P.gate.assign((size_t)NGATE * nc, 0.0);
  P.parent.assign(nc, -1);
  P.gpar.assign(nc, 0.0);
  std::vector<double> rad(nc), len(nc);

  for (int c = 0; c < n; c++) {
    const MorphRow *blk = rows.data() + (size_t)(nblocks == 1 ? 0 : c) * C;
    for (int k = 0; k < C; k++) {
      const MorphRow &q = blk[k];
      const int gi = c * C + q.idx;
      const double d = q.d;
      double a;
      rad[gi] = 0.5 * d * 1e-4;  // um -> cm
      P.compart[gi] = (int8_t)q.type;
      P.parent[gi] = (q.parent >= 0) ? c * C + q.parent : -1;
      if (q.type == SOMA) {
        len[gi] = d * 1e-4;
        a = 4.0 * M_PI * rad[gi] * rad[gi];  // sphere
        P.area[gi] = a;
        P.cm[gi] = CM * a;
        P.shell[gi] = (SHELLD - 2 * std::pow(SHELLD, 2) / (d * 1e-4) + 4 * std::pow(SHELLD, 3) / (3 * std::pow(d * 1e-4, 2))) * a;
      } else {
        len[gi] = std::sqrt(std::pow(q.x, 2) + std::pow(q.y, 2) + std::pow(q.z, 2)) * 1e-4;
        a = 2.0 * M_PI * rad[gi] * len[gi];  // cylinder
        P.area[gi] = a;
        P.shell[gi] = (SHELLD - std::pow(SHELLD, 2) / (d * 1e-4)) * a;
        P.cm[gi] = (q.type == SPINY) ? CM * (a + len[gi] * SPINE_AREA * SPINE_DENS) : CM * a;
      }
      const int cls = (q.type == SOMA) ? 0 : (q.type == MAIN) ? 1 : 2;
      const double leak_area = (q.type == SPINY) ? a + len[gi] * SPINE_AREA * SPINE_DENS : a;
      P.cond[(size_t)g_leak * nc + gi] = 1.0 / ((q.type == SOMA) ? RM_S : RM_D) * leak_area;
      for (int g = 0; g < 10; g++) P.cond[(size_t)(g_NaF + g) * nc + gi] = DENSITY[cls][g] * a;
    }
  }

  // axial conductances child <-> parent
  for (int i = 0; i < nc; i++) {
    int d = P.parent[i];
    if (d < 0) continue;
    P.gpar[i] = 2.0 / ((RI * len[i]) / (rad[i] * rad[i] * M_PI) + (RI * len[d]) / (rad[d] * rad[d] * M_PI));
  }

  // gap junctions between horizontally adjacent cells (only for a grid)
  std::vector<std::pair<int, int>> gj;  // directed (row, col) entries
  if (n > 1 && gj_pre && gj_post && !gj_pre->empty()) {
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx - 1; x++)
        for (size_t i = 0; i < gj_pre->size(); i++) {
          int a = (*gj_pre)[i] + (x + y * nx) * C;
          int b = (*gj_post)[i] + ((x + 1) + y * nx) * C;
          gj.push_back({a, b});
          gj.push_back({b, a});
        }
  }
  P.n_gj = (int)gj.size();

  // CSR rows: neighbours and the stored diagonal, sorted by column
  std::vector<std::vector<std::pair<int, double>>> adj(nc);
  for (int i = 0; i < nc; i++) {
    int d = P.parent[i];
    if (d < 0) continue;
    adj[i].push_back({d, P.gpar[i]});
    adj[d].push_back({i, P.gpar[i]});
  }
  std::vector<int> ngj(nc, 0);
  for (auto &e : gj) {
    for (auto &q : adj[e.first])
      if (q.first == e.second) {
        err = "duplicate Purkinje gap junction " + std::to_string(e.first) + " -> " + std::to_string(e.second);
        return false;
      }
    adj[e.first].push_back({e.second, G_GJ});
    ngj[e.first]++;
  }
  P.row.assign(nc + 1, 0);
  P.col.clear();
  P.val.clear();
  for (int i = 0; i < nc; i++) {
    auto &a = adj[i];
    a.push_back({i, 0.0});
    std::sort(a.begin(), a.end(), [](const std::pair<int, double> &l, const std::pair<int, double> &r) {
      return l.first < r.first;
    });
    // diagonal magnitude: cable conductances in column order, then one G_GJ per junction
    double s = 0.0;
    for (auto &q : a)
      if (q.first != i && q.first / C == i / C) s += q.second;
    for (int k = 0; k < ngj[i]; k++) s += G_GJ;
    for (auto &q : a) {
      P.col.push_back(q.first);
      P.val.push_back(q.first == i ? -s : q.second);
    }
    P.row[i + 1] = (int)P.col.size();
  }

  // initial state
  double jitter = 0.0;
  for (int i = 0; i < nc; i++) {
    if (!v0 && i % C == 0) jitter = (((double)std::rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 0.5;
    P.v[i] = v0 ? v0[i / C] : V_INIT + 2.0 * jitter;
    P.ca[i] = CA_0;
    P.rev[i] = V_CA;
    P.iext[i] = 0.0;
    double g[NGATE];
    steady_state<double>(P.v[i], P.ca[i], g);
    for (int k = 0; k < NGATE; k++) P.gate[(size_t)k * nc + i] = g[k];
  }
  return true;
}

namespace {

// largest eigenvalue of the symmetric tridiagonal (a[0..k), b[0..k-1)) by
// bisection on Sturm counts
double tridiag_max_eig(const std::vector<double> &a, const std::vector<double> &b) {
  const int k = (int)a.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; i++) {
    double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {   // eigenvalues < x
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < k; i++) {
      double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      d = a[i] - x - (i > 0 ? bb / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) c++;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-13 * std::max(1.0, std::fabs(hi)); it++) {
    double mid = 0.5 * (lo + hi);
    if (count_below(mid) == k) hi = mid;
    else lo = mid;
  }
  return hi;
}

}  // namespace

void cable_spectral_radius(const PkjHost &P, std::vector<double> &rho) {
  const int C = P.comp;
  rho.assign(P.n, 0.0);
  const int kmax = std::min(C, 80);
  const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<int> next{0};
  auto worker = [&]() {
   for (int c = next++; c < P.n; c = next++) {
    const int base = c * C;
    std::vector<double> isq(C);
    for (int i = 0; i < C; i++) isq[i] = 1.0 / std::sqrt(P.cm[base + i]);
    // y = S x, S = C^-1/2 (-L) C^-1/2, (-L x)_i = sum_j g_ij (x_i - x_j) within the cell
    auto apply = [&](const std::vector<double> &x, std::vector<double> &y) {
      for (int i = 0; i < C; i++) {
        const int gi = base + i;
        double s = 0.0, xi = x[i] * isq[i];
        for (int e = P.row[gi]; e < P.row[gi + 1]; e++) {
          const int j = P.col[e] - base;
          if (j == i || j < 0 || j >= C) continue;
          s += P.val[e] * (xi - x[j] * isq[j]);
        }
        y[i] = s * isq[i];
      }
    };
    std::vector<std::vector<double>> Q;
    std::vector<double> al, be, q(C), w(C);
    double nrm = 0.0;
    for (int i = 0; i < C; i++) {
      q[i] = ((i & 1) ? -1.0 : 1.0) * (1.0 + 0.1 * (double)((i * 7919) % 13) / 13.0);
      nrm += q[i] * q[i];
    }
    nrm = std::sqrt(nrm);
    for (double &x : q) x /= nrm;
    double bprev = 0.0;
    for (int k = 0; k < kmax; k++) {
      Q.push_back(q);
      apply(q, w);
      double a = 0.0;
      for (int i = 0; i < C; i++) a += q[i] * w[i];
      al.push_back(a);
      for (int i = 0; i < C; i++) w[i] -= a * q[i] + (k > 0 ? bprev * Q[k - 1][i] : 0.0);
      for (int pass = 0; pass < 2; pass++)   // full reorthogonalisation
        for (const auto &v : Q) {
          double d = 0.0;
          for (int i = 0; i < C; i++) d += v[i] * w[i];
          for (int i = 0; i < C; i++) w[i] -= d * v[i];
        }
      double bnew = 0.0;
      for (double x : w) bnew += x * x;
      bnew = std::sqrt(bnew);
      if (bnew < 1e-12 * std::fabs(a) || k + 1 == kmax) break;
      be.push_back(bnew);
      for (int i = 0; i < C; i++) q[i] = w[i] / bnew;
      bprev = bnew;
    }
    be.resize(al.size() - 1);
    rho[c] = tridiag_max_eig(al, be);
   }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; t++) pool.emplace_back(worker);
  for (auto &t : pool) t.join();
}

}  // namespace cbrkc
