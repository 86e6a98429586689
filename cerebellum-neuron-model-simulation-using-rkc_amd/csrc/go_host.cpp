// Golgi-cell population builder (host).  Built with -ffp-contract=off so the
// initial state and the cable conductances are the reference's expressions.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <tuple>

#include "go_model.hpp"
#include "morph.hpp"

namespace cbrkc {

bool build_go_host(const std::string &param_path, int nx, int ny, const double *v0, GoHost &H, std::string &err,
                   int comp, bool grid_gj) {
  using namespace go;
  std::FILE *f = std::fopen(param_path.c_str(), "r");
  if (!f) {
    err = "no such file " + param_path;
    return false;
  }
  const int n = nx * ny, nc = n * comp;
  if (comp < 1 || comp > 4096) {
    std::fclose(f);
    err = "Golgi compartment count out of range";
    return false;
  }
  // go.cu:30-63 / go_solve.cu:21-38: idx parent diam len Ra Cm type + 13
  // densities; one block of comp rows for every cell, or one block per cell
  // (idx and parent local to the block)
  std::vector<double> rad, len, Ra, cmd, dens;
  std::vector<int> compart, l_connect;
  std::vector<char> seen;
  int nblk = 0;
  for (;; nblk++) {
    const size_t o = (size_t)nblk * comp;
    rad.resize(o + comp, 0.0), len.resize(o + comp, 0.0), Ra.resize(o + comp, 0.0), cmd.resize(o + comp, 0.0);
    dens.resize((o + comp) * NCOND, 0.0);
    compart.resize(o + comp, 0), l_connect.resize(o + comp, -1), seen.resize(o + comp, 0);
    for (int i = 0; i < comp; i++) {
      int i1, i2, i3;
      double d1, d2, d3, d4, x[NCOND];
      const int got = std::fscanf(f, "%d %d %lf %lf %lf %lf %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf",
                                  &i1, &i2, &d1, &d2, &d3, &d4, &i3, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6],
                                  &x[7], &x[8], &x[9], &x[10], &x[11], &x[12]);
      if (got == EOF && i == 0 && nblk > 0) goto done;
      if (got != 20 || i1 < 0 || i1 >= comp || i2 < -1 || i2 >= comp) {
        std::fclose(f);
        err = "GO_PARAM_FILE_READING_ERROR in " + param_path;
        return false;
      }
      rad[o + i1] = 0.5 * d1 * 1e-4;   // um -> cm
      len[o + i1] = d2 * 1e-4;
      Ra[o + i1] = d3 * 1e-3;          // kOhm cm
      cmd[o + i1] = d4;
      compart[o + i1] = i3;
      seen[o + i1] = 1;
      for (int q = 0; q < NCOND; q++) dens[(o + i1) * NCOND + q] = x[q];
      l_connect[o + i] = i2;           // connected by line, go_solve.cu:37
    }
  }
done:
  std::fclose(f);
  if (nblk != 1 && nblk != n) {
    err = "GO_PARAM_FILE_READING_ERROR: " + std::to_string(nblk) + " blocks of " + std::to_string(comp) +
          " rows in " + param_path + " for " + std::to_string(n) + " cells";
    return false;
  }
  for (size_t i = 0; i < (size_t)nblk * comp; i++)
    if (!seen[i]) {
      err = "GO_PARAM_FILE_READING_ERROR: compartment missing in " + param_path;
      return false;
    }

  H.n = n;
  H.nx = nx;
  H.ny = ny;
  H.comp = comp;
  H.nc = nc;
  H.cm.assign(nc, 0.0);
  H.area.assign(nc, 0.0);
  H.iext.assign(nc, 0.0);
  H.v.assign(nc, 0.0);
  H.ca.assign(nc, CA1_0);
  H.ca2.assign(nc, CA1_0);
  H.rev.assign(nc, V_CA);
  H.ca_old.assign(nc, CA1_0);
  H.cond.assign((size_t)NCOND * nc, 0.0);
  H.gate.assign((size_t)NGATE * nc, 0.0);
  // go.cu:47-62: cylinder area, Cm and conductances as density x area
  for (int c = 0; c < n; c++)
    for (int k = 0; k < comp; k++) {
      const size_t i = (size_t)c * comp + k, o = (size_t)(nblk == 1 ? 0 : c) * comp + k;
      const double a = 2.0 * M_PI * rad[o] * len[o];
      H.area[i] = a;
      H.cm[i] = cmd[o] * a;
      for (int q = 0; q < NCOND; q++) H.cond[(size_t)q * nc + i] = dens[o * NCOND + q] * a;
    }

  // grid gap junctions between compartments 0 (go_solve.cu:52-78): x
  // neighbours row by row, then y neighbours column by column, both directions
  std::vector<int> gpre, gpost;
  if (n > 1 && grid_gj) {
    for (int y = 0; y < ny; y++)
      for (int x = 0; x + 1 < nx; x++) {
        gpre.push_back((y * nx + x) * comp);
        gpre.push_back((y * nx + x + 1) * comp);
      }
    for (int x = 0; x < nx; x++)
      for (int y = 0; y + 1 < ny; y++) {
        gpre.push_back((y * nx + x) * comp);
        gpre.push_back(((y + 1) * nx + x) * comp);
      }
    gpost.resize(gpre.size());
    for (size_t i = 0; i < gpre.size(); i += 2) {
      gpost[i + 1] = gpre[i];
      gpost[i] = gpre[i + 1];
    }
  }
  H.n_gj = (int)gpre.size();
  // gap-junction partners per cell (every junction sits on compartment 0)
  std::vector<std::vector<std::pair<int, double>>> extra(gpre.empty() ? 0 : n);
  for (size_t i = 0; i < gpre.size(); i++) extra[gpre[i] / comp].push_back({gpost[i], G_GJ});
  // CSR (go_solve.cu:108-213): each row holds its in-cell entries, its
  // diagonal and its gap-junction partners in ascending column order; RKC
  // sign: off-diagonal +g, diagonal -(row sum).  Cable conductance of each
  // connected pair, go_solve.cu:40-46 (row i, column d), one block at a time.
  std::vector<double> mat((size_t)comp * comp, 0.0), dsum(comp, 0.0);
  H.row.assign(nc + 1, 0);
  H.col.clear();
  H.val.clear();
  for (int cell = 0; cell < n; cell++) {
    const size_t ob = (size_t)(nblk == 1 ? 0 : cell) * comp;
    if (cell == 0 || nblk > 1) {
      std::fill(mat.begin(), mat.end(), 0.0);
      for (int i = 0; i < comp; i++) {
        const int d = l_connect[ob + i];
        if (d >= 0) {
          const size_t pi = ob + i, pd = ob + d;
          mat[d + (size_t)comp * i] = (2.0 / ((Ra[pi] * len[pi]) / (rad[pi] * rad[pi] * M_PI) +
                                              (Ra[pd] * len[pd]) / (rad[pd] * rad[pd] * M_PI)));
          mat[i + (size_t)comp * d] = mat[d + (size_t)comp * i];
        }
      }
      std::fill(dsum.begin(), dsum.end(), 0.0);
      for (int r = 0; r < comp; r++)
        for (int c = 0; c < comp; c++)
          if (c != r) dsum[r] += mat[c + (size_t)comp * r];
    }
    for (int r = 0; r < comp; r++) {
      const int gr = cell * comp + r;
      std::vector<std::pair<int, double>> ent;
      for (int c = 0; c < comp; c++) {
        const double m = mat[c + (size_t)comp * r];
        if (m > 0.0) {
          ent.push_back({cell * comp + c, m});
        } else if (c == r) {
          double dv = -dsum[r];
          if (!extra.empty() && r == 0)
            for (size_t q = 0; q < extra[cell].size(); q++) dv -= G_GJ;   // val[dig] += G_GJ_GO per junction
          ent.push_back({gr, dv});
        }
      }
      if (!extra.empty() && r == 0)
        for (auto &e : extra[cell]) ent.push_back(e);
      std::stable_sort(ent.begin(), ent.end(),
                       [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
      for (auto &e : ent) {
        H.col.push_back(e.first);
        H.val.push_back(e.second);
      }
      H.row[gr + 1] = (int)H.col.size();
    }
  }

  // initial state, go_ion.cu:590-634
  double init_v_rand = 0.0;
  double g[NGATE];
  for (int i = 0; i < nc; i++) {
    if (i % comp == 0 && !v0) init_v_rand = ((double)std::rand() / RAND_MAX) - 0.5;
    const double vv = v0 ? v0[i / comp] : V_INIT + 2.0 * init_v_rand;
    H.v[i] = vv;
    steady_state(vv, H.ca[i], g);
    for (int q = 0; q < NGATE; q++) H.gate[(size_t)q * nc + i] = g[q];
  }
  return true;
}

}  // namespace cbrkc
