// Inferior-olive population builder (host).  Built with -ffp-contract=off so
// the initial state and the coupling values are the reference's expressions.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "io_model.hpp"
#include "morph.hpp"

namespace cbrkc {

bool build_io_host(const std::string &param_path, int nx, int ny, const double *v0, IoHost &H, std::string &err) {
  using namespace io;
  std::FILE *f = std::fopen(param_path.c_str(), "r");
  if (!f) {
    err = "no such file " + param_path;
    return false;
  }
  int connect[COMP], compart[COMP];
  double cmv[COMP];
  for (int k = 0; k < COMP; k++) {
    int i1, i2, i3;
    double d1, d2, d4;
    if (std::fscanf(f, "%d %d %lf %lf %lf %d ", &i1, &i2, &d1, &d2, &d4, &i3) != 6 || i1 < 0 || i1 >= COMP) {
      std::fclose(f);
      err = "IO_PARAM_FILE_READING_ERROR in " + param_path;
      return false;
    }
    connect[i1] = i2;
    compart[i1] = i3;
    cmv[i1] = d4;
    if (i3 < 0 || i3 > 2) {
      std::fclose(f);
      err = "io.cu error: compartment type out of range";
      return false;
    }
  }
  std::fclose(f);

  const int n = nx * ny, nc = n * COMP;
  H.n = n;
  H.nx = nx;
  H.ny = ny;
  H.comp = COMP;
  H.nc = nc;
  H.cm.assign(nc, 0.0);
  H.iext.assign(nc, 0.0);
  H.v.assign(nc, 0.0);
  H.ca.assign(nc, 0.0);
  H.compart.assign(nc, 0);
  H.cond.assign((size_t)NCOND * nc, 0.0);
  H.gate.assign((size_t)NGATE * nc, 0.0);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < COMP; k++) {
      const int i = c * COMP + k, t = compart[k];
      H.compart[i] = (int8_t)t;
      H.cm[i] = cmv[k];
      double *g = &H.cond[i];
      auto set = [&](int q, double x) { g[(size_t)q * nc] = x; };
      set(g_leak, G_LEAK);
      if (t == DEND) {
        set(g_CaH, DNDR_G_CAH);
        set(g_KCa, DNDR_G_KCA);
        set(g_H, DNDR_G_H);
      } else if (t == SOMA) {
        set(g_CaL, SOMA_G_CAL);
        set(g_Na, SOMA_G_NA);
        set(g_Kdr, SOMA_G_KDR);
        set(g_K, SOMA_G_K);
      } else {
        set(g_Na, AXON_G_NA);
        set(g_K, AXON_G_K);
      }
    }

  // coupling of one cell (io_solve.cu:36-66): row i, column d
  double mat[COMP * COMP] = {0};
  for (int i = 0; i < COMP; i++) {
    const int d = connect[i], i3 = compart[i];
    if (d < 0) continue;
    double lg, lg_rev;
    if (i3 == 0 && d == 1) {
      lg = G_INT / (1.0 - G_P_DS);
      lg_rev = G_INT / G_P_DS;
    } else if (i3 == 1 && d == 2) {
      lg = G_INT / G_P_SA;
      lg_rev = G_INT / (1.0 - G_P_SA);
    } else if (i3 == 0 && d < 1) {
      lg = lg_rev = G_INT / 0.5;
    } else {
      err = "io_solve Error";
      return false;
    }
    mat[d + COMP * i] = lg;
    mat[i + COMP * d] = lg_rev;
  }
  // CSR per cell, columns ascending, stored diagonal -(row sum) (RKC sign)
  H.row.assign(nc + 1, 0);
  H.col.clear();
  H.val.clear();
  for (int cell = 0; cell < n; cell++)
    for (int r = 0; r < COMP; r++) {
      double dsum = 0.0;
      for (int c = 0; c < COMP; c++)
        if (c != r) dsum += mat[c + COMP * r];
      for (int c = 0; c < COMP; c++) {
        if (c == r) {
          H.col.push_back(cell * COMP + c);
          H.val.push_back(-dsum);
        } else if (mat[c + COMP * r] > 0.0) {
          H.col.push_back(cell * COMP + c);
          H.val.push_back(mat[c + COMP * r]);
        }
      }
      H.row[cell * COMP + r + 1] = (int)H.col.size();
    }

  // initial state, io_ion.cu:163-194
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % COMP == 0 && !v0) init_v_rand = ((double)std::rand() / RAND_MAX) - 0.5;
    const double vv = v0 ? v0[i / COMP] : V_INIT + 10.0 * init_v_rand;
    const int lc = H.compart[i];
    double *g = &H.gate[i];
    auto gate = [&](int q) -> double & { return g[(size_t)q * nc]; };
    H.v[i] = vv;
    gate(k_CaL) = k_inf(vv);
    gate(l_CaL) = l_inf(vv);
    gate(n_Kdr) = n_inf(vv);
    gate(p_Kdr) = p_inf(vv);
    gate(x_K) = x_inf(vv);
    gate(r_CaH) = r_inf(vv);
    gate(q_H) = q_inf(vv);
    const double I_Ca1 = H.cond[(size_t)g_CaH * nc + i] * gate(r_CaH) * gate(r_CaH) * (vv - V_CA);
    H.ca[i] = -I_Ca1 * CA_PH / BETA_CA;
    gate(s_KCa) = s_inf(H.ca[i]);
    gate(m_Na) = m_inf(vv);
    gate(h_Na) = h_inf(vv, lc);
  }
  // partners outside the grid (non-square grids, where the reference's
  // id / ny row reaches past its arrays) are treated as absent
  H.nbr.assign((size_t)n * 4, -1);
  if (n > 1)
    for (int c = 0; c < n; c++) {
      int *o = &H.nbr[(size_t)c * 4];
      neighbours(c, nx, ny, o);
      for (int s = 0; s < 4; s++)
        if (o[s] < 0 || o[s] >= n) o[s] = -1;
    }
  return true;
}

}  // namespace cbrkc
