// Granule-cell population builder (host).  Built with -ffp-contract=off so the
// initial state and the cable conductances are the reference's expressions.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "gr_model.hpp"
#include "morph.hpp"

namespace cbrkc {
namespace gr {

// gr_ion.cu:57-69: types below GR_hill carry placeholder ones (no Na scheme)
NaTable na_table() {
  const double n1[] = {1.0, 1.0, 1.0, 1.0, 1.0, 5.422, 5.422, 5.422, 5.422};
  const double n2[] = {1.0, 1.0, 1.0, 1.0, 1.0, 3.279, 3.279, 3.279, 3.279};
  const double n3[] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.83, 1.83, 1.83, 1.83};
  const double n4[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.738, 0.738, 0.738, 0.738};
  const double Aalpha[] = {1.0, 1.0, 1.0, 1.0, 1.0, 8.85, 8.85, 8.85, 8.85};
  const double Abeta[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.0318, 0.0318, 0.0318, 0.0318};
  const double Vshift[] = {1.0, 1.0, 1.0, 1.0, 1.0, -22.0, -22.0, -25.0, -25.0};
  const double Con[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.003464, 0.003464, 0.09353, 0.09353};
  const double Coff[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.4157, 0.4157, 0.04676, 0.04676};
  const double Oon[] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.8186, 1.8186, 2.5981, 2.5981};
  const double Ooff[] = {1.0, 1.0, 1.0, 1.0, 1.0, 0.006928, 0.006928, 0.000173, 0.000173};
  NaTable T;
  for (int t = 0; t < NTYPE; t++) {
    NaType &x = T.t[t];
    x.n1 = n1[t];
    x.n2 = n2[t];
    x.n3 = n3[t];
    x.n4 = n4[t];
    x.Aalpha = Aalpha[t];
    x.Abeta = Abeta[t];
    x.Vshift = Vshift[t];
    x.Con = Con[t];
    x.Coff = Coff[t];
    x.Oon = Oon[t];
    x.Ooff = Ooff[t];
    x.a = std::pow((Oon[t] / Con[t]), 0.25);
    x.b = std::pow((Ooff[t] / Coff[t]), 0.25);
    x.a4 = std::pow(x.a, 4.0);
    x.b4 = std::pow(x.b, 4.0);
  }
  return T;
}

}  // namespace gr

bool build_gr_host(const std::string &param_path, int nx, int ny, const double *v0, GrHost &H, std::string &err,
                   int comp) {
  using namespace gr;
  std::FILE *f = std::fopen(param_path.c_str(), "r");
  if (!f) {
    err = "no such file " + param_path;
    return false;
  }
  const int n = nx * ny, nc = n * comp;
  if (comp < 1 || comp > COMP) {
    std::fclose(f);
    err = "granule compartment count out of range";
    return false;
  }
  // gr.cu:29-58 / gr_solve.cu:20-35: idx parent diam len Ra Cm type + 10
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
      const int got = std::fscanf(f, "%d %d %lf %lf %lf %lf %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf ", &i1, &i2, &d1,
                                  &d2, &d3, &d4, &i3, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6], &x[7], &x[8],
                                  &x[9]);
      if (got == EOF && i == 0 && nblk > 0) goto done;
      if (got != 17 || i1 < 0 || i1 >= comp || i2 < -1 || i2 >= comp || i3 < 0 || i3 >= NTYPE) {
        std::fclose(f);
        err = "GR_PARAM_FILE_READING_ERROR in " + param_path;
        return false;
      }
      rad[o + i1] = 0.5 * d1 * 1e-4;
      len[o + i1] = d2 * 1e-4;
      Ra[o + i1] = d3 * 1e-3;
      cmd[o + i1] = d4;
      compart[o + i1] = i3;
      seen[o + i1] = 1;
      for (int q = 0; q < NCOND; q++) dens[(o + i1) * NCOND + q] = x[q];
      l_connect[o + i] = i2;   // connected by line, gr_solve.cu:34
    }
  }
done:
  std::fclose(f);
  if (nblk != 1 && nblk != n) {
    err = "GR_PARAM_FILE_READING_ERROR: " + std::to_string(nblk) + " blocks of " + std::to_string(comp) +
          " rows in " + param_path + " for " + std::to_string(n) + " cells";
    return false;
  }
  for (size_t i = 0; i < (size_t)nblk * comp; i++)
    if (!seen[i]) {
      err = "GR_PARAM_FILE_READING_ERROR: compartment missing in " + param_path;
      return false;
    }

  H.n = n;
  H.nx = nx;
  H.ny = ny;
  H.comp = comp;
  H.nc = nc;
  H.cm.assign(nc, 0.0);
  H.iext.assign(nc, 0.0);
  H.v.assign(nc, 0.0);
  H.ca.assign(nc, CA1_0);
  H.compart.assign(nc, 0);
  H.cond.assign((size_t)NCOND * nc, 0.0);
  H.gate.assign((size_t)NGATE * nc, 0.0);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < comp; k++) {
      const size_t i = (size_t)c * comp + k, o = (size_t)(nblk == 1 ? 0 : c) * comp + k;
      const double a = 2.0 * M_PI * rad[o] * len[o];   // gr.cu:44-57
      H.cm[i] = cmd[o] * a;
      H.compart[i] = (int8_t)compart[o];
      for (int q = 0; q < NCOND; q++) H.cond[(size_t)q * nc + i] = dens[o * NCOND + q] * a;
    }
  // cable conductances (gr_solve.cu:37-43), CSR per row in column order with
  // the stored diagonal, RKC sign (gr_solve.cu:73-133), one block at a time
  H.row.assign(nc + 1, 0);
  H.col.clear();
  H.val.clear();
  std::vector<double> mat((size_t)comp * comp, 0.0);
  std::vector<int> rcol, rptr(comp + 1, 0);
  std::vector<double> rval;
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
      rcol.clear();
      rval.clear();
      for (int r = 0; r < comp; r++) {
        double dsum = 0.0;
        for (int c = 0; c < comp; c++)
          if (c != r) dsum += mat[c + (size_t)comp * r];
        for (int c = 0; c < comp; c++) {
          const double m = mat[c + (size_t)comp * r];
          if (m > 0.0) {
            rcol.push_back(c);
            rval.push_back(m);
          } else if (c == r) {
            rcol.push_back(c);
            rval.push_back(-dsum);
          }
        }
        rptr[r + 1] = (int)rcol.size();
      }
    }
    const size_t base = H.col.size();
    for (size_t j = 0; j < rcol.size(); j++) {
      H.col.push_back(rcol[j] + cell * comp);
      H.val.push_back(rval[j]);
    }
    for (int r = 0; r < comp; r++) H.row[cell * comp + r + 1] = (int)(base + rptr[r + 1]);
  }

  // initial state, gr_ion.cu:635-672
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % comp == 0 && !v0) init_v_rand = (((double)std::rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 0.5;
    const double vv = v0 ? v0[i / comp] : V_INIT + 2.0 * init_v_rand;
    H.v[i] = vv;
    auto gate = [&](int q) -> double & { return H.gate[(size_t)q * nc + i]; };
    gate(n_KV) = inf_n_KV(vv);
    gate(a_KA) = inf_a_KA(vv);
    gate(b_KA) = inf_b_KA(vv);
    gate(ir_KIR) = inf_ir_KIR(vv);
    gate(s_KM) = inf_s_KM(vv);
    gate(c_KCa) = inf_c_KCa(vv, H.ca[i]);
    gate(ch_Ca) = inf_ch_Ca(vv);
    gate(ci_Ca) = inf_ci_Ca(vv);
    gate(c1_Na) = 1.0;   // the other Na states start at 0
  }
  return true;
}

}  // namespace cbrkc
