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

bool build_gr_host(const std::string &param_path, int nx, int ny, const double *v0, GrHost &H, std::string &err) {
  using namespace gr;
  std::FILE *f = std::fopen(param_path.c_str(), "r");
  if (!f) {
    err = "no such file " + param_path;
    return false;
  }
  // gr.cu:29-58 / gr_solve.cu:20-35: idx parent diam len Ra Cm type + 10 densities
  std::vector<double> rad(COMP, 0.0), len(COMP, 0.0), Ra(COMP, 0.0), cmd(COMP, 0.0), dens((size_t)COMP * NCOND, 0.0);
  std::vector<int> compart(COMP, 0), l_connect(COMP, -1);
  std::vector<char> seen(COMP, 0);
  for (int i = 0; i < COMP; i++) {
    int i1, i2, i3;
    double d1, d2, d3, d4, x[NCOND];
    if (std::fscanf(f, "%d %d %lf %lf %lf %lf %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf ", &i1, &i2, &d1, &d2, &d3,
                    &d4, &i3, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6], &x[7], &x[8], &x[9]) != 17 ||
        i1 < 0 || i1 >= COMP || i2 < -1 || i2 >= COMP || i3 < 0 || i3 >= NTYPE) {
      std::fclose(f);
      err = "GR_PARAM_FILE_READING_ERROR in " + param_path;
      return false;
    }
    rad[i1] = 0.5 * d1 * 1e-4;
    len[i1] = d2 * 1e-4;
    Ra[i1] = d3 * 1e-3;
    cmd[i1] = d4;
    compart[i1] = i3;
    seen[i1] = 1;
    for (int q = 0; q < NCOND; q++) dens[(size_t)i1 * NCOND + q] = x[q];
    l_connect[i] = i2;   // connected by line, gr_solve.cu:34
  }
  std::fclose(f);
  for (int i = 0; i < COMP; i++)
    if (!seen[i]) {
      err = "GR_PARAM_FILE_READING_ERROR: compartment missing in " + param_path;
      return false;
    }

  const int n = nx * ny, nc = n * COMP;
  H.n = n;
  H.nx = nx;
  H.ny = ny;
  H.comp = COMP;
  H.nc = nc;
  H.cm.assign(nc, 0.0);
  H.iext.assign(nc, 0.0);
  H.v.assign(nc, 0.0);
  H.ca.assign(nc, CA1_0);
  H.compart.assign(nc, 0);
  H.cond.assign((size_t)NCOND * nc, 0.0);
  H.gate.assign((size_t)NGATE * nc, 0.0);
  for (int c = 0; c < n; c++)
    for (int k = 0; k < COMP; k++) {
      const size_t i = (size_t)c * COMP + k;
      const double a = 2.0 * M_PI * rad[k] * len[k];   // gr.cu:44-57
      H.cm[i] = cmd[k] * a;
      H.compart[i] = (int8_t)compart[k];
      for (int q = 0; q < NCOND; q++) H.cond[(size_t)q * nc + i] = dens[(size_t)k * NCOND + q] * a;
    }
  // cable conductances (gr_solve.cu:37-43), CSR per row in column order with
  // the stored diagonal, RKC sign (gr_solve.cu:73-133)
  std::vector<double> mat((size_t)COMP * COMP, 0.0);
  for (int i = 0; i < COMP; i++) {
    const int d = l_connect[i];
    if (d >= 0) {
      mat[d + (size_t)COMP * i] =
          (2.0 / ((Ra[i] * len[i]) / (rad[i] * rad[i] * M_PI) + (Ra[d] * len[d]) / (rad[d] * rad[d] * M_PI)));
      mat[i + (size_t)COMP * d] = mat[d + (size_t)COMP * i];
    }
  }
  std::vector<int> rcol;
  std::vector<double> rval;
  std::vector<int> rptr(COMP + 1, 0);
  for (int r = 0; r < COMP; r++) {
    double dsum = 0.0;
    for (int c = 0; c < COMP; c++)
      if (c != r) dsum += mat[c + (size_t)COMP * r];
    for (int c = 0; c < COMP; c++) {
      const double m = mat[c + (size_t)COMP * r];
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
  const size_t nnz1 = rcol.size();
  H.row.assign(nc + 1, 0);
  H.col.resize(nnz1 * n);
  H.val.resize(nnz1 * n);
  for (int cell = 0; cell < n; cell++) {
    for (size_t j = 0; j < nnz1; j++) {
      H.col[cell * nnz1 + j] = rcol[j] + cell * COMP;
      H.val[cell * nnz1 + j] = rval[j];
    }
    for (int r = 0; r < COMP; r++) H.row[cell * COMP + r + 1] = (int)(cell * nnz1 + rptr[r + 1]);
  }

  // initial state, gr_ion.cu:635-672
  double init_v_rand = 0.0;
  for (int i = 0; i < nc; i++) {
    if (i % COMP == 0 && !v0) init_v_rand = (((double)std::rand() + 1.0) / ((double)RAND_MAX + 2.0)) - 0.5;
    const double vv = v0 ? v0[i / COMP] : V_INIT + 2.0 * init_v_rand;
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
