// Chebyshev stage coefficients of the RKC step (solve_rkc.cu:815-906),
// tabulated per stage count m.  They depend on m alone, so the per-cell
// kernel reads them (uniform loads, one row per step) instead of running the
// recurrence and log/sinh/cosh in every wave of every step.  Built with
// -ffp-contract=off, so every entry is the reference's expression evaluated
// in the reference's order.
#include <cmath>

#include "engine.hpp"

namespace cbrkc {

std::vector<double> rkc_stage_table(int mtab) {
  std::vector<double> t((size_t)(mtab + 1) * mtab * 4, 0.0);
  for (int m = 2; m <= mtab; m++) {
    double *r = &t[(size_t)m * mtab * 4];
    const double w0 = 1.0 + 2.0 / (13.0 * m * m);
    const double temp1 = w0 * w0 - 1.0;
    const double temp2 = std::sqrt(temp1);
    const double arg = m * std::log(w0 + temp2);
    const double w1 = std::sinh(arg) * temp1 / (std::cosh(arg) * m * temp2 - w0 * std::sinh(arg));
    double bjm1 = 1.0 / ((2.0 * w0) * (2.0 * w0)), bjm2 = bjm1;
    r[0] = w1 * bjm1;   // mu~_1, the first stage's multiplier of h F_0
    double zjm1 = w0, zjm2 = 1.0, dzjm1 = 1.0, dzjm2 = 0.0, d2zjm1 = 0.0, d2zjm2 = 0.0;
    for (int j = 1; j < m; j++) {
      const double zj = 2.0 * w0 * zjm1 - zjm2;
      const double dzj = 2.0 * w0 * dzjm1 - dzjm2 + 2.0 * zjm1;
      const double d2zj = 2.0 * w0 * d2zjm1 - d2zjm2 + 4.0 * dzjm1;
      const double bj = d2zj / (dzj * dzj);
      const double ajm1 = 1.0 - zjm1 * bjm1;
      const double mu = 2.0 * w0 * bj / bjm1;
      const double nu = -bj / bjm2;
      const double mus = mu * w1 / w0;
      r[4 * j + 0] = mu;
      r[4 * j + 1] = nu;
      r[4 * j + 2] = mus;
      r[4 * j + 3] = ajm1;
      bjm2 = bjm1; bjm1 = bj;
      zjm2 = zjm1; zjm1 = zj;
      dzjm2 = dzjm1; dzjm1 = dzj;
      d2zjm2 = d2zjm1; d2zjm1 = d2zj;
    }
  }
  return t;
}

}  // namespace cbrkc
