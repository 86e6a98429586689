// Crank-Nicolson (CN) engine: the reference's second solver (solve_cnm.cu)
// for every cell type, FP64, one population-wide linear system per step.
//
// One CN step of CN_DT = 0.025 ms (xx_solve_by_cnm) is
//   1. one fused per-compartment kernel: the type's CN ion update
//      (pkj_update_ion_2nd, io_update_ion_2nd, go_update_ion_exp_imp + the
//      KAHP step, gr_Na_update_2order + gr_update_ion_exp_imp), the half
//      conductance / reversal sums gamma2 / omega2, the system diagonal
//      Cm/DT + gamma2 on top of the halved cable matrix, the right-hand side
//      (xx_update_matrix) and, for the inferior olive, the explicit gap term
//      (io_gap_update + add_io_gap_val);
//   2. the reference's global Krylov solve from x = 0 (cg_cusparse_crs for
//      tree cells, bicg_cusparse_crs for the unsymmetric inferior-olive
//      coupling): at most 100 iterations, stopping when |r| <= 1e-15 |b|.
//      For 1600-compartment Purkinje trees the 100 iterations never reach
//      that tolerance, so the result is the truncated iterate -- the solve
//      must be the same global iteration, not a per-cell direct solve.
//
// The backward-Euler (BE) solver of solve_bem.cu shares the Krylov core: one
// fused kernel runs the type's BE ion update (pkj_update_ion, io_update_ion,
// go_KAHP_update + go_update_ion, gr_Na_update + gr_update_ion: exponential
// Euler gates and implicit Markov steps at the step's starting voltage) and
// builds the BE system on the reference's unhalved matrix, diagonal
// val_ori + Cm/DT + conductances, right-hand side Cm/DT V + reversal-weighted
// conductances + I_ext (minus the inferior olive's gap currents); the solve is
// the same CG / BiCG.
//
// The Krylov loop runs on the device without host round trips: each
// iteration is two launches, every dot product is reduced deterministically
// (256-wide stride-halving trees, the partials reduced again with the same
// tree by whichever block finishes last), and a device flag set on
// convergence turns the remaining launches into no-ops.  p_k = beta p_{k-1} + r
// is formed on the fly inside the next matrix-vector product (p double
// buffered), so an iteration needs no separate vector update launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "detsum.hpp"
#include "engine.hpp"
#include "record.hpp"
#include "go_model.hpp"
#include "gr_model.hpp"
#include "io_model.hpp"
#include "morph.hpp"
#include "pkj_model.hpp"

namespace cbrkc {
namespace {

constexpr double CN_DT = 0.025;   // param.h:11
constexpr double BE_DT = 0.025;   // param.h:10
constexpr double H_MAX = 0.125;   // param.h:12, the output interval
constexpr int NB = 256;

enum Scal { S_BNORM = 0, S_RR, S_ALPHA, S_BETA, S_NSCAL };
enum Flag { F_DONE = 0, F_TICKET, F_ITERS, F_TOTAL, F_NFLAG };

// The system and the Krylov vectors (device pointers)
struct CnSys {
  const int *row, *col, *dpos;       // CSR of the halved cable matrix (reference sign)
  const double *val;
  const int *trow, *tcol, *tsrc;     // transpose: row start, column, source entry
  const double *cm;
  double *v, *vold, *gamma1, *omega1, *iext;
  double *diag, *b, *r, *ap, *rs, *atps;
  double *p[2], *ps[2];
  double *part, *scal;
  int *flag;
  int nc, bicg;
  int direct;                        // 1: the per-cell tree solve replaces the Krylov loop
  double dt;                         // the step [ms]: CN_DT / BE_DT, or a whole fraction of H_MAX
  SynIn syn;                         // incoming synapses (none: syn.on() false)
};

// Deterministic sum of the block's values x (one per thread) into
// part[blockIdx.x]; the block that finishes last reduces all partials level by
// level with the same tree and returns true with the total in *out (thread 0).
// The partials cross workgroups (and the XCDs' separate L2 caches), so they
// are stored and loaded with device-scope atomics; every block releases its
// partial with the ticket increment, and only the block that turns out last
// acquires: its thread 0, right after reading the ticket, so the acquire
// follows the atomic read it pairs with; the workgroup barrier then orders
// the other threads' loads after it (the other blocks read nothing).
__device__ bool block_sum_last(const CnSys &S, double x, double *lds, double *out) {
  const int nb = gridDim.x;
  const double bs = block_tree_sum(x, lds);
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&S.part[blockIdx.x], bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int prev = __hip_atomic_fetch_add(&S.flag[F_TICKET], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    last = (prev == nb - 1);
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  if (!last) return false;
  int m = nb;
  while (m > 1) {
    const int chunks = (m + NB - 1) / NB;
    for (int c = 0; c < chunks; c++) {
      const int k = c * NB + threadIdx.x;
      const double y = (k < m) ? __hip_atomic_load(&S.part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
      __syncthreads();
      const double t = block_tree_sum(y, lds);
      __syncthreads();
      if (threadIdx.x == 0) __hip_atomic_store(&S.part[c], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
    }
    m = chunks;
  }
  if (threadIdx.x == 0) {
    *out = __hip_atomic_load(&S.part[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    S.flag[F_TICKET] = 0;
  }
  return true;
}

// b = r = p = bb (and the shadow vectors of BiCG); returns bb
__device__ inline double set_rhs(const CnSys &S, int i, double bb) {
  S.b[i] = bb;
  S.r[i] = bb;
  S.p[0][i] = bb;
  if (S.bicg) {
    S.rs[i] = bb;
    S.ps[0][i] = bb;
  }
  return bb;
}

// the generic part of xx_update_matrix for compartment i given the step's
// conductance and reversal sums (gs, es) and the initial omega2 (the IO gap
// term, else 0); writes the diagonal and the right-hand side; returns b_i
__device__ inline double cn_row(const CnSys &S, int i, double gs, double es, double om0) {
  const double DT = S.dt;
  // add_*_val (solve_cnm.cu:100-128, 245-257, 406-418): gamma2 += val / 2,
  // omega2 += val E / 2 per synapse, before the ion sums
  double g0 = 0.0, o0 = om0;
  if (S.syn.on())
    for (int j = S.syn.row[i]; j < S.syn.row[i + 1]; j++) {
      double E;
      const double val = syn_entry(S.syn, j, E);
      g0 += 0.5 * val;
      o0 += 0.5 * val * E;
    }
  const double gamma2 = g0 + gs / 2.0;
  const double omega2 = o0 + es / 2.0;
  S.diag[i] = S.val[S.dpos[i]] + ((S.cm[i] / DT) + gamma2);
  double bb = 0.0;
  for (int j = S.row[i]; j < S.row[i + 1]; j++) bb -= S.v[S.col[j]] * S.val[j];
  const double vi = S.v[i];
  bb += (S.cm[i] / DT - S.gamma1[i]) * vi + S.omega1[i] + omega2;
  S.gamma1[i] = gamma2;
  S.omega1[i] = omega2;
  return set_rhs(S, i, bb);
}

// the BE row: diagonal val_ori[d] (+ each synapse's val, add_*_val,
// solve_bem.cu:229-255, 306-320) + gs (gs carries Cm/DT first); b = the
// synapses' val E from 0, + es
__device__ inline double be_row(const CnSys &S, int i, double gs, double es) {
  double d0 = S.val[S.dpos[i]], b0 = 0.0;
  if (S.syn.on())
    for (int j = S.syn.row[i]; j < S.syn.row[i + 1]; j++) {
      double E;
      const double val = syn_entry(S.syn, j, E);
      d0 += val;
      b0 += val * E;
    }
  S.diag[i] = d0 + gs;
  return set_rhs(S, i, b0 + es);
}

// the BE row without synapses and with b given (IO: the gap currents follow es)
__device__ inline double be_row_b(const CnSys &S, int i, double gs, double bb) {
  S.diag[i] = S.val[S.dpos[i]] + gs;
  return set_rhs(S, i, bb);
}

// end of the prep kernel: |b|^2 (= the first r.r) and the solve's flags
__device__ inline void cn_prep_finish(const CnSys &S, double bb, double *lds) {
  if (S.direct) return;              // uniform over the grid: no Krylov state to set up
  double tot;
  if (block_sum_last(S, bb * bb, lds, &tot) && threadIdx.x == 0) {
    S.scal[S_BNORM] = tot;
    S.scal[S_RR] = tot;
    S.flag[F_DONE] = 0;
    S.flag[F_ITERS] = 0;
  }
}

// ---- the cell types' CN ion updates (fused with the matrix row) ----

struct PkjCn {
  double *ca, *rev;
  const double *shell, *area, *cond;
  double *gate;
  const int8_t *compart;
};

__global__ void __launch_bounds__(NB) k_cn_prep_pkj(CnSys S, PkjCn M) {
  using namespace pkj;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // pkj_update_ion_2nd (pkj_ion.cu:246-302): Ca current at V, gates at the
    // voltage extrapolated from the last two steps
    const double vm = (1.5 * v - 0.5 * S.vold[i]);
    double ca = M.ca[i], rev = M.rev[i];
    const double caq = ca_rk4(v, ca, rev, M.shell[i], c[g_CaP] * g[m_CaP] * g[h_CaP], c[g_CaT] * g[m_CaT] * g[h_CaT],
                              S.dt, i % 1599 == 0);
    gates_cn<double>(vm, caq, S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    M.rev[i] = rev;
    S.vold[i] = v;
    double gs, es;
    cn_ge(c, g, rev, S.iext[i], gs, es);
    bb = cn_row(S, i, gs, es, 0.0);
  }
  cn_prep_finish(S, bb, lds);
}

struct IoCn {
  double *ca;
  const double *cond;
  double *gate;
  const int8_t *compart;
  const int *nbr;
  int n;
};

__global__ void __launch_bounds__(NB) k_cn_prep_io(CnSys S, IoCn M) {
  using namespace io;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // io_update_ion_2nd (io_ion.cu:53-89)
    double ca = M.ca[i];
    update_ion_at<double>(v, 1.5 * v - 0.5 * S.vold[i], ca, c[g_CaH], M.compart[i], S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    S.vold[i] = v;
    // io_gap_update + add_io_gap_val (gap.cu:142-160, solve_cnm.cu:729-739):
    // minus the gap current into the dendrite, partners in slot order
    double om0 = 0.0;
    if (M.n > 1 && i % COMP == 0) {
      const int cell = i / COMP;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int pre = M.nbr[cell * 4 + s];
        const double diff_v = (pre > -1) ? (v - S.v[pre * COMP]) : 0.0;
        om0 += -1.0 * gap_current<double>(diff_v);
      }
    }
    double gs, es;
    cn_ge(c, g, S.iext[i], gs, es);
    bb = cn_row(S, i, gs, es, om0);
  }
  cn_prep_finish(S, bb, lds);
}

struct GoCn {
  double *ca, *ca2, *rev, *ca_old;
  const double *area, *cond;
  double *gate;
};

__global__ void __launch_bounds__(NB) k_cn_prep_go(CnSys S, GoCn M) {
  using namespace go;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i], vo = S.vold[i];
    // go_update_ion_exp_imp (go_ion.cu:319-377): Ca currents at V_old, gates
    // at (V + V_old) / 2; then the KAHP step (go_ion.cu:88-213)
    double ca = M.ca[i], ca2 = M.ca2[i], rev = M.rev[i];
    update_ion_at(vo, v, vo, M.area[i], c, ca, ca2, rev, S.dt, g);
    kahp_step(M.ca_old[i], ca, S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    M.ca2[i] = ca2;
    M.rev[i] = rev;
    M.ca_old[i] = ca;
    S.vold[i] = v;
    double gs, es;
    cn_ge(c, g, rev, S.iext[i], gs, es);
    bb = cn_row(S, i, gs, es, 0.0);
  }
  cn_prep_finish(S, bb, lds);
}

struct GrCn {
  const double *ca, *cond;
  double *gate;
  const int8_t *compart;
  const gr::NaType *na;
};

__global__ void __launch_bounds__(NB) k_cn_prep_gr(CnSys S, GrCn M) {
  using namespace gr;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i], vo = S.vold[i];
    // gr_Na_update_2order from V_old to V, then gr_update_ion_exp_imp at
    // (V + V_old) / 2 (solve_cnm.cu:221-226, gr_ion.cu:438-459)
    const int t = M.compart[i];
    if (t >= HILL) {
      const NaType T = M.na[t];
      double x[12];
      na_gather(g, x);
      if (na_step(T, vo, v, S.dt, x)) na_scatter(x, g);
    }
    update_gates(v, vo, M.ca[i], S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    S.vold[i] = v;
    double gs, es;
    cn_ge(c, g, S.iext[i], gs, es);
    bb = cn_row(S, i, gs, es, 0.0);
  }
  cn_prep_finish(S, bb, lds);
}

// ---- the cell types' BE ion updates (fused with the matrix row) ----

__global__ void __launch_bounds__(NB) k_be_prep_pkj(CnSys S, PkjCn M) {
  using namespace pkj;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // pkj_update_ion (pkj_ion.cu:94-151) then pkj_update_matrix
    // (solve_bem.cu:407-445)
    double ca = M.ca[i], rev = M.rev[i];
    update_ion_be(v, ca, rev, M.shell[i], M.area[i], M.compart[i] == SOMA, c[g_CaP] * g[m_CaP] * g[h_CaP],
                  c[g_CaT] * g[m_CaT] * g[h_CaT], S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    M.rev[i] = rev;
    double gs, es;
    be_ge(c, g, rev, S.iext[i], S.cm[i] / S.dt, v, gs, es);
    bb = be_row_b(S, i, gs, 0.0 + es);   // pkj_solve_by_bem couples no synapses
  }
  cn_prep_finish(S, bb, lds);
}

__global__ void __launch_bounds__(NB) k_be_prep_io(CnSys S, IoCn M) {
  using namespace io;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // io_update_ion (io_ion.cu:128-161) then io_update_matrix
    // (solve_bem.cu:476-508)
    double ca = M.ca[i];
    update_ion_be<double>(v, ca, c[g_CaH], M.compart[i], S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    double gs, es;
    be_ge(c, g, S.iext[i], S.cm[i] / S.dt, v, gs, es);
    bb = 0.0 + es;
    // io_gap_update + add_io_gap_val (gap.cu:142-160, solve_bem.cu:463-473):
    // minus each gap current into the dendrite, partners in slot order
    if (M.n > 1 && i % COMP == 0) {
      const int cell = i / COMP;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int pre = M.nbr[cell * 4 + s];
        const double diff_v = (pre > -1) ? (v - S.v[pre * COMP]) : 0.0;
        bb += -1.0 * gap_current<double>(diff_v);
      }
    }
    bb = be_row_b(S, i, gs, bb);
  }
  cn_prep_finish(S, bb, lds);
}

__global__ void __launch_bounds__(NB) k_be_prep_go(CnSys S, GoCn M) {
  using namespace go;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // go_KAHP_update at the present Ca (go_ion.cu:214-317), go_update_ion
    // (go_ion.cu:438-511), then go_update_matrix (solve_bem.cu:323-347)
    double ca = M.ca[i], ca2 = M.ca2[i], rev = M.rev[i];
    kahp_be(ca, S.dt, g);
    update_ion_be(v, M.area[i], c, ca, ca2, rev, S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    M.ca[i] = ca;
    M.ca2[i] = ca2;
    M.rev[i] = rev;
    double gs, es;
    be_ge(c, g, rev, S.iext[i], S.cm[i] / S.dt, v, gs, es);
    bb = be_row(S, i, gs, es);
  }
  cn_prep_finish(S, bb, lds);
}

__global__ void __launch_bounds__(NB) k_be_prep_gr(CnSys S, GrCn M) {
  using namespace gr;
  __shared__ double lds[NB];
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  double bb = 0.0;
  if (i < nc) {
    double g[NGATE], c[NCOND];
#pragma unroll
    for (int k = 0; k < NGATE; k++) g[k] = M.gate[(size_t)k * nc + i];
#pragma unroll
    for (int k = 0; k < NCOND; k++) c[k] = M.cond[(size_t)k * nc + i];
    const double v = S.v[i];
    // gr_Na_update and gr_update_ion at V (gr_ion.cu:209-357, 579-600), then
    // gr_update_matrix (solve_bem.cu:258-291)
    const int t = M.compart[i];
    if (t >= HILL) {
      const NaType T = M.na[t];
      double x[12];
      na_gather(g, x);
      if (na_be(T, v, S.dt, x)) na_scatter(x, g);
    }
    update_gates_be(v, M.ca[i], S.dt, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) M.gate[(size_t)k * nc + i] = g[k];
    double gs, es;
    be_ge(c, g, S.iext[i], S.cm[i] / S.dt, v, gs, es);
    bb = be_row(S, i, gs, es);
  }
  cn_prep_finish(S, bb, lds);
}

// ---- Krylov iteration k: A (or cg) half ----
// p_k = beta p_{k-1} + r (k > 0) on the fly, Ap = A p_k (BiCG: also A^T ps_k),
// alpha = rr / (p_k . Ap) (BiCG: ps_k . Ap)
__global__ void __launch_bounds__(NB) k_kry_a(CnSys S, int k) {
  __shared__ double lds[NB];
  if (S.flag[F_DONE]) return;
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  const double beta = S.scal[S_BETA];
  const double *pp = S.p[(k + 1) & 1], *pps = S.ps[(k + 1) & 1];   // p_{k-1}
  double *pk = S.p[k & 1], *psk = S.ps[k & 1];
  double dotv = 0.0;
  if (i < nc) {
    if (k == 0) S.v[i] = 0.0 * S.v[i];   // x = 0 (cublasDscal by 0, solve_cnm.cu:65)
    double s = 0.0, mine = 0.0;
    for (int j = S.row[i]; j < S.row[i + 1]; j++) {
      const int cj = S.col[j];
      const double pj = (k == 0) ? pk[cj] : (beta * pp[cj] + S.r[cj]);
      if (cj == i) mine = pj;
      s += ((j == S.dpos[i]) ? S.diag[i] : S.val[j]) * pj;
    }
    if (k > 0) pk[i] = mine;
    S.ap[i] = s;
    if (S.bicg) {
      double st = 0.0, mines = 0.0;
      for (int t = S.trow[i]; t < S.trow[i + 1]; t++) {
        const int cj = S.tcol[t], src = S.tsrc[t];
        const double pj = (k == 0) ? psk[cj] : (beta * pps[cj] + S.rs[cj]);
        if (cj == i) mines = pj;
        st += ((src == S.dpos[cj]) ? S.diag[cj] : S.val[src]) * pj;
      }
      if (k > 0) psk[i] = mines;
      S.atps[i] = st;
      dotv = mines * s;
    } else {
      dotv = mine * s;
    }
  }
  double tot;
  if (block_sum_last(S, dotv, lds, &tot) && threadIdx.x == 0) S.scal[S_ALPHA] = S.scal[S_RR] / tot;
}

// ---- Krylov iteration k: update half ----
// x += alpha p_k, r -= alpha Ap (BiCG: rs -= alpha A^T ps_k), rr' = r.r
// (BiCG: r.rs); stop when sqrt(rr') <= 1e-15 sqrt(|b|^2), else beta = rr'/rr
__global__ void __launch_bounds__(NB) k_kry_b(CnSys S, int k) {
  __shared__ double lds[NB];
  if (S.flag[F_DONE]) return;
  const int nc = S.nc, i = blockIdx.x * NB + threadIdx.x;
  const double alpha = S.scal[S_ALPHA];
  const double nalpha = -1.0 * alpha;
  double dotv = 0.0;
  if (i < nc) {
    S.v[i] = alpha * S.p[k & 1][i] + S.v[i];
    const double ri = nalpha * S.ap[i] + S.r[i];
    S.r[i] = ri;
    if (S.bicg) {
      const double rsi = nalpha * S.atps[i] + S.rs[i];
      S.rs[i] = rsi;
      dotv = ri * rsi;
    } else {
      dotv = ri * ri;
    }
  }
  double tot;
  if (block_sum_last(S, dotv, lds, &tot) && threadIdx.x == 0) {
    S.flag[F_ITERS] = k + 1;
    if (sqrt(tot) <= 1.0e-15 * sqrt(S.scal[S_BNORM])) {
      S.flag[F_DONE] = 1;
      S.flag[F_TOTAL] += k + 1;
    } else {
      S.scal[S_BETA] = tot / S.scal[S_RR];
      S.scal[S_RR] = tot;
      if (k == 99) S.flag[F_TOTAL] += 100;
    }
  }
}

// ---- the whole Krylov solve of a small system in one workgroup ----
// For populations of at most KRY_SMALL_BLOCKS x 256 compartments the 200
// launches of a step cost more than the arithmetic, and the reference's
// solves often converge early (the olive's BiCG in about eight iterations),
// after which the multi-block launches only return.  This kernel runs the
// same iteration in one workgroup: rows in 256-wide chunks (chunk c = block c
// of k_kry_a / k_kry_b), every dot product as the chunk partials of
// block_tree_sum reduced once more with the same tree (the order of
// block_sum_last), so the iterates are bit-identical to the multi-block path;
// it stops at convergence.
constexpr int KRY_SMALL_BLOCKS = 16;

// The total of the values the threads left in v[c * NB + t] for chunks
// c < nch: every chunk's stride-halving tree (the additions of
// block_tree_sum), all chunks' levels at once, then the chunk partials by the
// same tree as block_sum_last.  Returned to every thread.
__device__ inline double chunks_total(double *v, int nch, double *lds, double *bc) {
  __syncthreads();
  for (int stride = NB / 2; stride >= 1; stride >>= 1) {
    for (int q = threadIdx.x; q < nch * stride; q += NB) {
      const int c = q / stride, t = q - c * stride;
      v[c * NB + t] = v[c * NB + t] + v[c * NB + t + stride];
    }
    __syncthreads();
  }
  if (nch > 1) {
    const double y = (threadIdx.x < nch) ? v[threadIdx.x * NB] : 0.0;
    const double t = block_tree_sum(y, lds);
    if (threadIdx.x == 0) *bc = t;
  } else if (threadIdx.x == 0) {
    *bc = v[0];
  }
  __syncthreads();
  return *bc;
}

__global__ void __launch_bounds__(NB) k_kry_small(CnSys S) {
  __shared__ double lds[NB];
  __shared__ double part[KRY_SMALL_BLOCKS * NB];
  __shared__ double bc;
  const int nc = S.nc, nch = (nc + NB - 1) / NB;
  const double bnorm = S.scal[S_BNORM];
  double rr = S.scal[S_RR], alpha = 0.0, beta = 0.0;
  int k = 0, done = 0;
  for (; k < 100 && !done; k++) {
    const double *pp = S.p[(k + 1) & 1], *pps = S.ps[(k + 1) & 1];   // p_{k-1}
    double *pk = S.p[k & 1], *psk = S.ps[k & 1];
    double tot = 0.0;
    for (int c = 0; c < nch; c++) {                  // the k_kry_a half
      const int i = c * NB + threadIdx.x;
      double dotv = 0.0;
      if (i < nc) {
        if (k == 0) S.v[i] = 0.0 * S.v[i];
        double sa = 0.0, mine = 0.0;
        for (int j = S.row[i]; j < S.row[i + 1]; j++) {
          const int cj = S.col[j];
          const double pj = (k == 0) ? pk[cj] : (beta * pp[cj] + S.r[cj]);
          if (cj == i) mine = pj;
          sa += ((j == S.dpos[i]) ? S.diag[i] : S.val[j]) * pj;
        }
        if (k > 0) pk[i] = mine;
        S.ap[i] = sa;
        if (S.bicg) {
          double st = 0.0, mines = 0.0;
          for (int t = S.trow[i]; t < S.trow[i + 1]; t++) {
            const int cj = S.tcol[t], src = S.tsrc[t];
            const double pj = (k == 0) ? psk[cj] : (beta * pps[cj] + S.rs[cj]);
            if (cj == i) mines = pj;
            st += ((src == S.dpos[cj]) ? S.diag[cj] : S.val[src]) * pj;
          }
          if (k > 0) psk[i] = mines;
          S.atps[i] = st;
          dotv = mines * sa;
        } else {
          dotv = mine * sa;
        }
      }
      part[c * NB + threadIdx.x] = dotv;
    }
    tot = chunks_total(part, nch, lds, &bc);
    alpha = rr / tot;
    const double nalpha = -1.0 * alpha;
    for (int c = 0; c < nch; c++) {                  // the k_kry_b half
      const int i = c * NB + threadIdx.x;
      double dotv = 0.0;
      if (i < nc) {
        S.v[i] = alpha * S.p[k & 1][i] + S.v[i];
        const double ri = nalpha * S.ap[i] + S.r[i];
        S.r[i] = ri;
        if (S.bicg) {
          const double rsi = nalpha * S.atps[i] + S.rs[i];
          S.rs[i] = rsi;
          dotv = ri * rsi;
        } else {
          dotv = ri * ri;
        }
      }
      part[c * NB + threadIdx.x] = dotv;
    }
    tot = chunks_total(part, nch, lds, &bc);
    if (sqrt(tot) <= 1.0e-15 * sqrt(bnorm)) {
      done = 1;
    } else {
      beta = tot / rr;
      rr = tot;
    }
  }
  if (threadIdx.x == 0) {
    S.flag[F_ITERS] = k;
    S.flag[F_DONE] = 1;
    S.flag[F_TOTAL] += k;
    S.scal[S_ALPHA] = alpha;
    S.scal[S_BETA] = beta;
    S.scal[S_RR] = rr;
  }
}

__global__ void k_cn_stim(double *__restrict__ iext, const double *__restrict__ amp, int comp, int stim_comp,
                          double t0, double t1, double t, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    double *x = &iext[stim_comp + (size_t)c * comp];
    *x = 0;
    if (t0 <= t && t < t1) *x = amp[c];
  }
}

__global__ void k_cn_record(const double *__restrict__ v, double *__restrict__ out, int comp, int rec, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = v[rec + (size_t)c * comp];
}

// ---- the exact per-cell solve (Hines elimination on each cell's tree) ----
//
// Each cell's compartments are renumbered breadth-first from the centre of
// its tree (the middle of a longest path, which halves the number of levels
// against a leaf root), so every level is a contiguous range of positions and
// a node's children are contiguous too.  One 64-lane wavefront per cell
// stages the cell's diagonal, right-hand side and off-diagonal pairs in LDS,
// eliminates level by level from the deepest parents to the root (each parent
// folds in its children in position order: d_p -= A(p,c) A(c,p) / d_c,
// r_p -= A(p,c) r_c / d_c), then substitutes back from the root
// (x_c = (r_c - A(c,p) x_p) / d_c).  No fill-in, no pivoting (the CN and BE
// systems are diagonally dominant), a fixed order, so the result does not
// depend on the launch.
struct HinesPlan {
  const int *perm;       // position -> global compartment, [cell * comp + pos]
  const double *up, *dn; // A(node, parent), A(parent, node) per position (root: 0)
  const int *meta;       // first child position << 8 | child count
  const int *lev;        // per cell: level start positions, then comp
  const int *lev_off;    // cell -> its first entry in lev (n + 1 entries)
  int comp;
};

constexpr int HINES_WAVE = 64;
constexpr int CHAIN_SEGS = 64;   // segments per cell of the chain solve (one per lane)
constexpr int HINES_MAX_COMP = 65536 / (4 * 8 + 4);   // the LDS a cell's staging takes (64 KB)

// SYM: A(node, parent) = A(parent, node) everywhere (the CN and BE tree
// matrices of the Purkinje, granule and Golgi cells), so one off-diagonal
// plane is staged: 28 instead of 36 bytes of LDS per compartment, three
// Purkinje cells per CU instead of two
template <bool SYM>
__global__ void __launch_bounds__(HINES_WAVE) k_hines(HinesPlan P, const double *__restrict__ diag,
                                                      const double *__restrict__ b, double *__restrict__ v) {
  extern __shared__ double sm[];
  const int comp = P.comp, lane = threadIdx.x;
  double *d = sm, *r = sm + comp, *up = sm + 2 * comp;
  double *dn = SYM ? up : sm + 3 * comp;
  int *meta = (int *)(sm + (SYM ? 3 : 4) * comp);
  const size_t base = (size_t)blockIdx.x * comp;
  for (int k = lane; k < comp; k += HINES_WAVE) {
    const int g = P.perm[base + k];
    d[k] = diag[g];
    r[k] = b[g];
    up[k] = P.up[base + k];
    if (!SYM) dn[k] = P.dn[base + k];
    meta[k] = P.meta[base + k];
  }
  __syncthreads();
  const int *lev = P.lev + P.lev_off[blockIdx.x];
  const int nl = P.lev_off[blockIdx.x + 1] - P.lev_off[blockIdx.x] - 1;   // levels
  for (int l = nl - 2; l >= 0; l--) {
    for (int k = lev[l] + lane; k < lev[l + 1]; k += HINES_WAVE) {
      const int m = meta[k], fc = m >> 8, nch = m & 255;
      double dd = d[k], rr = r[k];
      for (int j = fc; j < fc + nch; j++) {
        const double f = dn[j] / d[j];
        dd -= f * up[j];
        rr -= f * r[j];
      }
      d[k] = dd;
      r[k] = rr;
    }
    __syncthreads();
  }
  if (lane == 0) r[0] = r[0] / d[0];
  __syncthreads();
  for (int l = 0; l < nl - 1; l++) {
    for (int k = lev[l] + lane; k < lev[l + 1]; k += HINES_WAVE) {
      const int m = meta[k], fc = m >> 8, nch = m & 255;
      const double xp = r[k];
      for (int j = fc; j < fc + nch; j++) r[j] = (r[j] - up[j] * xp) / d[j];
    }
    __syncthreads();
  }
  for (int k = lane; k < comp; k += HINES_WAVE) v[P.perm[base + k]] = r[k];
}

// The same elimination by chains, for trees made of few long unbranched
// runs (the granule cell's 578 compartments are 9 chains, the Golgi cell's 131
// are 6; the level walk pays one barrier per compartment of depth there).  A
// segment is a maximal run bottom -> top in which every compartment above the
// bottom has exactly one child; positions are numbered segment by segment
// (bottom first), the segments by height (0: the bottom is a leaf, else one
// more than its child segments).  One lane per segment: at height h it folds
// the tops of its child segments into its bottom (child order as in k_hines),
// then walks up its chain in registers, one division per compartment, no
// barrier; back substitution walks down from the top.  Every compartment sees
// the same operations in the same order as in k_hines, so the two agree bit
// for bit.
struct ChainPlan {
  const int *perm;       // position -> global compartment
  const double *up, *dn; // per position
  const int *seg;        // [cell][CHAIN_SEGS][4]: start, len, kf | kc << 16, (ptop + 1) | h << 16
  const int *kids;       // [cell][CHAIN_SEGS]: positions of the child segments' tops
  const int *nseg, *nh;  // per cell: segments, heights
  int comp;
};

__global__ void __launch_bounds__(HINES_WAVE) k_hines_chain(ChainPlan P, const double *__restrict__ diag,
                                                            const double *__restrict__ b, double *__restrict__ v) {
  extern __shared__ double sm[];
  const int comp = P.comp, lane = threadIdx.x, cell = blockIdx.x;
  double *d = sm, *r = sm + comp, *up = sm + 2 * comp, *dn = sm + 3 * comp;
  int *kids = (int *)(sm + 4 * comp);
  const size_t base = (size_t)cell * comp;
  for (int k = lane; k < comp; k += HINES_WAVE) {
    const int g = P.perm[base + k];
    d[k] = diag[g];
    r[k] = b[g];
    up[k] = P.up[base + k];
    dn[k] = P.dn[base + k];
  }
  kids[lane] = P.kids[(size_t)cell * CHAIN_SEGS + lane];
  const int nseg = P.nseg[cell], nh = P.nh[cell];
  int start = 0, len = 0, kf = 0, kc = 0, ptop = -1, h = -1;
  if (lane < nseg) {
    const int *q = P.seg + ((size_t)cell * CHAIN_SEGS + lane) * 4;
    start = q[0];
    len = q[1];
    kf = q[2] & 0xffff;
    kc = q[2] >> 16;
    ptop = (q[3] & 0xffff) - 1;
    h = q[3] >> 16;
  }
  __syncthreads();
  for (int hh = 0; hh < nh; hh++) {
    if (h == hh) {
      int c = start;
      double dd = d[c], rr = r[c];
      for (int q = kf; q < kf + kc; q++) {
        const int j = kids[q];
        const double f = dn[j] / d[j];
        dd -= f * up[j];
        rr -= f * r[j];
      }
      for (int i = 1; i < len; i++, c++) {
        const double f = dn[c] / dd;
        d[c] = dd;
        r[c] = rr;
        dd = d[c + 1] - f * up[c];
        rr = r[c + 1] - f * rr;
      }
      d[c] = dd;
      r[c] = rr;
    }
    __syncthreads();
  }
  for (int hh = nh - 1; hh >= 0; hh--) {
    if (h == hh) {
      int c = start + len - 1;
      double x = (ptop >= 0) ? (r[c] - up[c] * r[ptop]) / d[c] : r[c] / d[c];
      r[c] = x;
      for (c = c - 1; c >= start; c--) {
        x = (r[c] - up[c] * x) / d[c];
        r[c] = x;
      }
    }
    __syncthreads();
  }
  for (int k = lane; k < comp; k += HINES_WAVE) v[P.perm[base + k]] = r[k];
}

struct HinesHost {
  std::vector<int> perm, meta, lev, lev_off;
  std::vector<double> up, dn;
  bool sym = true;   // up == dn at every position
  // the chain plan (every cell at most CHAIN_SEGS segments), else empty
  std::vector<int> cperm, seg, kids, nseg, nh;
  std::vector<double> cup, cdn;
};

// breadth-first distances from s over adj (ascending neighbour order)
void tree_bfs(const std::vector<std::vector<int>> &adj, int s, std::vector<int> &dist, std::vector<int> &par,
              std::vector<int> &order) {
  const int n = (int)adj.size();
  dist.assign(n, -1);
  par.assign(n, -1);
  order.clear();
  dist[s] = 0;
  order.push_back(s);
  for (size_t h = 0; h < order.size(); h++) {
    const int u = order[h];
    for (int w : adj[u])
      if (dist[w] < 0) {
        dist[w] = dist[u] + 1;
        par[w] = u;
        order.push_back(w);
      }
  }
}

// the chain plan of cell c from its tree (par, breadth-first order from the
// root); false when the cell has more than CHAIN_SEGS segments
template <typename At>
bool chain_plan(int c, int comp, const std::vector<std::vector<int>> &adj, const std::vector<int> &par,
                const std::vector<int> &order, At &at, HinesHost &H) {
  std::vector<int> nch(comp, 0), seg_of(comp, -1);
  for (int u = 0; u < comp; u++)
    if (par[u] >= 0) nch[par[u]]++;
  std::vector<std::vector<int>> segs;   // bottom -> top, deepest bottoms first
  for (int k = comp - 1; k >= 0; k--) {
    const int u = order[k];
    if (nch[u] == 1) continue;
    std::vector<int> nodes{u};
    for (int w = u; par[w] >= 0 && nch[par[w]] == 1;) {
      w = par[w];
      nodes.push_back(w);
    }
    if ((int)segs.size() == CHAIN_SEGS) return false;
    for (int w : nodes) seg_of[w] = (int)segs.size();
    segs.push_back(std::move(nodes));
  }
  const int ns = (int)segs.size();
  std::vector<int> ht(ns, 0);
  for (int s = 0; s < ns; s++)              // child segments were made earlier
    for (int w : adj[segs[s][0]])
      if (par[w] == segs[s][0]) ht[s] = std::max(ht[s], ht[seg_of[w]] + 1);
  std::vector<int> ord(ns);
  for (int s = 0; s < ns; s++) ord[s] = s;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return ht[x] < ht[y]; });
  std::vector<int> pos(comp), start(ns);
  int next = 0;
  for (int s : ord) {
    start[s] = next;
    for (int w : segs[s]) pos[w] = next++;
  }
  const size_t base = (size_t)c * comp;
  for (int u = 0; u < comp; u++) {
    const size_t k = base + pos[u];
    H.cperm[k] = (int)base + u;
    if (par[u] >= 0) {
      H.cup[k] = at((int)base + u, (int)base + par[u]);
      H.cdn[k] = at((int)base + par[u], (int)base + u);
    }
  }
  int nkid = 0, hmax = 0;
  for (int i = 0; i < ns; i++) {
    const int s = ord[i], bottom = segs[s][0], top = segs[s].back();
    const int kf = nkid;
    for (int w : adj[bottom])                 // the bottom's children in the level walk's order
      if (par[w] == bottom) H.kids[(size_t)c * CHAIN_SEGS + nkid++] = pos[w];
    int *q = &H.seg[((size_t)c * CHAIN_SEGS + i) * 4];
    q[0] = start[s];
    q[1] = (int)segs[s].size();
    q[2] = kf | ((nkid - kf) << 16);
    q[3] = (par[top] >= 0 ? pos[par[top]] + 1 : 0) | (ht[s] << 16);
    hmax = std::max(hmax, ht[s]);
  }
  H.nseg[c] = ns;
  H.nh[c] = hmax + 1;
  return true;
}

// the plan of every cell from the system's CSR (val = the path's signed
// matrix with its stored diagonal); throws when a cell is not a tree or is
// coupled to another cell
void build_hines(int n, int comp, const std::vector<int> &row, const std::vector<int> &col,
                 const std::vector<double> &val, HinesHost &H) {
  if (comp > HINES_MAX_COMP) throw Error("the per-cell tree solve is provided up to 1820 compartments per cell", -5);
  const size_t nc = (size_t)n * comp;
  H.perm.resize(nc);
  H.meta.resize(nc);
  H.up.assign(nc, 0.0);
  H.dn.assign(nc, 0.0);
  H.lev.clear();
  H.lev_off.assign(n + 1, 0);
  bool chain = true;
  H.cperm.assign(nc, 0);
  H.cup.assign(nc, 0.0);
  H.cdn.assign(nc, 0.0);
  H.seg.assign((size_t)n * CHAIN_SEGS * 4, 0);
  H.kids.assign((size_t)n * CHAIN_SEGS, 0);
  H.nseg.assign(n, 0);
  H.nh.assign(n, 0);
  std::vector<std::vector<int>> adj(comp);
  std::vector<int> dist, par, order, pos(comp);
  auto at = [&](int gi, int gj) -> double {
    for (int j = row[gi]; j < row[gi + 1]; j++)
      if (col[j] == gj) return val[j];
    throw Error("the per-cell tree solve needs a structurally symmetric matrix", -5);
  };
  for (int c = 0; c < n; c++) {
    const int base = c * comp;
    int edges = 0;
    for (int i = 0; i < comp; i++) {
      adj[i].clear();
      for (int j = row[base + i]; j < row[base + i + 1]; j++) {
        const int k = col[j] - base;
        if (k == i) continue;
        if (k < 0 || k >= comp)
          throw Error("the per-cell tree solve needs cells without intercell coupling (gap junctions)", -5);
        adj[i].push_back(k);
        edges++;
      }
      std::sort(adj[i].begin(), adj[i].end());
    }
    tree_bfs(adj, 0, dist, par, order);
    if ((int)order.size() != comp || edges != 2 * (comp - 1))
      throw Error("the per-cell tree solve needs every cell to be a tree", -5);
    // centre: the middle of a longest path (two sweeps)
    const int a = order.back();
    tree_bfs(adj, a, dist, par, order);
    std::vector<int> path;
    for (int u = order.back(); u >= 0; u = par[u]) path.push_back(u);
    const int root = path[path.size() / 2];
    tree_bfs(adj, root, dist, par, order);
    for (int k = 0; k < comp; k++) pos[order[k]] = k;
    H.lev_off[c] = (int)H.lev.size();
    int next_child = 1;
    for (int k = 0; k < comp; k++) {
      const int u = order[k];
      if (k == 0 || dist[u] != dist[order[k - 1]]) H.lev.push_back(k);
      int nch = 0;
      for (int w : adj[u])
        if (par[w] == u) nch++;
      if (nch > 255) throw Error("the per-cell tree solve takes at most 255 children per compartment", -5);
      H.meta[(size_t)base + k] = (next_child << 8) | nch;
      next_child += nch;
      H.perm[(size_t)base + k] = base + u;
      if (par[u] >= 0) {
        H.up[(size_t)base + k] = at(base + u, base + par[u]);
        H.dn[(size_t)base + k] = at(base + par[u], base + u);
        if (H.up[(size_t)base + k] != H.dn[(size_t)base + k]) H.sym = false;
      }
    }
    H.lev.push_back(comp);
    if (chain) chain = chain_plan(c, comp, adj, par, order, at, H);
  }
  H.lev_off[n] = (int)H.lev.size();
  if (!chain) {
    for (auto *x : {&H.cperm, &H.seg, &H.kids, &H.nseg, &H.nh}) x->clear();
    H.cup.clear();
    H.cdn.clear();
  }
}

template <typename T>
T *dal(size_t n) {
  void *p = nullptr;
  CB_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
  return (T *)p;
}
template <typename T>
void up(T *d, const std::vector<T> &h) {
  CB_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}
template <typename T>
T *dup(const std::vector<T> &h) {
  T *d = dal<T>(h.size());
  up(d, h);
  return d;
}

class CnEngine : public Engine {
 public:
  CnEngine(int n, int comp, const std::vector<double> &cm, const std::vector<double> &v,
           const std::vector<double> &iext, const std::vector<int> &row, const std::vector<int> &col,
           const std::vector<double> &rkc_val, const StimSpec &stim, bool bicg, bool be, bool direct, double dt)
      : n_(n), comp_(comp), nc_((int)cm.size()), be_(be), direct_(direct), dt_(be ? BE_DT : CN_DT), stim_(stim) {
    if (dt > 0.0) {
      const double k = H_MAX / dt;
      if (k < 1.0 || k > 4096.0 || std::fabs(k - std::nearbyint(k)) > 1e-9)
        throw Error("the CN / BE step must divide the 0.125 ms output interval into 1 to 4096 steps", -2);
      dt_ = dt;
      nsub_ = (int)std::nearbyint(k);
    }
    const int nnz = (int)col.size();
    // the reference's unflipped matrix = -(RKC-signed); halved for CN
    // (pkj_solve.cu:238-245 and the other xx_solve.cu files)
    std::vector<double> val(nnz);
    for (int j = 0; j < nnz; j++) val[j] = be ? -rkc_val[j] : -rkc_val[j] / 2.0;
    std::vector<int> dpos(nc_, -1);
    for (int i = 0; i < nc_; i++)
      for (int j = row[i]; j < row[i + 1]; j++)
        if (col[j] == i) dpos[i] = j;
    for (int i = 0; i < nc_; i++)
      if (dpos[i] < 0) throw Error("the CN and BE paths need a stored diagonal in every row");
    HinesHost hh;
    if (direct_) build_hines(n_, comp_, row, col, val, hh);   // before any device allocation
    create_stream(&st_);
    if (direct_ && !hh.nseg.empty()) {
      chain_ = has_chain_ = true;
      Q_.perm = dup(hh.cperm);
      Q_.up = dup(hh.cup);
      Q_.dn = dup(hh.cdn);
      Q_.seg = dup(hh.seg);
      Q_.kids = dup(hh.kids);
      Q_.nseg = dup(hh.nseg);
      Q_.nh = dup(hh.nh);
      Q_.comp = comp_;
    }
    if (direct_) {
      sym_ = hh.sym;
      P_.perm = dup(hh.perm);
      P_.up = dup(hh.up);
      P_.dn = dup(hh.dn);
      P_.meta = dup(hh.meta);
      P_.lev = dup(hh.lev);
      P_.lev_off = dup(hh.lev_off);
      P_.comp = comp_;
      S_.direct = 1;
    }
    // transpose for BiCG: rows of A in order, so transposed rows are in column order
    std::vector<int> trow(nc_ + 1, 0), tcol(nnz), tsrc(nnz), fill(nc_, 0);
    for (int j = 0; j < nnz; j++) trow[col[j] + 1]++;
    for (int i = 0; i < nc_; i++) trow[i + 1] += trow[i];
    for (int i = 0; i < nc_; i++)
      for (int j = row[i]; j < row[i + 1]; j++) {
        const int c = col[j], t = trow[c] + fill[c]++;
        tcol[t] = i;
        tsrc[t] = j;
      }
    S_.row = dup(row);
    S_.col = dup(col);
    S_.dpos = dup(dpos);
    S_.val = dup(val);
    S_.trow = dup(trow);
    S_.tcol = dup(tcol);
    S_.tsrc = dup(tsrc);
    S_.cm = dup(cm);
    S_.v = dup(v);
    S_.vold = dup(v);
    S_.iext = dup(iext);
    S_.gamma1 = dal<double>(nc_);
    S_.omega1 = dal<double>(nc_);
    for (double **p : {&S_.diag, &S_.b, &S_.r, &S_.ap, &S_.rs, &S_.atps, &S_.p[0], &S_.p[1], &S_.ps[0], &S_.ps[1]})
      *p = dal<double>(nc_);
    nb_ = (nc_ + NB - 1) / NB;
    S_.part = dal<double>(nb_);
    S_.scal = dal<double>(S_NSCAL);
    S_.flag = dal<int>(F_NFLAG);
    CB_HIP(hipMemset(S_.scal, 0, sizeof(double) * S_NSCAL));
    CB_HIP(hipMemset(S_.flag, 0, sizeof(int) * F_NFLAG));
    S_.nc = nc_;
    S_.bicg = bicg ? 1 : 0;
    S_.dt = dt_;
    amp_ = dup(stim_.amp);
    rec_ = dal<double>(n_);
  }
  ~CnEngine() override {
    (void)hipStreamSynchronize(st_);
    for (const void *p : {(const void *)S_.row, (const void *)S_.col, (const void *)S_.dpos, (const void *)S_.val,
                          (const void *)S_.trow, (const void *)S_.tcol, (const void *)S_.tsrc, (const void *)S_.cm,
                          (const void *)S_.v, (const void *)S_.vold, (const void *)S_.iext, (const void *)S_.gamma1,
                          (const void *)S_.omega1, (const void *)S_.diag, (const void *)S_.b, (const void *)S_.r,
                          (const void *)S_.ap, (const void *)S_.rs, (const void *)S_.atps, (const void *)S_.p[0],
                          (const void *)S_.p[1], (const void *)S_.ps[0], (const void *)S_.ps[1],
                          (const void *)S_.part, (const void *)S_.scal, (const void *)S_.flag, (const void *)amp_,
                          (const void *)rec_, (const void *)P_.perm, (const void *)P_.up, (const void *)P_.dn,
                          (const void *)P_.meta, (const void *)P_.lev, (const void *)P_.lev_off,
                          (const void *)Q_.perm, (const void *)Q_.up, (const void *)Q_.dn, (const void *)Q_.seg,
                          (const void *)Q_.kids, (const void *)Q_.nseg, (const void *)Q_.nh})
      if (p) (void)hipFree(const_cast<void *>(p));
    (void)hipStreamDestroy(st_);
  }
  int nc() const override { return nc_; }
  int ncell() const override { return n_; }
  void sync() override { CB_HIP(hipStreamSynchronize(st_)); }
  Stats stats() const override {
    Stats s = st;
    int f[F_NFLAG];
    CB_HIP(hipStreamSynchronize(st_));
    CB_HIP(hipMemcpy(f, S_.flag, sizeof(f), hipMemcpyDeviceToHost));
    s.stages = f[F_TOTAL];
    return s;
  }
  void set_current(double t) override {
    k_cn_stim<<<(n_ + 127) / 128, 128, 0, st_>>>(S_.iext, amp_, comp_, stim_.comp, stim_.t0, stim_.t1, t, n_);
  }
  // one exact tree solve of (diag, b) into v: by unbranched runs or level
  // by level (the same operations per compartment, bit-identical results)
  void tree_solve(bool by_runs, const double *diag, const double *b, double *v) {
    if (by_runs) {
      const size_t lds = (size_t)comp_ * 4 * sizeof(double) + CHAIN_SEGS * sizeof(int);
      k_hines_chain<<<n_, HINES_WAVE, lds, st_>>>(Q_, diag, b, v);
    } else if (sym_) {
      const size_t lds = (size_t)comp_ * (3 * sizeof(double) + sizeof(int));
      k_hines<true><<<n_, HINES_WAVE, lds, st_>>>(P_, diag, b, v);
    } else {
      const size_t lds = (size_t)comp_ * (4 * sizeof(double) + sizeof(int));
      k_hines<false><<<n_, HINES_WAVE, lds, st_>>>(P_, diag, b, v);
    }
  }

  // Tree engine: when every cell is at most CHAIN_SEGS unbranched runs both
  // solves apply; each runs trial_steps times on a copy of the current system
  // (the population's state is not touched) and the faster is kept.
  // Returns the lanes per cell (one wavefront either way).
  int autotune(int trial_steps, double t0) override {
    (void)t0;
    if (!direct_) return 0;
    if (!has_chain_ || trial_steps <= 0) return HINES_WAVE;
    double *d = dal<double>(nc_), *b = dal<double>(nc_), *v = dal<double>(nc_);
    CB_HIP(hipMemcpyAsync(d, S_.diag, sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    CB_HIP(hipMemcpyAsync(b, S_.b, sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    hipEvent_t e0, e1;
    CB_HIP(hipEventCreate(&e0));
    CB_HIP(hipEventCreate(&e1));
    float ms[2] = {0.0f, 0.0f};
    for (int by_runs = 0; by_runs < 2; by_runs++) {
      tree_solve(by_runs != 0, d, b, v);   // warm
      CB_HIP(hipEventRecord(e0, st_));
      for (int k = 0; k < trial_steps; k++) tree_solve(by_runs != 0, d, b, v);
      CB_HIP(hipEventRecord(e1, st_));
      CB_HIP(hipEventSynchronize(e1));
      CB_HIP(hipEventElapsedTime(&ms[by_runs], e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d);
    (void)hipFree(b);
    (void)hipFree(v);
    chain_ = ms[1] <= ms[0];
    tune_ms_[0] = ms[0];
    tune_ms_[1] = ms[1];
    return HINES_WAVE;
  }
  std::string mapping() const override {
    if (!direct_) return nb_ <= KRY_SMALL_BLOCKS ? "Krylov solve in one workgroup" : "Krylov solve, two launches per iteration";
    std::string m = chain_ ? "tree solve by unbranched runs, one lane per run, one wavefront per cell"
                           : "tree solve level by level, one wavefront per cell";
    if (tune_ms_[0] > 0.0 || tune_ms_[1] > 0.0)
      m += " (trial: level walk " + std::to_string(tune_ms_[0]) + " ms, run walk " + std::to_string(tune_ms_[1]) +
           " ms)";
    return m;
  }

  // one CN (BE) step of CN_DT (BE_DT): xx_solve_update_v with that solver
  void advance() override {
    prep();
    if (direct_) {
      tree_solve(chain_, S_.diag, S_.b, S_.v);
    } else if (nb_ <= KRY_SMALL_BLOCKS) {
      k_kry_small<<<1, NB, 0, st_>>>(S_);
    } else {
      for (int k = 0; k < 100; k++) {
        k_kry_a<<<nb_, NB, 0, st_>>>(S_, k);
        k_kry_b<<<nb_, NB, 0, st_>>>(S_, k);
      }
    }
    st.steps++;
  }
  // the reference's sub-step loop for each H_MAX output interval
  // (main.cu:193-200: set_current(t + l_dt), one CN step, l_dt += DT while
  // l_dt < H_MAX); records compartment rec_comp of every cell after it
  void run(int nsteps, double t0, int rec_comp, double *rec_out) override {
    if (nsteps <= 0) return;
    const int per = rec_comp >= 0 ? 1 : comp_;
    const int words = (n_ + 31) / 32;
    const bool raster = raster_comp_ >= 0;
    double *drec = nullptr;
    uint32_t *dbits = nullptr;
    if (rec_out) drec = dal<double>((size_t)nsteps * n_ * per);
    if (raster) {
      dbits = dal<uint32_t>((size_t)nsteps * words);
      CB_HIP(hipMemsetAsync(dbits, 0, sizeof(uint32_t) * (size_t)nsteps * words, st_));
    }
    for (int k = 0; k < nsteps; k++) {
      const double t = t0 + k * H_MAX;
      if (nsub_ > 0) {
        // a set step: exactly H_MAX / dt steps, at whole multiples of dt
        for (int j = 0; j < nsub_; j++) {
          set_current(t + j * dt_);
          advance();
        }
      } else {
        double l_dt = 0.0;
        while (l_dt < H_MAX) {
          set_current(t + l_dt);
          advance();
          l_dt += dt_;
        }
      }
      st.macro_steps++;
      if (drec || dbits)
        record_interval(st_, S_.v, drec ? drec + (size_t)k * n_ * per : nullptr, rec_comp,
                        dbits ? dbits + (size_t)k * words : nullptr);
    }
    std::vector<uint32_t> bits((size_t)(raster ? nsteps : 0) * words);
    if (drec)
      CB_HIP(hipMemcpyAsync(rec_out, drec, sizeof(double) * (size_t)nsteps * n_ * per, hipMemcpyDeviceToHost, st_));
    if (dbits) CB_HIP(hipMemcpyAsync(bits.data(), dbits, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, st_));
    CB_HIP(hipStreamSynchronize(st_));
    if (drec) CB_HIP(hipFree(drec));
    if (dbits) CB_HIP(hipFree(dbits));
    if (raster) append_spikes(bits, nsteps, words, t0, H_MAX);
  }
  const double *device_vnew() const override { return S_.v; }
  const double *device_v() const override { return S_.v; }
  hipStream_t stream() const override { return st_; }
  // add_*_val of the CN and BE paths (GR, GO, PKJ; the BE Purkinje path
  // couples none); the inferior olive has no synapses
  void set_synapses(const SynIn &in) override {
    if (in.on() && !takes_synapses()) throw Error("synaptic input is not provided for this population", -5);
    S_.syn = in;
  }

  void read(int field, double *dst) override {
    const double *src = nullptr;
    switch (field) {
      case F_VNEW:
      case F_V: src = S_.v; break;
      case F_IEXT: src = S_.iext; break;
      case F_B: src = S_.b; break;
      case F_VOLD: src = S_.vold; break;
      default: src = model_field(field);
    }
    if (!src) throw Error("field not available for this population");
    CB_HIP(hipStreamSynchronize(st_));
    CB_HIP(hipMemcpy(dst, src, sizeof(double) * nc_, hipMemcpyDeviceToHost));
  }

 protected:
  virtual void prep() = 0;
  virtual const double *model_field(int) const { return nullptr; }
  virtual bool takes_synapses() const { return true; }
  // the half conductance / reversal sums of the initial state
  // (xx_cnm_vec_initialize): gamma1 = gs / 2, omega1 = es / 2
  void init_half_sums(const std::vector<double> &gs, const std::vector<double> &es) {
    std::vector<double> g1(nc_), o1(nc_);
    for (int i = 0; i < nc_; i++) {
      g1[i] = gs[i] / 2.0;
      o1[i] = es[i] / 2.0;
    }
    up(S_.gamma1, g1);
    up(S_.omega1, o1);
  }
  int n_, comp_, nc_, nb_ = 1;
  bool be_, direct_;
  double dt_;
  int nsub_ = 0;   // steps per output interval when the step is set (0: the reference's loop)
  StimSpec stim_;
  CnSys S_{};
  HinesPlan P_{};
  ChainPlan Q_{};
  bool sym_ = false;     // the level walk stages one off-diagonal plane
  bool chain_ = false;   // the solve by unbranched runs is in use
  bool has_chain_ = false;   // every cell at most CHAIN_SEGS runs: the run walk applies
  double tune_ms_[2] = {0.0, 0.0};
  double *amp_ = nullptr, *rec_ = nullptr;
  hipStream_t st_{};
  Stats st;
};

class CnPkj : public CnEngine {
 public:
  CnPkj(const PkjHost &H, const StimSpec &stim, bool be, bool direct, double dt)
      : CnEngine(H.n, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, stim, false, be, direct, dt) {
    using namespace pkj;
    M_.ca = dup(H.ca);
    M_.rev = dup(H.rev);
    M_.shell = dup(H.shell);
    M_.area = dup(H.area);
    M_.cond = dup(H.cond);
    M_.gate = dup(H.gate);
    M_.compart = dup(H.compart);
    if (be) return;
    std::vector<double> gs(nc_), es(nc_);
    double c[NCOND], g[NGATE];
    for (int i = 0; i < nc_; i++) {
      for (int k = 0; k < NCOND; k++) c[k] = H.cond[(size_t)k * nc_ + i];
      for (int k = 0; k < NGATE; k++) g[k] = H.gate[(size_t)k * nc_ + i];
      cn_ge(c, g, H.rev[i], H.iext[i], gs[i], es[i]);
    }
    init_half_sums(gs, es);
  }
  ~CnPkj() override {
    (void)hipStreamSynchronize(st_);
    for (const void *p : {(const void *)M_.ca, (const void *)M_.rev, (const void *)M_.shell, (const void *)M_.area,
                          (const void *)M_.cond, (const void *)M_.gate, (const void *)M_.compart})
      (void)hipFree(const_cast<void *>(p));
  }

 private:
  void prep() override {
    if (be_) k_be_prep_pkj<<<nb_, NB, 0, st_>>>(S_, M_);
    else k_cn_prep_pkj<<<nb_, NB, 0, st_>>>(S_, M_);
  }
  const double *model_field(int f) const override {
    if (f == F_CA) return M_.ca;
    if (f == F_REV) return M_.rev;
    if (f >= F_GATE0 && f < F_GATE0 + pkj::NGATE) return M_.gate + (size_t)(f - F_GATE0) * nc_;
    return nullptr;
  }
  PkjCn M_{};
};

class CnIo : public CnEngine {
 public:
  CnIo(const IoHost &H, const StimSpec &stim, bool be, bool direct, double dt)
      : CnEngine(H.n, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, stim, true, be, direct, dt) {
    using namespace io;
    M_.ca = dup(H.ca);
    M_.cond = dup(H.cond);
    M_.gate = dup(H.gate);
    M_.compart = dup(H.compart);
    M_.nbr = dup(H.nbr);
    M_.n = H.n;
    if (be) return;
    std::vector<double> gs(nc_), es(nc_);
    double c[NCOND], g[NGATE];
    for (int i = 0; i < nc_; i++) {
      for (int k = 0; k < NCOND; k++) c[k] = H.cond[(size_t)k * nc_ + i];
      for (int k = 0; k < NGATE; k++) g[k] = H.gate[(size_t)k * nc_ + i];
      cn_ge(c, g, H.iext[i], gs[i], es[i]);
    }
    init_half_sums(gs, es);
  }
  ~CnIo() override {
    (void)hipStreamSynchronize(st_);
    for (const void *p : {(const void *)M_.ca, (const void *)M_.cond, (const void *)M_.gate,
                          (const void *)M_.compart, (const void *)M_.nbr})
      (void)hipFree(const_cast<void *>(p));
  }

 private:
  void prep() override {
    if (be_) k_be_prep_io<<<nb_, NB, 0, st_>>>(S_, M_);
    else k_cn_prep_io<<<nb_, NB, 0, st_>>>(S_, M_);
  }
  bool takes_synapses() const override { return false; }
  const double *model_field(int f) const override {
    if (f == F_CA) return M_.ca;
    if (f >= F_GATE0 && f < F_GATE0 + io::NGATE) return M_.gate + (size_t)(f - F_GATE0) * nc_;
    return nullptr;
  }
  IoCn M_{};
};

class CnGo : public CnEngine {
 public:
  CnGo(const GoHost &H, const StimSpec &stim, bool be, bool direct, double dt)
      : CnEngine(H.n, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, stim, false, be, direct, dt) {
    using namespace go;
    M_.ca = dup(H.ca);
    M_.ca2 = dup(H.ca2);
    M_.rev = dup(H.rev);
    M_.ca_old = dup(H.ca_old);
    M_.area = dup(H.area);
    M_.cond = dup(H.cond);
    M_.gate = dup(H.gate);
    if (be) return;
    std::vector<double> gs(nc_), es(nc_);
    double c[NCOND], g[NGATE];
    for (int i = 0; i < nc_; i++) {
      for (int k = 0; k < NCOND; k++) c[k] = H.cond[(size_t)k * nc_ + i];
      for (int k = 0; k < NGATE; k++) g[k] = H.gate[(size_t)k * nc_ + i];
      cn_ge(c, g, H.rev[i], H.iext[i], gs[i], es[i]);
    }
    init_half_sums(gs, es);
  }
  ~CnGo() override {
    (void)hipStreamSynchronize(st_);
    for (const void *p : {(const void *)M_.ca, (const void *)M_.ca2, (const void *)M_.rev, (const void *)M_.ca_old,
                          (const void *)M_.area, (const void *)M_.cond, (const void *)M_.gate})
      (void)hipFree(const_cast<void *>(p));
  }

 private:
  void prep() override {
    if (be_) k_be_prep_go<<<nb_, NB, 0, st_>>>(S_, M_);
    else k_cn_prep_go<<<nb_, NB, 0, st_>>>(S_, M_);
  }
  const double *model_field(int f) const override {
    switch (f) {
      case F_CA: return M_.ca;
      case F_REV: return M_.rev;
      case F_CA2: return M_.ca2;
      case F_CA_OLD: return M_.ca_old;
    }
    if (f >= F_GATE0 && f < F_GATE0 + go::NGATE) return M_.gate + (size_t)(f - F_GATE0) * nc_;
    return nullptr;
  }
  GoCn M_{};
};

class CnGr : public CnEngine {
 public:
  CnGr(const GrHost &H, const StimSpec &stim, bool be, bool direct, double dt)
      : CnEngine(H.n, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, stim, false, be, direct, dt) {
    using namespace gr;
    ca_ = dup(H.ca);
    M_.ca = ca_;
    cond_ = dup(H.cond);
    M_.cond = cond_;
    M_.gate = dup(H.gate);
    compart_ = dup(H.compart);
    M_.compart = compart_;
    const NaTable T = na_table();
    na_ = dal<NaType>(NTYPE);
    CB_HIP(hipMemcpy(na_, T.t, sizeof(T.t), hipMemcpyHostToDevice));
    M_.na = na_;
    if (be) return;
    std::vector<double> gs(nc_), es(nc_);
    double c[NCOND], g[NGATE];
    for (int i = 0; i < nc_; i++) {
      for (int k = 0; k < NCOND; k++) c[k] = H.cond[(size_t)k * nc_ + i];
      for (int k = 0; k < NGATE; k++) g[k] = H.gate[(size_t)k * nc_ + i];
      cn_ge(c, g, H.iext[i], gs[i], es[i]);
    }
    init_half_sums(gs, es);
  }
  ~CnGr() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)ca_, (void *)cond_, (void *)M_.gate, (void *)compart_, (void *)na_}) (void)hipFree(p);
  }

 private:
  void prep() override {
    if (be_) k_be_prep_gr<<<nb_, NB, 0, st_>>>(S_, M_);
    else k_cn_prep_gr<<<nb_, NB, 0, st_>>>(S_, M_);
  }
  const double *model_field(int f) const override {
    if (f == F_CA) return ca_;
    if (f >= F_GATE0 && f < F_GATE0 + gr::NGATE) return M_.gate + (size_t)(f - F_GATE0) * nc_;
    return nullptr;
  }
  GrCn M_{};
  double *ca_ = nullptr, *cond_ = nullptr;
  int8_t *compart_ = nullptr;
  gr::NaType *na_ = nullptr;
};

}  // namespace

Engine *make_cn_pkj(const PkjHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnPkj(h, s, false, direct, dt);
}
Engine *make_cn_io(const IoHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnIo(h, s, false, direct, dt);
}
Engine *make_cn_go(const GoHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnGo(h, s, false, direct, dt);
}
Engine *make_cn_gr(const GrHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnGr(h, s, false, direct, dt);
}
Engine *make_be_pkj(const PkjHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnPkj(h, s, true, direct, dt);
}
Engine *make_be_io(const IoHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnIo(h, s, true, direct, dt);
}
Engine *make_be_go(const GoHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnGo(h, s, true, direct, dt);
}
Engine *make_be_gr(const GrHost &h, const StimSpec &s, bool direct, double dt) {
  return new CnGr(h, s, true, direct, dt);
}

}  // namespace cbrkc
