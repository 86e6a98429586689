// Plain FP64 RKC with one population-global controller: the reference's
// semantics (solve_rkc.cu) on MI355X.
//
// One thread per compartment, structure-of-arrays FP64 state in HBM.  The
// differences from the reference are in how the work is laid out, not in
// what is computed:
//   - the RHS (dydt) and the stage combination (calcY) are one kernel, so a
//     stage reads each vector once (solve_rkc.cu:857-900 uses two launches);
//   - the error map, f(Y_m) and the first reduction level are one kernel;
//   - every norm is a deterministic 256-wide block tree (no atomics, no
//     library dot products), so runs are bit-reproducible;
//   - the ion update and the frozen-current rebuild are one kernel;
//   - the host controller reads back one double per step (the error norm)
//     and one per power-method iteration, nothing else.
// The controller (rkclow/step/rkcrho) is restated on the host, in the
// reference's scalar arithmetic order, with the same quirks (see the oracle
// header for the list).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include <cmath>
#include <cstring>
#include <memory>

#include "detsum.hpp"
#include "engine.hpp"
#include "philox.hpp"
#include "go_model.hpp"
#include "gr_model.hpp"
#include "io_model.hpp"
#include "ion_state.hpp"
#include "pkj_model.hpp"
#include "record.hpp"

namespace cbrkc {

__global__ void __launch_bounds__(256) k_reduce256(const double *__restrict__ in, double *__restrict__ out, int n) {
  __shared__ double s[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double x = (i < n) ? in[i] : 0.0;
  double r = block_tree_sum(x, s);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

namespace {

// the last stage of the deterministic reduction (n <= 256 partials, the tree
// of k_reduce256's single block) written straight into mapped host memory,
// with the flag word beside it (null: 0), so that a step's read-back is one
// stream synchronisation and no copy commands
__global__ void __launch_bounds__(256) k_reduce_to_host(const double *__restrict__ in, int n,
                                                        const int *__restrict__ bad, double *hout) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  const double x = (t < n) ? in[t] : 0.0;
  const double r = block_tree_sum(x, s);
  if (t == 0) {
    hout[0] = r;
    hout[1] = bad ? (double)*bad : 0.0;
  }
}

constexpr double UROUND = 2.22e-16;   // solve_rkc.cuh:22
constexpr double H_MAX = 0.125;       // solve_rkc.cuh:23

__device__ inline double rhs(int i, const int *__restrict__ row, const int *__restrict__ col,
                             const double *__restrict__ val, const double *__restrict__ b,
                             const double *__restrict__ cm, const double *__restrict__ y) {
  // (-b_i + sum_j val_j (y_col_j - y_i)) / Cm_i in CSR column order (solve_rkc.cu:177-196)
  double d = -b[i];
  const double yi = y[i];
  const int e = row[i + 1];
  for (int j = row[i]; j < e; j++) d += val[j] * (y[col[j]] - yi);
  return d / cm[i];
}

struct Csr {
  const int *row, *col;
  const double *val;
};

// vt2 = yn; vt1 = (h mus) fn + yn   (solve_rkc.cu:843-847)
__global__ void __launch_bounds__(256) k_first(const double *__restrict__ fn, const double *__restrict__ yn,
                                               double *__restrict__ vt1, double *__restrict__ vt2, double hmus, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) {
    double y0 = yn[i];
    vt2[i] = y0;
    vt1[i] = hmus * fn[i] + y0;
  }
}

// Y_j = mu Y_{j-1} + nu Y_{j-2} + (1-mu-nu) Y_0 + h mus (F(Y_{j-1}) - a_{j-1} F_0)
__global__ void __launch_bounds__(256) k_stage(Csr A, const double *__restrict__ b, const double *__restrict__ cm,
                                               const double *__restrict__ vt1, const double *__restrict__ vt2,
                                               const double *__restrict__ yn, const double *__restrict__ fn,
                                               double *__restrict__ y, double mu, double nu, double omn, double hmus,
                                               double ajm1, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) {
    double F = rhs(i, A.row, A.col, A.val, b, cm, vt1);
    y[i] = mu * vt1[i] + nu * vt2[i] + omn * yn[i] + hmus * (F - ajm1 * fn[i]);
  }
}

// vt1 = F(y); partial[block] = sum ((0.8 (yn - y) + 0.4 h (fn + vt1)) / wt)^2
// (solve_rkc.cu:979-982, 58-73)
__global__ void __launch_bounds__(256) k_err(Csr A, const double *__restrict__ b, const double *__restrict__ cm,
                                             const double *__restrict__ y, const double *__restrict__ yn,
                                             const double *__restrict__ fn, double *__restrict__ vt1, double at,
                                             double rt, double h, double *__restrict__ partial, int *__restrict__ bad,
                                             int nc) {
  __shared__ double s[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  double term = 0.0;
  if (i < nc) {
    double F = rhs(i, A.row, A.col, A.val, b, cm, y);
    vt1[i] = F;
    double yi = y[i], yni = yn[i];
    double ay = fabs(yi), ayn = fabs(yni);
    double wt = at + rt * ((ay > ayn) ? ay : ayn);
    double est = 0.8 * (yni - yi) + 0.4 * h * (fn[i] + F);
    if (wt == 0.0) *bad = 1;
    term = (est / wt) * (est / wt);
  }
  double r = block_tree_sum(term, s);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// partial[block] = sum x^2
__global__ void __launch_bounds__(256) k_sumsq(const double *__restrict__ x, double *__restrict__ partial, int nc) {
  __shared__ double s[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  double t = 0.0;
  if (i < nc) {
    double v = x[i];
    t = v * v;
  }
  double r = block_tree_sum(t, s);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// C = s A + B (vec_sA_plus_B, solve_rkc.cu:143-148)
__global__ void __launch_bounds__(256) k_saxpb(double s, const double *__restrict__ A, const double *__restrict__ B,
                                               double *__restrict__ C, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) C[i] = s * A[i] + B[i];
}

__global__ void __launch_bounds__(256) k_scale(double s, const double *__restrict__ A, double *__restrict__ C, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) C[i] = s * A[i];
}

__global__ void __launch_bounds__(256) k_fill(double s, double *__restrict__ C, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) C[i] = s;
}

// one power-method iteration: vt2 = F(vt1); dammy = -fn + vt2;
// partial = sum dammy^2 (solve_rkc.cu:645-661)
__global__ void __launch_bounds__(256) k_rho_iter(Csr A, const double *__restrict__ b, const double *__restrict__ cm,
                                                  const double *__restrict__ vt1, const double *__restrict__ fn,
                                                  double *__restrict__ vt2, double *__restrict__ dammy,
                                                  double *__restrict__ partial, int nc) {
  __shared__ double s[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  double t = 0.0;
  if (i < nc) {
    double F = rhs(i, A.row, A.col, A.val, b, cm, vt1);
    vt2[i] = F;
    double d = -1.0 * fn[i] + F;
    dammy[i] = d;
    t = d * d;
  }
  double r = block_tree_sum(t, s);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// vt1 = alpha (-vt2 + fn) + yn  (the two vec_sA_plus_B of solve_rkc.cu:725-726)
__global__ void __launch_bounds__(256) k_rho_next(double alpha, const double *__restrict__ vt2,
                                                  const double *__restrict__ fn, const double *__restrict__ yn,
                                                  double *__restrict__ vt1, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) {
    double t = -1.0 * vt2[i] + fn[i];
    vt1[i] = alpha * t + yn[i];
  }
}

__global__ void k_flip1(double *__restrict__ v, const double *__restrict__ y, int id) {
  v[id] = 2.0 * y[id] - v[id];   // rand_elements_change, solve_rkc.cu:22-27
}

// set_initial_absh weights: dammy = (vt2 - fn) / (at + rt |yn|) as the
// reference's host loop + Hadamard_product (solve_rkc.cu:525-543)
__global__ void __launch_bounds__(256) k_absh_w(const double *__restrict__ vt2, const double *__restrict__ fn,
                                                const double *__restrict__ yn, double at, double rt,
                                                double *__restrict__ dammy, int *__restrict__ bad, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) {
    double wt = at + rt * fabs(yn[i]);
    if (wt == 0.0) *bad = 1;
    double w = 1.0 / wt;
    dammy[i] = (vt2[i] - fn[i]) * w;
  }
}

// rkcint, solve_rkc.cu:794-813
__global__ void __launch_bounds__(256) k_rkcint(double *__restrict__ vnew, double a1, double a2, double b1, double b2,
                                                const double *__restrict__ vt1, const double *__restrict__ vt2,
                                                const double *__restrict__ yn, const double *__restrict__ fn, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) vnew[i] = a1 * vt1[i] + a2 * yn[i] + b1 * vt2[i] + b2 * fn[i];
}

struct PkjState {
  double *v, *ca, *rev, *iext, *shell;
  const double *cond;   // [NCOND][nc]
  double *gate;         // [NGATE][nc]
  double *b;
  const double *syn_g;  // may be null
  double syn_e;
  int nc;
  const double *net_s;  // synaptic current of the macro step (granule input), may be null
  double net_decay;     // exp(-dt / 0.6), dt the ion interval
};

// pkj_update_ion_RKC (pkj_ion.cu:304-358) followed by pkj_update_matrix
// (solve_rkc.cu:369-395), fused: one pass over the compartment state.
// With do_ion == 0 only the current is rebuilt (RKC_vec_initialize).
__global__ void __launch_bounds__(256) k_pkj_ion_b(PkjState S, const double *__restrict__ vnew, double DT, int do_ion) {
  using namespace pkj;
  const int nc = S.nc;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  double g[NGATE], c[NCOND];
#pragma unroll
  for (int k = 0; k < NGATE; k++) g[k] = S.gate[(size_t)k * nc + i];
#pragma unroll
  for (int k = 0; k < NCOND; k++) c[k] = S.cond[(size_t)k * nc + i];
  const double vn = vnew[i];
  double rev = S.rev[i];
  if (do_ion) {
    const double vo = S.v[i];
    const double vm = 0.5 * (vn + vo);
    double ca = S.ca[i];
    double caq = ca_rk4(vo, ca, rev, S.shell[i], c[g_CaP] * g[m_CaP] * g[h_CaP], c[g_CaT] * g[m_CaT] * g[h_CaT], DT,
                        i % 1599 == 0);
    gates_cn<double>(vm, caq, DT, g);
#pragma unroll
    for (int k = 0; k < NGATE; k++) S.gate[(size_t)k * nc + i] = g[k];
    S.ca[i] = ca;
    S.rev[i] = rev;
    S.v[i] = vn;
  }
  double bb = current<double>(vn, S.iext[i], rev, c, g);
  if (S.syn_g) bb += S.syn_g[i] * (vn - S.syn_e);
  // pkj_update_matrix_new (solve_rkc.cu:358-367, 1245-1249)
  if (S.net_s) bb += S.net_s[i] * S.net_decay;
  S.b[i] = bb;
}


// io_update_ion_RKC (io_ion.cu:90-127), io_update_matrix (solve_rkc.cu:412-434)
// and, on each dendrite, the gap currents of io_gap_update / add_io_gap_val
// (gap.cu:149-160, solve_rkc.cu:400-410), fused.  The gap current uses the
// partners' voltages after the update (elem[v] = vnew there) and is added in
// partner order 0..3 (the reference adds with atomics in no fixed order).
__global__ void __launch_bounds__(256) k_io_ion_b(IoState S, const double *__restrict__ vnew, double DT, int do_ion) {
  using namespace io;
  const int nc = S.nc;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const double vn = vnew[i];
  double bb = io_ion_current(S, i, vn, DT, do_ion != 0);
  if (do_ion && S.n > 1 && i % COMP == 0) {
    const int cell = i / COMP;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int pre = S.nbr[cell * 4 + s];
      const double diff_v = (pre > -1) ? (vn - vnew[pre * COMP]) : 0.0;
      bb += gap_current<double>(diff_v);
    }
  }
  if (S.syn_g) bb += S.syn_g[i] * (vn - S.syn_e);
  S.b[i] = bb;
}


// go_update_ion_RKC_exp_imp (go_ion.cu:379-437), go_KAHP_update_2order
// (go_ion.cu:88-213) and go_update_matrix (solve_rkc.cu:288-314), fused: the
// three reference kernels touch one compartment's state each, in this order
// (solve_rkc.cu:1174-1179).  The intercell gap junctions are in the cable
// matrix, not here.
__global__ void __launch_bounds__(256) k_go_ion_b(GoState S, const double *__restrict__ vnew, double DT, int do_ion) {
  const int nc = S.nc;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const double vn = vnew[i];
  double bb = go_ion_current(S, i, vn, DT, do_ion != 0);
  if (S.syn_g) bb += S.syn_g[i] * (vn - S.syn_e);
  // go_update_matrix_new (solve_rkc.cu:279-286, 1181)
  if (S.net_s) bb += S.net_s[i];
  S.b[i] = bb;
}


// gr_Na_update_2order (gr_ion.cu:45-207), gr_update_ion_RKC_exp_imp
// (gr_ion.cu:461-480) and gr_update_matrix (solve_rkc.cu:226-249), fused in
// the reference's order (solve_rkc.cu:1111-1116): the Na scheme sees the
// voltage of the previous ion update and the new one, then the gates move and
// the stored voltage becomes the new one.
__global__ void __launch_bounds__(256) k_gr_ion_b(GrState S, const double *__restrict__ vnew, double DT, int do_ion) {
  const int nc = S.nc;
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const double vn = vnew[i];
  double bb = gr_ion_current(S, i, vn, DT, do_ion != 0);
  if (S.syn_g) bb += S.syn_g[i] * (vn - S.syn_e);
  // add_mfgr_val / add_gogr_val (solve_rkc.cu:200-224, 1117-1118), one
  // synapse after the other
  if (S.net.on())
    for (int j = S.net.row[i]; j < S.net.row[i + 1]; j++) {
      double E;
      const double gs = syn_entry(S.net, j, E);
      bb += gs * (vn - E);
    }
  S.b[i] = bb;
}

// xx_set_current shape: one compartment per cell, a time window (pkj.cu:333-344)
__global__ void k_set_current(double *__restrict__ iext, const double *__restrict__ amp, int comp, int stim_comp,
                              double t0, double t1, double t, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    double *x = &iext[stim_comp + (size_t)c * comp];
    *x = 0;
    if (t0 <= t && t < t1) *x = amp[c];
  }
}

// Poisson conductance drive, advanced once per macro step k:
//   g <- g exp(-H/tau) + g_event [u < p],  u from Philox(ctr = (i, k, 0x5359, 0), key = seed)
__global__ void __launch_bounds__(256) k_poisson(double *__restrict__ g, double decay, double p, double g_event,
                                                 uint32_t k0, uint32_t k1, uint32_t step, int nc) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) {
    u32x4 r = philox4x32_10((uint32_t)i, step, 0x5359u, 0u, k0, k1);
    double u = u01_53(r.v[0], r.v[1]);
    g[i] = g[i] * decay + ((u < p) ? g_event : 0.0);
  }
}

inline int nblk(int n) { return (n + 255) / 256; }

template <typename T>
T *dalloc(size_t n) {
  void *p = nullptr;
  CB_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
  return (T *)p;
}

template <typename T>
void h2d(T *d, const std::vector<T> &h) {
  CB_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

double fmax_ori(double a, double b) { return (a > b) ? a : b; }
int nint_ori(double x) {
  if (x - (double)(int)x < 0.5) return (int)(double)(int)x;
  return (int)(double)(int)(x + 0.9);
}

enum { Y = 0, YN, FN, VT1, VT2, EDE, VNEW, NVEC };
enum { MMAX = 0, ABSH, SPRAD, NACCPT, ATOL, RTOL, IDID, GT, GT_OLD, LT, TEND, ERROLD, HOLD, NSTSIG, NOTHERS };

// The RKC machinery shared by every cell type: vectors, the cable CSR, the
// host controller (rkclow / step / rkcrho in the reference's scalar order)
// and the macro-step driver.  A cell type supplies the frozen-current and
// ion-update kernel (model_b) and its ion-update trigger.
class GlobalRkc : public Engine {
 public:
  GlobalRkc(int n, int nc, int comp, const std::vector<double> &cm, const std::vector<double> &v,
            const std::vector<double> &iext, const std::vector<int> &row, const std::vector<int> &col,
            const std::vector<double> &val, const StimSpec &stim, const PoissonSpec &syn)
      : n_(n), nc_(nc), comp_(comp), stim_(stim), syn_(syn) {
    create_stream(&st_);
    v_ = dalloc<double>(nc_); h2d(v_, v);
    iext_ = dalloc<double>(nc_); h2d(iext_, iext);
    cm_ = dalloc<double>(nc_); h2d(cm_, cm);
    b_ = dalloc<double>(nc_);
    dammy_ = dalloc<double>(nc_);
    row_ = dalloc<int>(nc_ + 1); h2d(row_, row);
    col_ = dalloc<int>(col.size()); h2d(col_, col);
    val_ = dalloc<double>(val.size()); h2d(val_, val);
    for (int k = 0; k < NVEC; k++) vec_[k] = dalloc<double>(nc_);
    red_a_ = dalloc<double>(nblk(nc_) + 1);
    red_b_ = dalloc<double>(nblk(nc_) + 1);
    bad_ = dalloc<int>(1);
    CB_HIP(hipMemset(bad_, 0, sizeof(int)));
    // coherent mapped host memory: the last reduction block writes the sum here
    CB_HIP(hipHostMalloc((void **)&hres_, sizeof(double) * 2, hipHostMallocMapped | hipHostMallocCoherent));
    CB_HIP(hipHostGetDevicePointer((void **)&hres_dev_, hres_, 0));
    amp_ = dalloc<double>(n_); h2d(amp_, stim_.amp);
    if (syn_.on) {
      syn_g_ = dalloc<double>(nc_);
      CB_HIP(hipMemset(syn_g_, 0, sizeof(double) * nc_));
    }
    hmax_ = H_MAX;
  }

  ~GlobalRkc() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)v_, (void *)iext_, (void *)cm_, (void *)b_, (void *)dammy_, (void *)row_, (void *)col_,
                    (void *)val_, (void *)red_a_, (void *)red_b_, (void *)bad_, (void *)amp_, (void *)syn_g_,
                    (void *)zero_, (void *)net_s_})
      if (p) (void)hipFree(p);
    for (int k = 0; k < NVEC; k++) (void)hipFree(vec_[k]);
    (void)hipHostFree(hres_);
    (void)hipStreamDestroy(st_);
  }

  int nc() const override { return nc_; }
  int ncell() const override { return n_; }
  Stats stats() const override { return st; }
  void controller(double *o, double *hw) const override {
    std::memcpy(o, o_, sizeof(o_));
    std::memcpy(hw, hw_, sizeof(hw_));
  }
  void sync() override { CB_HIP(hipStreamSynchronize(st_)); }

  void set_current(double t) override {
    k_set_current<<<(n_ + 127) / 128, 128, 0, st_>>>(iext_, amp_, comp_, stim_.comp, stim_.t0, stim_.t1, t, n_);
  }

  void advance() override {
    if (syn_.on) {
      double lam = syn_.rate_hz * H_MAX * 1e-3;
      double p = 1.0 - std::exp(-lam);
      k_poisson<<<nblk(nc_), 256, 0, st_>>>(syn_g_, std::exp(-H_MAX / syn_.tau_ms), p, syn_.g_event,
                                            (uint32_t)syn_.seed, (uint32_t)(syn_.seed >> 32),
                                            (uint32_t)st.macro_steps, nc_);
    }
    // the synaptic current of this macro step from its starting y (Golgi and
    // Purkinje cells: reset_b + add_*_val_new, solve_rkc.cu:1199-1200, 1267-1269)
    if (net_.on() && net_s_) syn_macro_current(net_, vec_[Y], net_s_, nc_, st_);
    if (sum_dt_ != 0.0) {
      ion_and_b(sum_dt_);
      sum_dt_ = 0.0;
    }
    while (true) {
      rkclow();
      sum_dt_ += o_[GT] - o_[GT_OLD];
      o_[GT_OLD] = o_[GT];
      if (trigger() && o_[IDID] == 2.0) {
        ion_and_b(sum_dt_);
        sum_dt_ = 0.0;
      }
      if (o_[IDID] == 2.0) continue;
      if (o_[IDID] == 1.0) break;
      throw Error("RKC failed: idid = " + std::to_string((int)o_[IDID]), (int)o_[IDID]);
    }
    // rkcint at gt_old (solve_rkc.cu:794-813)
    double hlast = hw_[0], lt = hw_[1], gtv = o_[GT_OLD];
    double tlast = lt - hlast;
    double s = (gtv - tlast) / hlast;
    double a1 = (1.0 + 2.0 * s) * (s - 1.0) * (s - 1.0);
    double a2 = (3.0 - 2.0 * s) * s * s;
    double b1 = hlast * s * (s - 1.0) * (s - 1.0);
    double b2 = hlast * (s - 1.0) * s * s;
    k_rkcint<<<nblk(nc_), 256, 0, st_>>>(vec_[VNEW], a1, a2, b1, b2, vec_[VT1], vec_[VT2], vec_[YN], vec_[FN], nc_);
    o_[TEND] += H_MAX;
    st.macro_steps++;
  }

  const double *device_v() const override { return v_; }
  const double *device_vnew() const override { return vec_[VNEW]; }
  hipStream_t stream() const override { return st_; }

  void read(int field, double *dst) override {
    const double *src = nullptr;
    switch (field) {
      case F_VNEW: src = vec_[VNEW]; break;
      case F_V: src = v_; break;
      case F_B: src = b_; break;
      case F_IEXT: src = iext_; break;
      case F_Y: src = vec_[Y]; break;
      case F_YN: src = vec_[YN]; break;
      case F_FN: src = vec_[FN]; break;
      default: src = model_field(field);
    }
    if (!src) throw Error("unknown field " + std::to_string(field));
    CB_HIP(hipMemcpyAsync(dst, src, sizeof(double) * nc_, hipMemcpyDeviceToHost, st_));
    CB_HIP(hipStreamSynchronize(st_));
  }

 protected:
  // b from vv; with do_ion, first the ion update over dt with vv as the new
  // voltage (the reference's xx_update_ion_and_matrix)
  virtual void model_b(const double *vv, bool do_ion, double dt) = 0;
  // accumulated time reached the ion-update interval: '>=' for GO/PKJ,
  // '>' for GR/IO (solve_rkc.cu:1148, 1212, 1284, 1347)
  virtual bool trigger() const { return sum_dt_ >= integ_; }
  virtual const double *model_field(int) const { return nullptr; }

  // Golgi / Purkinje coupling: keep the synapse view and a buffer for the
  // macro step's synaptic current
  void attach_net(const SynIn &in) {
    net_ = in;
    if (in.on() && !net_s_) {
      net_s_ = dalloc<double>(nc_);
      CB_HIP(hipMemset(net_s_, 0, sizeof(double) * nc_));
    }
  }

  // RKC_vec_initialize and the configuration of the controller; called by the
  // cell type's constructor once its state is on the device
  void setup(const SolverSpec &sol, double hmin) {
    hmin_ = hmin;
    // one controller for the whole population: the floor is the largest cell's
    for (double r : sol.rho_floor) rho_floor_ = std::max(rho_floor_, r);
    init(sol.atol, sol.rtol);
    if (sol.ion_interval > 0) integ_ = sol.ion_interval;
    refresh_fn_ = sol.refresh_fn;
  }

  Csr csr() const { return Csr{row_, col_, val_}; }

  void ion_and_b(double dt) {
    st.ion_updates++;
    model_b(vec_[Y], true, dt);
    if (refresh_fn_) k_stage_rhs(vec_[YN], vec_[FN]);
  }

  // the sum of nparts block partials in red_a_ (the deterministic tree), and
  // with flag the zero-weight flag of the same step, in one read-back
  double finish_sum(int nparts, int *flag = nullptr) {
    double *in = red_a_, *out = red_b_;
    while (nparts > RED_BLOCK) {
      int nb = nblk(nparts);
      k_reduce256<<<nb, 256, 0, st_>>>(in, out, nparts);
      std::swap(in, out);
      nparts = nb;
    }
    k_reduce_to_host<<<1, 256, 0, st_>>>(in, nparts, flag ? bad_ : nullptr, hres_dev_);
    CB_HIP(hipGetLastError());
    CB_HIP(hipStreamSynchronize(st_));
    if (flag) *flag = hres_[1] != 0.0;
    return hres_[0];
  }

  double sumsq(const double *x) {
    k_sumsq<<<nblk(nc_), 256, 0, st_>>>(x, red_a_, nc_);
    return finish_sum(nblk(nc_));
  }

  void init(double atol, double rtol) {
    // RKC_vec_initialize (solve_rkc.cu:756-791): b from the initial voltage
    model_b(v_, false, 0.0);
    // set_initial_vec (solve_rkc.cu:437-487)
    CB_HIP(hipMemcpyAsync(vec_[VNEW], v_, sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    CB_HIP(hipMemcpyAsync(vec_[Y], v_, sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    CB_HIP(hipMemcpyAsync(vec_[YN], v_, sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    for (int k : {FN, EDE, VT1, VT2}) k_fill<<<nblk(nc_), 256, 0, st_>>>(0.0, vec_[k], nc_);
    for (double &x : hw_) x = 0.0;
    for (double &x : o_) x = 0.0;
    hw_[2] = nc_;
    hw_[3] = UROUND;
    hw_[4] = std::sqrt(UROUND);
    hw_[5] = hmax_;
    hw_[6] = hmin_;
    o_[ATOL] = atol;
    o_[RTOL] = rtol;
    o_[TEND] = hmax_;
    last_ = true;
    newspc_ = true;
    jacatt_ = false;
    k_stage_rhs(vec_[YN], vec_[FN]);
    CB_HIP(hipMemcpyAsync(vec_[EDE], vec_[FN], sizeof(double) * nc_, hipMemcpyDeviceToDevice, st_));
    rkcrho();
    // set_initial_absh (solve_rkc.cu:489-560)
    double absh = hmax_;
    if (o_[SPRAD] * absh > 1.0) absh = 1.0 / o_[SPRAD];
    absh = fmax_ori(absh, hw_[6]);
    k_saxpb<<<nblk(nc_), 256, 0, st_>>>(absh, vec_[FN], vec_[YN], vec_[VT1], nc_);
    k_stage_rhs(vec_[VT1], vec_[VT2]);
    k_absh_w<<<nblk(nc_), 256, 0, st_>>>(vec_[VT2], vec_[FN], vec_[YN], o_[ATOL], o_[RTOL], dammy_, bad_, nc_);
    if (read_bad()) throw Error("RKC initial step: zero error weight (idid = 3)", 3);
    double est = sumsq(dammy_);
    est = absh * std::sqrt(est / nc_);
    if (0.1 * absh < hmax_ * std::sqrt(est))
      absh = fmax_ori(0.1 * absh / std::sqrt(est), hw_[6]);
    else
      absh = hmax_;
    o_[ABSH] = absh;
    o_[MMAX] = nint_ori(std::sqrt(o_[RTOL] / (10.0 * UROUND)));
    sum_dt_ = 0.0;
    integ_ = std::pow(2.0, std::log10(o_[ATOL]));   // solve_rkc.cu:1265
  }

  int read_bad() {
    int h = 0;
    CB_HIP(hipMemcpyAsync(&h, bad_, sizeof(int), hipMemcpyDeviceToHost, st_));
    CB_HIP(hipStreamSynchronize(st_));
    return h;
  }

  // dy = F(y) alone: a stage with mu = 0 is not the same arithmetic, so use
  // the power-iteration kernel's RHS with fn = 0 and discard the norm.
  void k_stage_rhs(const double *y, double *dy) {
    k_rho_iter<<<nblk(nc_), 256, 0, st_>>>(csr(), b_, cm_, y, zero_fn(), dy, dammy_, red_a_, nc_);
  }

  const double *zero_fn() {
    if (!zero_) {
      zero_ = dalloc<double>(nc_);
      CB_HIP(hipMemsetAsync(zero_, 0, sizeof(double) * nc_, st_));
    }
    return zero_;
  }

  // rkcrho, solve_rkc.cu:562-754
  // spectral radius estimate, raised to the floor when one is configured
  void rkcrho() {
    rkcrho_power();
    if (o_[SPRAD] < rho_floor_) o_[SPRAD] = rho_floor_;
  }

  void rkcrho_power() {
    const int itmax = 1000;
    double sprad_old[3] = {0.0, 0.0, 0.0};
    const double sqrtu = std::sqrt(UROUND);
    const double small = 1.0 / 0.125;
    double ynrm = std::sqrt(sumsq(vec_[YN]));
    double vnrm = std::sqrt(sumsq(vec_[EDE]));
    double dynrm;
    if (ynrm != 0.0 && vnrm != 0.0) {
      dynrm = ynrm * sqrtu;
      k_saxpb<<<nblk(nc_), 256, 0, st_>>>(dynrm / vnrm, vec_[EDE], vec_[YN], vec_[VT1], nc_);
    } else if (ynrm != 0.0) {
      dynrm = ynrm * sqrtu;
      k_saxpb<<<nblk(nc_), 256, 0, st_>>>(sqrtu, vec_[YN], vec_[YN], vec_[VT1], nc_);
    } else if (vnrm != 0.0) {
      dynrm = UROUND;
      k_scale<<<nblk(nc_), 256, 0, st_>>>(dynrm / vnrm, vec_[EDE], vec_[VT1], nc_);
    } else {
      dynrm = UROUND;
      k_fill<<<nblk(nc_), 256, 0, st_>>>(dynrm, vec_[VT1], nc_);
    }
    double sigma = 0.0, sigmal;
    for (int iter = 0; iter < itmax; iter++) {
      st.rho_iters++;
      k_rho_iter<<<nblk(nc_), 256, 0, st_>>>(csr(), b_, cm_, vec_[VT1], vec_[FN], vec_[VT2], dammy_, red_a_, nc_);
      double dfnrm = std::sqrt(finish_sum(nblk(nc_)));
      sigmal = sigma;
      sigma = dfnrm / dynrm;
      sprad_old[0] = sprad_old[1];
      sprad_old[1] = sprad_old[2];
      sprad_old[2] = o_[SPRAD];
      o_[SPRAD] = 1.2 * sigma;
      if (iter >= 1 && std::fabs(sigma - sigmal) <= fmax_ori(sigma, small) * 0.001) {
        k_saxpb<<<nblk(nc_), 256, 0, st_>>>(-1.0, vec_[YN], vec_[VT1], vec_[EDE], nc_);
        return;
      }
      if (iter > itmax / 2) {
        if ((std::fabs(sprad_old[2] - sprad_old[0]) < 0.00001) && (std::fabs(sprad_old[1] - o_[SPRAD]) < 0.00001)) {
          o_[SPRAD] = (o_[SPRAD] + sprad_old[2]) / 2.0;
          k_saxpb<<<nblk(nc_), 256, 0, st_>>>(-1.0, vec_[YN], vec_[VT1], vec_[EDE], nc_);
          return;
        }
      }
      if (dfnrm != 0.0) {
        k_rho_next<<<nblk(nc_), 256, 0, st_>>>(dynrm / dfnrm, vec_[VT2], vec_[FN], vec_[YN], vec_[VT1], nc_);
      } else {
        k_flip1<<<1, 1, 0, st_>>>(vec_[VT1], vec_[YN], iter % nc_);
      }
    }
    o_[IDID] = 6;
  }

  // step, solve_rkc.cu:815-906
  void step(double h, int m) {
    double ajm1, arg, bj, bjm1, bjm2, dzj, dzjm1, dzjm2, d2zj, d2zjm1, d2zjm2, mu, mus, nu, temp1, temp2, w0, w1, zj,
        zjm1, zjm2;
    w0 = 1.0 + 2.0 / (13.0 * m * m);
    temp1 = w0 * w0 - 1.0;
    temp2 = std::sqrt(temp1);
    arg = m * std::log(w0 + temp2);
    w1 = std::sinh(arg) * temp1 / (std::cosh(arg) * m * temp2 - w0 * std::sinh(arg));
    bjm1 = 1.0 / ((2.0 * w0) * (2.0 * w0));
    bjm2 = bjm1;
    mus = w1 * bjm1;
    k_first<<<nblk(nc_), 256, 0, st_>>>(vec_[FN], vec_[YN], vec_[VT1], vec_[VT2], h * mus, nc_);
    st.stages++;
    zjm1 = w0; zjm2 = 1.0; dzjm1 = 1.0; dzjm2 = 0.0; d2zjm1 = 0.0; d2zjm2 = 0.0;
    for (int j = 1; j < m; j++) {
      zj = 2.0 * w0 * zjm1 - zjm2;
      dzj = 2.0 * w0 * dzjm1 - dzjm2 + 2.0 * zjm1;
      d2zj = 2.0 * w0 * d2zjm1 - d2zjm2 + 4.0 * dzjm1;
      bj = d2zj / (dzj * dzj);
      ajm1 = 1.0 - zjm1 * bjm1;
      mu = 2.0 * w0 * bj / bjm1;
      nu = -bj / bjm2;
      mus = mu * w1 / w0;
      k_stage<<<nblk(nc_), 256, 0, st_>>>(csr(), b_, cm_, vec_[VT1], vec_[VT2], vec_[YN], vec_[FN], vec_[Y], mu, nu,
                                          (1.0 - mu - nu), h * mus, ajm1, nc_);
      st.stages++;
      double *tmp = vec_[VT2];
      vec_[VT2] = vec_[VT1];
      vec_[VT1] = vec_[Y];
      vec_[Y] = tmp;
      bjm2 = bjm1; bjm1 = bj;
      zjm2 = zjm1; zjm1 = zj;
      dzjm2 = dzjm1; dzjm1 = dzj;
      d2zjm2 = d2zjm1; d2zjm1 = d2zj;
    }
    double *tmp = vec_[Y];
    vec_[Y] = vec_[VT1];
    vec_[VT1] = vec_[VT2];
    vec_[VT2] = tmp;
  }

  // rkclow, solve_rkc.cu:908-1096
  void rkclow() {
    const double hmin = hw_[6];
    int loop_num = 0;
    while (true) {
      if (++loop_num > 1000) return;
      if (newspc_) {
        rkcrho();
        jacatt_ = true;
      }
      last_ = false;
      if (1.1 * o_[ABSH] >= std::fabs(o_[TEND] - o_[LT])) {
        o_[ABSH] = std::fabs(o_[TEND] - o_[LT]);
        last_ = true;
      }
      // (a non-finite h * rho takes the mmax branch instead of an undefined
      // conversion; its error estimate then stops the run below)
      const double mr = std::sqrt(1.54 * o_[ABSH] * o_[SPRAD] + 1.0);
      int m = (mr < o_[MMAX]) ? 1 + (int)mr : (int)o_[MMAX] + 1;
      if (m > o_[MMAX]) {
        m = (int)o_[MMAX];
        o_[ABSH] = (m * m - 1) / (1.54 * o_[SPRAD]);
        last_ = false;
      }
      double h = o_[ABSH];
      if (h <= hmin) h = hmin;
      step(h, m);
      const int nb = nblk(nc_);
      k_err<<<nb, 256, 0, st_>>>(csr(), b_, cm_, vec_[Y], vec_[YN], vec_[FN], vec_[VT1], o_[ATOL], o_[RTOL], h, red_a_,
                                 bad_, nc_);
      st.stages++;
      int bad = 0;
      double err = finish_sum(nb, &bad);
      if (bad) {
        o_[IDID] = 3;
        return;
      }
      err = std::sqrt(err / (double)nc_);
      // a NaN or infinite state: the reference's main stops on a NaN voltage
      // (main.cu:200, 228, 257, 287); here the run stops with the status
      if (!std::isfinite(err)) throw Error("RKC: non-finite state (error estimate " + std::to_string(err) + ")", NONFINITE);
      if (err > 1.0) {
        st.rejects++;
        o_[ABSH] = 0.8 * o_[ABSH] / (std::pow(err, 1.0 / 3.0));
        if (o_[ABSH] < hmin) {
          o_[IDID] = 4;   // and fall through to the accept branch, as the reference does
        } else {
          newspc_ = !jacatt_;
          continue;
        }
      }
      st.steps++;
      o_[NACCPT] += 1;
      o_[LT] += h;
      o_[GT] = o_[LT];
      jacatt_ = false;
      int l_nstsig = ((int)(o_[NSTSIG]) + 1) % 25;   // nstsig is never advanced
      newspc_ = false;
      if (l_nstsig == 0) newspc_ = !jacatt_;
      hw_[0] = h;
      hw_[1] = o_[LT];
      double *tmp = vec_[VT2];
      vec_[VT2] = vec_[FN];
      vec_[FN] = vec_[VT1];
      vec_[VT1] = vec_[YN];
      vec_[YN] = vec_[Y];
      vec_[Y] = tmp;
      double fac = 10.0, temp1, temp2;
      if (o_[NACCPT] == 1) {
        temp2 = std::pow(err, 1.0 / 3.0);
        if (0.8 < fac * temp2) fac = 0.8 / temp2;
      } else {
        temp1 = 0.8 * o_[ABSH] * std::pow(o_[ERROLD], 1.0 / 3.0);
        temp2 = std::fabs(o_[HOLD]) * std::pow(err, 2.0 / 3.0);
        if (temp1 < fac * temp2) fac = temp1 / temp2;
      }
      o_[ABSH] = fmax_ori(0.1, fac) * o_[ABSH];
      o_[ABSH] = fmax_ori(hmin, std::fmin(hmax_, o_[ABSH]));
      o_[ERROLD] = err;
      o_[HOLD] = h;
      o_[IDID] = last_ ? 1 : 2;
      return;
    }
  }

  int n_, nc_, comp_;
  StimSpec stim_;
  PoissonSpec syn_;
  hipStream_t st_;
  double *v_, *iext_, *cm_, *b_, *dammy_, *val_;
  int *row_, *col_;
  double *vec_[NVEC];
  double *red_a_, *red_b_, *hres_, *amp_;
  double *hres_dev_ = nullptr;   // hres_ as the device sees it
  double *syn_g_ = nullptr;
  SynIn net_;               // incoming synapses (none: net_.on() false)
  double *net_s_ = nullptr; // Golgi / Purkinje: the macro step's synaptic current
  double *zero_ = nullptr;
  int *bad_;
  double hw_[8], o_[NOTHERS];
  bool last_, newspc_, jacatt_, refresh_fn_ = false;
  double rho_floor_ = 0.0;
  double sum_dt_ = 0.0, integ_ = 0.0, hmin_, hmax_;
  Stats st;
};

// Purkinje cells (pkj.cu, pkj_solve.cu, pkj_ion.cu)
class GlobalPkj : public GlobalRkc {
 public:
  GlobalPkj(const PkjHost &H, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn)
      : GlobalRkc(H.n, H.nc, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, with_amp(stim, H.n), syn) {
    using namespace pkj;
    ca_ = dalloc<double>(nc_); h2d(ca_, H.ca);
    rev_ = dalloc<double>(nc_); h2d(rev_, H.rev);
    shell_ = dalloc<double>(nc_); h2d(shell_, H.shell);
    cond_ = dalloc<double>((size_t)NCOND * nc_); h2d(cond_, H.cond);
    gate_ = dalloc<double>((size_t)NGATE * nc_); h2d(gate_, H.gate);
    setup(sol, sol.hmin_or(std::pow(2.0, -10.0)));   // H_MIN_PKJ, solve_rkc.cuh:26
  }
  ~GlobalPkj() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)ca_, (void *)rev_, (void *)shell_, (void *)cond_, (void *)gate_}) (void)hipFree(p);
  }

 private:
  static StimSpec with_amp(StimSpec s, int n) {
    if (s.amp.size() != (size_t)n) s.amp.assign(n, pkj::CURRENT);
    return s;
  }
  void model_b(const double *vv, bool do_ion, double dt) override {
    // the granule input enters at ion updates only (pkj_update_ion_and_matrix)
    const bool net = do_ion && net_.on();
    PkjState S{v_, ca_, rev_, iext_, shell_, cond_, gate_, b_, syn_g_, syn_.e_syn, nc_,
               net ? net_s_ : nullptr, net ? std::exp(-dt / 0.6) : 0.0};
    k_pkj_ion_b<<<nblk(nc_), 256, 0, st_>>>(S, vv, dt, do_ion ? 1 : 0);
  }
 public:
  void set_synapses(const SynIn &in) override { attach_net(in); }

 private:
  const double *model_field(int field) const override {
    switch (field) {
      case F_CA: return ca_;
      case F_REV: return rev_;
    }
    if (field >= F_GATE0 && field < F_GATE0 + pkj::NGATE) return gate_ + (size_t)(field - F_GATE0) * nc_;
    return nullptr;
  }
  double *ca_, *rev_, *shell_, *cond_, *gate_;
};

// Inferior-olive cells with their grid gap junctions (io.cu, io_solve.cu,
// io_ion.cu, gap.cu)
class GlobalIo : public GlobalRkc {
 public:
  GlobalIo(const IoHost &H, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn)
      : GlobalRkc(H.n, H.nc, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, with_amp(stim, H.n), syn) {
    using namespace io;
    ca_ = dalloc<double>(nc_); h2d(ca_, H.ca);
    cond_ = dalloc<double>((size_t)NCOND * nc_); h2d(cond_, H.cond);
    gate_ = dalloc<double>((size_t)NGATE * nc_); h2d(gate_, H.gate);
    compart_ = dalloc<int8_t>(nc_); h2d(compart_, H.compart);
    nbr_ = dalloc<int>(H.nbr.size()); h2d(nbr_, H.nbr);
    setup(sol, sol.hmin_or(std::pow(2.0, -10.0)));   // H_MIN_IO, solve_rkc.cuh:27
  }
  ~GlobalIo() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)ca_, (void *)cond_, (void *)gate_, (void *)compart_, (void *)nbr_}) (void)hipFree(p);
  }

 private:
  static StimSpec with_amp(StimSpec s, int n) {
    if (s.amp.size() != (size_t)n) s.amp.assign(n, io::STIM);
    return s;
  }
  bool trigger() const override { return sum_dt_ > integ_; }   // strict for IO, solve_rkc.cu:1347
  void model_b(const double *vv, bool do_ion, double dt) override {
    IoState S{v_, ca_, iext_, cond_, gate_, compart_, nbr_, b_, syn_g_, syn_.e_syn, nc_, n_};
    k_io_ion_b<<<nblk(nc_), 256, 0, st_>>>(S, vv, dt, do_ion ? 1 : 0);
  }
  const double *model_field(int field) const override {
    if (field == F_CA) return ca_;
    if (field >= F_GATE0 && field < F_GATE0 + io::NGATE) return gate_ + (size_t)(field - F_GATE0) * nc_;
    return nullptr;
  }
  double *ca_, *cond_, *gate_;
  int8_t *compart_;
  int *nbr_;
};

// Golgi cells on a grid, gap junctions in the cable matrix (go.cu,
// go_solve.cu, go_ion.cu)
class GlobalGo : public GlobalRkc {
 public:
  GlobalGo(const GoHost &H, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn)
      : GlobalRkc(H.n, H.nc, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, with_amp(stim, H.n), syn) {
    using namespace go;
    ca_ = dalloc<double>(nc_); h2d(ca_, H.ca);
    ca2_ = dalloc<double>(nc_); h2d(ca2_, H.ca2);
    rev_ = dalloc<double>(nc_); h2d(rev_, H.rev);
    ca_old_ = dalloc<double>(nc_); h2d(ca_old_, H.ca_old);
    area_ = dalloc<double>(nc_); h2d(area_, H.area);
    cond_ = dalloc<double>((size_t)NCOND * nc_); h2d(cond_, H.cond);
    gate_ = dalloc<double>((size_t)NGATE * nc_); h2d(gate_, H.gate);
    setup(sol, sol.hmin_or(std::pow(2.0, -10.0)));   // H_MIN_GO, solve_rkc.cuh:25
  }
  ~GlobalGo() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)ca_, (void *)ca2_, (void *)rev_, (void *)ca_old_, (void *)area_, (void *)cond_,
                    (void *)gate_})
      (void)hipFree(p);
  }

 private:
  static StimSpec with_amp(StimSpec s, int n) {
    if (s.amp.size() != (size_t)n) s.amp.assign(n, go::STIM);
    return s;
  }
  void model_b(const double *vv, bool do_ion, double dt) override {
    const bool net = do_ion && net_.on();
    GoState S{v_, ca_, ca2_, rev_, ca_old_, iext_, area_, cond_, gate_, b_, syn_g_, syn_.e_syn, nc_,
              net ? net_s_ : nullptr};
    k_go_ion_b<<<nblk(nc_), 256, 0, st_>>>(S, vv, dt, do_ion ? 1 : 0);
  }
 public:
  void set_synapses(const SynIn &in) override { attach_net(in); }

 private:
  const double *model_field(int field) const override {
    switch (field) {
      case F_CA: return ca_;
      case F_REV: return rev_;
      case F_CA2: return ca2_;
      case F_CA_OLD: return ca_old_;
    }
    if (field >= F_GATE0 && field < F_GATE0 + go::NGATE) return gate_ + (size_t)(field - F_GATE0) * nc_;
    return nullptr;
  }
  double *ca_, *ca2_, *rev_, *ca_old_, *area_, *cond_, *gate_;
};

// Granule cells (gr.cu, gr_solve.cu, gr_ion.cu)
class GlobalGr : public GlobalRkc {
 public:
  GlobalGr(const GrHost &H, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn)
      : GlobalRkc(H.n, H.nc, H.comp, H.cm, H.v, H.iext, H.row, H.col, H.val, with_amp(stim, H.n), syn) {
    using namespace gr;
    ca_ = dalloc<double>(nc_); h2d(ca_, H.ca);
    compart_ = dalloc<int8_t>(nc_); h2d(compart_, H.compart);
    cond_ = dalloc<double>((size_t)NCOND * nc_); h2d(cond_, H.cond);
    gate_ = dalloc<double>((size_t)NGATE * nc_); h2d(gate_, H.gate);
    const NaTable T = na_table();
    na_ = dalloc<NaType>(NTYPE);
    CB_HIP(hipMemcpy(na_, T.t, sizeof(T.t), hipMemcpyHostToDevice));
    setup(sol, sol.hmin_or(std::pow(2.0, -12.0)));   // H_MIN_GR, solve_rkc.cuh:24
  }
  ~GlobalGr() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)ca_, (void *)compart_, (void *)cond_, (void *)gate_, (void *)na_}) (void)hipFree(p);
  }

 private:
  static StimSpec with_amp(StimSpec s, int n) {
    if (s.amp.size() != (size_t)n) s.amp.assign(n, gr::STIM);
    return s;
  }
  bool trigger() const override { return sum_dt_ > integ_; }   // strict for GR, solve_rkc.cu:1148
  void model_b(const double *vv, bool do_ion, double dt) override {
    GrState S{v_, iext_, ca_, compart_, cond_, gate_, na_, b_, syn_g_, syn_.e_syn, nc_,
              do_ion ? net_ : SynIn()};
    k_gr_ion_b<<<nblk(nc_), 256, 0, st_>>>(S, vv, dt, do_ion ? 1 : 0);
  }
 public:
  void set_synapses(const SynIn &in) override { net_ = in; }

 private:
  const double *model_field(int field) const override {
    if (field == F_CA) return ca_;
    if (field >= F_GATE0 && field < F_GATE0 + gr::NGATE) return gate_ + (size_t)(field - F_GATE0) * nc_;
    return nullptr;
  }
  double *ca_, *cond_, *gate_;
  int8_t *compart_;
  gr::NaType *na_;
};

}  // namespace

void create_stream(hipStream_t *st) {
  static char park[256];
  char *caller = initstate(1u, park, sizeof(park));
  const hipError_t e = hipStreamCreate(st);
  setstate(caller);
  CB_HIP(e);
}

void Engine::set_raster(int comp, double thr) {
  if (comp >= 0 && comp >= nc() / std::max(1, ncell())) throw Error("raster compartment out of range");
  raster_comp_ = comp;
  raster_thr_ = thr;
  if (raster_prev_) {   // a new raster starts from the reference's 100 mV
    std::vector<double> h(ncell(), 100.0);
    CB_HIP(hipMemcpy(raster_prev_, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  }
}

double *Engine::raster_prev() {
  if (!raster_prev_) {
    CB_HIP(hipMalloc(&raster_prev_, sizeof(double) * std::max(1, ncell())));
    std::vector<double> h(ncell(), 100.0);
    CB_HIP(hipMemcpy(raster_prev_, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  }
  return raster_prev_;
}

void Engine::free_raster() {
  if (raster_prev_) (void)hipFree(raster_prev_);
  raster_prev_ = nullptr;
}

int Engine::take_spikes(int cap, double *t_out, int *cell_out) {
  const int k = (int)std::min<size_t>(spikes_.size(), (size_t)std::max(0, cap));
  for (int i = 0; i < k; i++) {
    t_out[i] = spikes_[i].first;
    cell_out[i] = spikes_[i].second;
  }
  spikes_.erase(spikes_.begin(), spikes_.begin() + k);
  return k;
}

void Engine::append_spikes(const std::vector<uint32_t> &bits, int nsteps, int words, double t0, double interval) {
  const int n = ncell();
  for (int k = 0; k < nsteps; k++)
    for (int c = 0; c < n; c++)
      if (bits[(size_t)k * words + (c >> 5)] >> (c & 31) & 1u) spikes_.push_back({t0 + (k + 1) * interval, c});
}

void Engine::record_interval(hipStream_t st, const double *v, double *rec, int rec_comp, uint32_t *bits) {
  const int n = ncell();
  k_rec_step<<<(unsigned)((n + 127) / 128), 128, 0, st>>>(v, nc() / std::max(1, n), n, rec_comp, rec, raster_comp_,
                                                          raster_thr_, bits ? raster_prev() : nullptr, bits);
  CB_HIP(hipGetLastError());
}

// The macro-step loop of main.cu:182-311 for one population: set_current and
// advance per output interval, with the recorded compartment (rec_comp >= 0)
// or every compartment (rec_comp < 0) and the raster bits kept on the device
// and copied back once.
void Engine::run(int nsteps, double t0, int rec_comp, double *rec_out) {
  if (nsteps <= 0) return;
  const double *v = device_vnew();
  const int n = ncell(), per = rec_comp >= 0 ? 1 : nc() / std::max(1, n);
  const bool raster = raster_comp_ >= 0;
  if (!v) {   // an engine without a device output array: copy V per interval
    if (raster) throw Error("device raster not provided by this engine");
    std::vector<double> buf(nc());
    for (int k = 0; k < nsteps; k++) {
      set_current(t0 + k * H_MAX);
      advance();
      if (rec_out) {
        read(F_VNEW, buf.data());
        for (int c = 0; c < n; c++)
          for (int q = 0; q < per; q++)
            rec_out[((size_t)k * n + c) * per + q] = buf[(size_t)c * (nc() / n) + (rec_comp >= 0 ? rec_comp : q)];
      }
    }
    return;
  }
  const int words = (n + 31) / 32;
  double *drec = nullptr;
  uint32_t *dbits = nullptr;
  if (rec_out) CB_HIP(hipMalloc(&drec, sizeof(double) * (size_t)nsteps * n * per));
  if (raster) {
    CB_HIP(hipMalloc(&dbits, sizeof(uint32_t) * (size_t)nsteps * words));
    CB_HIP(hipMemsetAsync(dbits, 0, sizeof(uint32_t) * (size_t)nsteps * words, stream()));
  }
  try {
    for (int k = 0; k < nsteps; k++) {
      set_current(t0 + k * H_MAX);
      advance();
      if (drec || dbits)
        record_interval(stream(), device_vnew(), drec ? drec + (size_t)k * n * per : nullptr, rec_comp,
                        dbits ? dbits + (size_t)k * words : nullptr);
    }
    std::vector<uint32_t> bits((size_t)(raster ? nsteps : 0) * words);
    if (drec)
      CB_HIP(hipMemcpyAsync(rec_out, drec, sizeof(double) * (size_t)nsteps * n * per, hipMemcpyDeviceToHost, stream()));
    if (dbits)
      CB_HIP(hipMemcpyAsync(bits.data(), dbits, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, stream()));
    CB_HIP(hipStreamSynchronize(stream()));
    if (raster) append_spikes(bits, nsteps, words, t0, H_MAX);
  } catch (...) {
    if (drec) (void)hipFree(drec);
    if (dbits) (void)hipFree(dbits);
    throw;
  }
  if (drec) CB_HIP(hipFree(drec));
  if (dbits) CB_HIP(hipFree(dbits));
}

Engine *make_global_pkj(const PkjHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn) {
  return new GlobalPkj(host, sol, stim, syn);
}

Engine *make_global_io(const IoHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn) {
  return new GlobalIo(host, sol, stim, syn);
}

Engine *make_global_go(const GoHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn) {
  return new GlobalGo(host, sol, stim, syn);
}

Engine *make_global_gr(const GrHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn) {
  return new GlobalGr(host, sol, stim, syn);
}

}  // namespace cbrkc
