// Per-cell adaptive RKC in one persistent launch: one workgroup advances one
// Purkinje cell through many macro steps with its own step-size controller.
//
// Each cell is the reference's population of one cell (solve_rkc.cu run on
// a single neuron): the same Chebyshev recurrence (step, :815-906), the same
// error estimate and step control (rkclow, :908-1096), the same nonlinear
// power method (rkcrho, :562-754), the same operator-split ion update
// cadence (pkj_solve_by_rkc, :1252-1299) and quirks (Y_{m-2} as the voltage
// handed to the ion update; spectral radius refreshed only after a reject;
// steps at hmin always accepted).  What changes is the mapping:
//
//   * compartments are renumbered breadth-first from the soma, so the
//     children of a compartment are contiguous and the cable stencil is
//     parent + child range; thread t owns compartments t, t+NT, ...
//     (coalesced when the gate planes are streamed);
//   * the stage recurrence runs on increments D_j = Y_j - Y_0 kept in
//     registers, with D_{j-1} staged in LDS (double buffered, one barrier
//     per stage) for the neighbour gather; nothing touches HBM inside the
//     stage loop;
//   * y_n is a compensated pair (hi + lo) updated with TwoSum, so the
//     mixed-precision engine keeps ~48 bits of voltage while every stage,
//     gate and current is FP32;
//   * every norm is a fixed-order workgroup reduction in FP64, and the
//     controller arithmetic is FP64;
//   * gates, Ca and the reversal potential stay in HBM and are streamed
//     only at ion updates; Ca and its quantisation stay FP64.
//
// T = double gives the per-cell FP64 engine; T = float the mixed engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <mutex>
#include <type_traits>
#include <vector>

#include "engine.hpp"
#include "philox.hpp"
#include "ion_state.hpp"
#include "morph.hpp"
#include "pkj_model.hpp"

namespace cbrkc {
namespace {

constexpr double UROUND = 2.22e-16;
constexpr int NO_PARENT = 0x7FF;
// per-compartment flag byte: compartment type in the low bits, plus the
// reference's index-keyed fixed somatic Ca reversal (pkj_ion.cu:329-331)
constexpr int CLS_MASK = 0x7;
constexpr int FIXED_REV = 0x10;

// per-cell controller record (doubles)
enum Ctl {
  C_ABSH, C_SPRAD, C_NACCPT, C_ERROLD, C_HOLD, C_LT, C_TEND, C_GT, C_GTOLD, C_SUMDT, C_MMAX, C_HLAST, C_LTLAST,
  C_NEWSPC, C_JACATT, C_LAST, C_IDID, C_STEPS, C_REJECTS, C_STAGES, C_RHO, C_IONS, C_MACRO, C_STATUS,
  C_PHASE0 = 24, NCTL = 32
};

// Optional per-phase cycle counters, built with -DCB_PHASE_CLOCKS: thread 0
// of each workgroup accumulates clock64() deltas of its own wave into
// C_PHASE0 + {A rebuild, ion update, power method, step set-up, stage loop,
// error + controller}.  The default build compiles them out.
enum Phase { PH_A, PH_ION, PH_RHO, PH_SETUP, PH_STAGES, PH_ERR };
#ifdef CB_PHASE_CLOCKS
#define PHASE_T0(v) const unsigned long long v = clock64()
#define PHASE_ADD(q, v) \
  if (threadIdx.x == 0) s_ph[q] += (double)(clock64() - (v))
#else
#define PHASE_T0(v) (void)0
#define PHASE_ADD(q, v) (void)0
#endif

// Per-compartment planes of one cell.  All of a cell's per-compartment
// arrays sit in one slab: plane p of cell c starts at byte
//   c * slab_stride + p * PITCH * sizeof(T)
// in breadth-first compartment order, so a workgroup reaches every array of
// its cell through one buffer resource held in SGPRs plus one 32-bit lane
// offset, instead of one 64-bit vector address per array.  The pitch is a
// compile-time constant (the largest cell the 256-thread mapping holds), so
// every plane offset is a literal rather than a live register.
constexpr int PITCH = 1792;

// stage counts up to this read their coefficients from the table (64 x 64
// entries of 4 doubles, 128 KiB); larger ones run the recurrence in-kernel
constexpr int COEF_MMAX = 64;
enum Plane {
  P_GATE = 0,                                    // NGATE gate planes
  P_AREA = pkj::NGATE, P_LEAK, P_SHELL,          // static
  P_META,                                        // uint32: orig | flags << 16
  P_REV, P_VOLD, P_B, P_SYN,                     // ion and current state
  P_Y0HI, P_Y0LO, P_F0, P_VY, P_EDE,             // integrator state at launch edges
  P_CA                                           // FP64 Ca, 8 / sizeof(T) planes
};
// the network synaptic current of the macro step follows the Ca planes
template <typename T> constexpr int p_nets() { return P_CA + (int)(8 / sizeof(T)); }
template <typename T> constexpr int n_planes() { return p_nets<T>() + 1; }

template <typename T>
struct CellArgs {
  int ncell, comp;
  double atol, rtol, hmin, hmax, integ;
  int refresh_fn;         // 1: fn = F(y_n) after each ion update
  const double *rho_floor;  // per cell lower bound of the spectral radius (null: none)
  const double *coef;     // RKC stage coefficients by stage count (rkc_stage_table)
  int coef_mmax;          // largest tabulated stage count
  double interval;        // output / exchange interval [ms] (the reference's H_MAX = 0.125)
  // static, per compartment, cell-major, breadth-first order (read into LDS)
  const uint32_t *topo;   // parent | first_child << 11 | nchild << 22
  const T *gpar;          // conductance to the parent [mS]
  const T *invc;          // 1 / Cm
  unsigned char *slab;    // per-cell compartment planes
  uint32_t slab_stride;   // bytes per cell
  double *ctl;            // [ncell][NCTL]
  double *vnew_out;       // [ncell * comp] original order, written at launch end
  // stimulus
  int stim_orig;          // original local index of the stimulated compartment
  double stim_t0, stim_t1;
  const double *amp;      // per cell
  // Poisson drive
  int syn_on;
  T syn_decay, syn_g_event, syn_e;
  double syn_p;
  uint32_t syn_k0, syn_k1;
  // recording
  int rec_orig;           // original local index of the recorded compartment
  double *rec;            // [nsteps][ncell] or null
  // spikes: upward crossings of thr by the recorded compartment between two
  // intervals set bit (cell & 31) of spk[s * spk_words + cell / 32], s the
  // interval of the launch;
  // vprev holds each cell's previous value (null: no detection)
  double thr;
  double *vprev;
  uint32_t *spk;
  int spk_words;
  int *queue;             // work counter
  const int *order;       // queue position -> cell (the costliest cells of the last run first)
  // network synapses onto the cells (granule -> Purkinje): CSR by global
  // compartment cell * comp + original index; off when net.row is null
  SynIn net;
};

template <typename T> __device__ inline T absT(T x) { return x < T(0) ? -x : x; }

// x^(e/3), e = 1 or 2, for the step-size factor (solve_rkc.cu:1006-1066).
// The FP64 engine calls pow as the reference does.  The mixed engine takes a
// hardware FP32 cube-root seed (relative error ~1e-6) through two Newton
// steps in FP64 (~1e-12, then rounding level): three FP64 pow calls per step
// in every wave were a fifth of the controller's time.  Arguments outside
// (1e-30, 1e30), where the FP32 seed would under- or overflow, use pow.
template <typename T>
__device__ inline double ctl_pow(double x, int e) {
  if constexpr (sizeof(T) == 8) {
    return pow(x, e == 1 ? 1.0 / 3.0 : 2.0 / 3.0);
  } else {
    if (!(x > 1e-30 && x < 1e30)) return pow(x, e == 1 ? 1.0 / 3.0 : 2.0 / 3.0);
    double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (1.0f / 3.0f));
    y = (2.0 * y + x / (y * y)) * (1.0 / 3.0);
    y = (2.0 * y + x / (y * y)) * (1.0 / 3.0);
    return e == 1 ? y : y * y;
  }
}

// gfx9 buffer resource word 3 (32-bit data format, raw addressing)
constexpr int RSRC_WORD3 = 0x00020000;

// One cell's slab as a buffer resource: plane offsets go in the scalar
// offset, the lane's compartment index in the vector offset.
template <typename T>
struct Slab {
  static constexpr int pl = PITCH * (int)sizeof(T);   // bytes per plane
  __amdgpu_buffer_rsrc_t r;
  __device__ Slab(unsigned char *base, int bytes) : r(__builtin_amdgcn_make_buffer_rsrc(base, 0, bytes, RSRC_WORD3)) {}
  __device__ T ld(int p, int i) const {
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, p * pl, 0));
    else
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, p * pl, 0));
  }
  __device__ void st(int p, int i, T v) const {
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, i * 4, p * pl, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                            r, i * 8, p * pl, 0);
  }
  __device__ uint32_t meta(int i) const { return __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, P_META * pl, 0); }
  __device__ double ca(int i) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, P_CA * pl, 0));
  }
  __device__ void set_ca(int i, double v) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                          r, i * 8, P_CA * pl, 0);
  }
};

// The workgroup-wide values (reduction results, the cell index) are equal in
// every lane; saying so lets the compiler keep the controller in SGPRs.
__device__ inline double uniform(double x) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// fixed-order workgroup sum in FP64: xor butterfly inside each wave, then
// the wave totals in index order.  Returns the same value in every thread.
template <int NT>
__device__ inline double wg_sum(double x, double *red, int &slot) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
  const int w = threadIdx.x >> 6;
  double *r = red + slot * (NT / 64);
  slot ^= 1;
  if ((threadIdx.x & 63) == 0) r[w] = x;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; i++) s += r[i];
  return uniform(s);
}

template <typename T>
struct Lds {
  T *d0, *d1;        // stage buffers
  T *gp;             // gpar (FP64 engine)
  T *ic;             // invc
  T *bl;             // frozen current b (a copy of the slab plane, read by every A rebuild)
  uint32_t *tp;      // topology (FP64 engine)
  float4 *w4;        // FP32 engine: the four edge conductances (parent, children 0..2; 0 where none)
  uint2 *nb;         // FP32 engine: the four neighbour indices, 16 bits each (the compartment itself where none)
};

// sum_nb g (D_nb - D_i) with D in LDS buffer cur, where cur[i] holds this
// thread's own D_i.  Branch-free: a missing parent or child is read as the
// compartment itself (D_i - D_i = 0 exactly) with conductance 0, so every
// LDS read is issued unconditionally and together; the added terms are
// exact zeros, so the sum equals the one over the real neighbours.  The
// FP32 engine keeps the neighbour indices and edge conductances decoded in
// LDS (set once per cell), so a stage reads two words and the four values
// instead of decoding the topology word; the operands and their order are
// the same.
template <typename T>
__device__ inline T lap(const Lds<T> &L, const T *cur, int i, T di) {
  if constexpr (sizeof(T) == 4) {
    const uint2 q = L.nb[i];
    const float4 w = L.w4[i];
    T s = w.x * (cur[q.x & 0xFFFFu] - di);
    s += w.y * (cur[q.x >> 16] - di);
    s += w.z * (cur[q.y & 0xFFFFu] - di);
    s += w.w * (cur[q.y >> 16] - di);
    return s;
  }
  const uint32_t tp = L.tp[i];
  const int p = tp & 0x7FF, fc = (tp >> 11) & 0x7FF, nch = (tp >> 22) & 3;
  const int pp = (p != NO_PARENT) ? p : i;
  const T gi = L.gp[i];
  const T dp = cur[pp];
  T gc[3], dc[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {   // at most 3 children
    const int j = (c < nch) ? fc + c : i;
    gc[c] = L.gp[j];
    dc[c] = cur[j];
  }
  T s = ((p != NO_PARENT) ? gi : T(0)) * (dp - di);
#pragma unroll
  for (int c = 0; c < 3; c++) s += ((c < nch) ? gc[c] : T(0)) * (dc[c] - di);
  return s;
}

template <typename T>
__device__ inline void two_sum_acc(T &hi, T &lo, T d) {
  if constexpr (sizeof(T) == 8) {
    hi = hi + d;     // FP64 engine: plain accumulation, as the reference
    (void)lo;
  } else {
    T s = hi + d;
    T bp = s - hi;
    T e = (hi - (s - bp)) + (d - bp);
    T l = lo + e;
    hi = s + l;
    lo = l - (hi - s);
  }
}

// The body of both kernels: INIT runs RKC_vec_initialize for every cell,
// otherwise nsteps macro steps are advanced from t0.
template <typename T, int NT, int K, bool INIT, bool NET>
__device__ __forceinline__ void cell_body(const CellArgs<T> &a, int nsteps, double t0, int step0) {
  using namespace pkj;
  extern __shared__ __align__(16) unsigned char smem[];
  const int C = a.comp;
  Lds<T> L;
  // LDS layout (cell_lds_bytes gives the size):
  //   FP32: d0 d1 ic bl | w4 (16-byte aligned) | nb | red (16-byte aligned)
  //   FP64: d0 d1 gp ic bl tp | red (16-byte aligned)
  double *red;
  L.d0 = (T *)smem;
  L.d1 = L.d0 + C;
  if constexpr (sizeof(T) == 4) {
    L.gp = nullptr;
    L.tp = nullptr;
    L.ic = L.d1 + C;
    L.bl = L.ic + C;
    L.w4 = (float4 *)(((uintptr_t)(L.bl + C) + 15) & ~(uintptr_t)15);
    L.nb = (uint2 *)(L.w4 + C);
    red = (double *)(((uintptr_t)(L.nb + C) + 15) & ~(uintptr_t)15);
  } else {
    L.gp = L.d1 + C;
    L.ic = L.gp + C;
    L.bl = L.ic + C;
    L.tp = (uint32_t *)(L.bl + C);
    L.w4 = nullptr;
    L.nb = nullptr;
    red = (double *)(((uintptr_t)(L.tp + C) + 15) & ~(uintptr_t)15);
  }
  __shared__ int s_cell;
  __shared__ double s_cnt[6];   // steps, rejects, stages, rho iterations, ion updates, macro steps
  __shared__ double s_cs[NT / 64][14];   // per-wave copy of the controller state
#ifdef CB_PHASE_CLOCKS
  __shared__ double s_ph[8];    // cycles per phase, kept by thread 0
#endif
  const int tid = threadIdx.x;
  int slot = 0;

  while (true) {
    if (tid == 0) {
      const int q = atomicAdd(a.queue, 1);
      s_cell = (q < a.ncell) ? a.order[q] : q;
    }
    __syncthreads();
    const int cell = __builtin_amdgcn_readfirstlane(s_cell);
    __syncthreads();
    if (cell >= a.ncell) break;
    const size_t base = (size_t)cell * C;
    const Slab<T> S(a.slab + (size_t)cell * a.slab_stride, (int)a.slab_stride);
    double *ctl = a.ctl + (size_t)cell * NCTL;

    for (int i = tid; i < C; i += NT) {
      L.ic[i] = a.invc[base + i];
      L.bl[i] = S.ld(P_B, i);
      const uint32_t tp = a.topo[base + i];
      if constexpr (sizeof(T) == 4) {
        const int p = tp & 0x7FF, fc = (tp >> 11) & 0x7FF, nch = (tp >> 22) & 3;
        const int pp = (p != NO_PARENT) ? p : i;
        const int j0 = (0 < nch) ? fc : i, j1 = (1 < nch) ? fc + 1 : i, j2 = (2 < nch) ? fc + 2 : i;
        L.w4[i] = make_float4((p != NO_PARENT) ? a.gpar[base + i] : 0.0f, (0 < nch) ? a.gpar[base + j0] : 0.0f,
                              (1 < nch) ? a.gpar[base + j1] : 0.0f, (2 < nch) ? a.gpar[base + j2] : 0.0f);
        L.nb[i] = make_uint2((uint32_t)pp | ((uint32_t)j0 << 16), (uint32_t)j1 | ((uint32_t)j2 << 16));
      } else {
        L.gp[i] = a.gpar[base + i];
        L.tp[i] = tp;
      }
    }

    // register-resident per-compartment state
    T hi[K], lo[K], F0[K], A[K], D1[K], D2[K], VY[K];
    uint32_t rec_mask = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = tid + k * NT;
      if (i < C) {
        if ((int)(S.meta(i) & 0xFFFFu) == a.rec_orig) rec_mask |= 1u << k;
        hi[k] = S.ld(P_Y0HI, i);
        lo[k] = S.ld(P_Y0LO, i);
        F0[k] = S.ld(P_F0, i);
        VY[k] = S.ld(P_VY, i);
      } else {
        hi[k] = lo[k] = F0[k] = VY[k] = T(0);
      }
      A[k] = D1[k] = D2[k] = T(0);
    }
    __syncthreads();

    // ---- controller state (uniform across the workgroup) ----
    // Every wave runs the controller on its own copy, held in LDS rather
    // than in registers: FP64 values have no scalar registers, and kept as
    // 14 VGPR pairs across the stage loop they pushed the run kernel into
    // scratch spills that were reloaded at every step.
    double *cs = s_cs[tid >> 6];
    double &absh = cs[0], &sprad = cs[1], &naccpt = cs[2], &errold = cs[3], &hold = cs[4], &lt = cs[5],
           &tend = cs[6], &gt = cs[7], &gt_old = cs[8], &sum_dt = cs[9], &mmax = cs[10], &hlast = cs[11],
           &ltlast = cs[12], &idid = cs[13];
    absh = ctl[C_ABSH]; sprad = ctl[C_SPRAD]; naccpt = ctl[C_NACCPT]; errold = ctl[C_ERROLD];
    hold = ctl[C_HOLD]; lt = ctl[C_LT]; tend = ctl[C_TEND]; gt = ctl[C_GT]; gt_old = ctl[C_GTOLD];
    sum_dt = ctl[C_SUMDT]; mmax = ctl[C_MMAX]; hlast = ctl[C_HLAST]; ltlast = ctl[C_LTLAST];
    idid = ctl[C_IDID];
    bool newspc = ctl[C_NEWSPC] != 0.0, jacatt = ctl[C_JACATT] != 0.0, last = ctl[C_LAST] != 0.0;
    // work counters live in LDS, kept by thread 0 (they need no registers elsewhere)
    if (tid == 0)
      for (int q = 0; q < 6; q++) s_cnt[q] = ctl[C_STEPS + q];
    auto count = [&](int q, double v) {
      if (tid == 0) s_cnt[q] += v;
    };
#ifdef CB_PHASE_CLOCKS
    if (tid == 0)
      for (int q = 0; q < 8; q++) s_ph[q] = ctl[C_PHASE0 + q];
#endif
    int status = (int)ctl[C_STATUS];
    const double at = a.atol, rt = a.rtol, hmin = a.hmin, hmax = a.hmax;
    double t_cur = t0;            // time of the last set_current
    bool a_valid = false;         // A = F(y_n) with the current b

    // A = F(y_n) = (-b + L y_n) / C with y_n = hi + lo
    auto compute_A = [&]() {
      PHASE_T0(tA);
      T *h = L.d0, *l = L.d1;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (i < C) {
          h[i] = hi[k];
          if constexpr (sizeof(T) == 4) l[i] = lo[k];
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (i < C) {
          T s = lap<T>(L, h, i, hi[k]);
          if constexpr (sizeof(T) == 4) s += lap<T>(L, l, i, lo[k]);
          A[k] = (-L.bl[i] + s) * L.ic[i];
        }
      }
      __syncthreads();
      a_valid = true;
      PHASE_ADD(PH_A, tA);
    };

    // pkj_update_ion_RKC (pkj_ion.cu:304-358) when do_ion, then the frozen
    // current b (pkj_update_matrix, solve_rkc.cu:369-395).  Streams the
    // compartment's gates from HBM; one compartment per iteration so the
    // channel arithmetic does not multiply register pressure by K.  The
    // mixed engine updates and streams only the channels present in the
    // compartment's class (a dendrite: 9 of the 18 gates; zero-density
    // channels add exact zeros to the current and the Ca influx).
    auto ion_and_b = [&](double dt, bool do_ion, bool stim_live) {
      PHASE_T0(tI);
      // the synaptic current of the macro step, decayed over the interval
      // since the last ion update (solve_rkc.cu:344-367)
      const bool net_on = NET && a.net.row != nullptr;
      const T net_decay = net_on ? (T)exp(-dt / 0.6) : T(0);
      // the voltage for the update goes through LDS
      T *vyl = L.d1;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (i < C) vyl[i] = VY[k];
      }
      // FP64 engine: thread t updates its own compartments t, t + NT, ...
      // and reads back only what it wrote.  Mixed engine: the last, partial
      // block of compartments goes to the highest threads instead, so wave 0
      // (which holds the soma and the first dendrite, the compartments with
      // the long channel lists, in its first block) runs one block fewer;
      // the vyl exchange then needs a barrier on either side.
      constexpr bool REMAP = sizeof(T) == 4;
      const int kl = (C - 1) / NT, rem = C - kl * NT;
      if constexpr (REMAP) __syncthreads();
      // One compartment's inputs as streamed from the slab.  The mixed engine
      // streams the channels every class has (the dendrite's); the soma and
      // main-dendrite compartments fetch their extra gates when they are
      // updated.
      struct In {
        uint32_t meta;
        T ar, leak, rev, vo, shell, syn, net;
        double ca;
        T g[NGATE];
      };
      constexpr unsigned MS = (sizeof(T) == 4) ? channels_of_row(2) : CH_ALL;
      auto fetch = [&](int i, In &x) {
        // the flag word first: the loads return in order, so the class
        // select waits for it alone
        x.meta = S.meta(i);
        x.ar = S.ld(P_AREA, i);
        x.leak = S.ld(P_LEAK, i);
        x.rev = S.ld(P_REV, i);
        if (do_ion) {
          x.vo = S.ld(P_VOLD, i);
          x.ca = S.ca(i);
          x.shell = S.ld(P_SHELL, i);
        }
        if (a.syn_on) x.syn = S.ld(P_SYN, i);
        if (net_on) x.net = S.ld(p_nets<T>(), i);
#pragma unroll
        for (int q = 0; q < NGATE; q++)
          if (chan_of_gate(q) & MS) x.g[q] = S.ld(P_GATE + q, i);
      };
      auto body = [&](auto Mc, const In &x, int i) {
        constexpr unsigned M = decltype(Mc)::value;
        T g[NGATE];
#pragma unroll
        for (int q = 0; q < NGATE; q++)
          g[q] = (chan_of_gate(q) & M) ? ((chan_of_gate(q) & MS) ? x.g[q] : S.ld(P_GATE + q, i)) : T(0);
        const int og = (int)(x.meta & 0xFFFFu);
        const int fl = (int)(x.meta >> 16);
        const int cl = fl & CLS_MASK;
        const int row = (cl == SOMA) ? 0 : (cl == MAIN) ? 1 : 2;
        T c[NCOND];
        c[g_leak] = (T)(1.0 / ((cl == SOMA) ? RM_S : RM_D)) * x.leak;
        densities<T>(row, x.ar, c, std::make_index_sequence<10>{});
        T rev = x.rev;
        const T vn = vyl[i];
        if (do_ion) {
          const T vo = x.vo;
          double ca = x.ca;
          if constexpr (sizeof(T) == 4) {
            T gca = T(0);
            if constexpr ((M & CH_CAP) != 0) gca = c[g_CaP] * g[m_CaP] * g[h_CaP];
            if constexpr ((M & CH_CAT) != 0) gca = gca + c[g_CaT] * g[m_CaT] * g[h_CaT];
            // the shell plane holds 1 / (2 F shell) for the mixed engine
            const double caq = ca_rk4_mixed(vo, ca, rev, (double)x.shell, gca, dt, (fl & FIXED_REV) != 0);
            gates_cn_fast<M>(0.5f * (vn + vo), (float)caq, (float)dt, g);
          } else {
            double rv = rev;
            const double caq = ca_rk4(vo, ca, rv, x.shell, c[g_CaP] * g[m_CaP] * g[h_CaP],
                                      c[g_CaT] * g[m_CaT] * g[h_CaT], dt, (fl & FIXED_REV) != 0);
            rev = rv;
            gates_cn<T>(T(0.5) * (vn + vo), (T)caq, (T)dt, g);
          }
#pragma unroll
          for (int q = 0; q < NGATE; q++)
            if (chan_of_gate(q) & M) S.st(P_GATE + q, i, g[q]);
          S.set_ca(i, ca);
          S.st(P_REV, i, rev);
          S.st(P_VOLD, i, vn);
        }
        const T iext = (stim_live && og == a.stim_orig && a.stim_t0 <= t_cur && t_cur < a.stim_t1)
                           ? (T)a.amp[cell] : T(0);
        T bb = current_m<M, T>(vn, iext, rev, c, g);
        if (a.syn_on) bb += x.syn * (vn - a.syn_e);
        if (net_on) bb += x.net * net_decay;
        S.st(P_B, i, bb);
        L.bl[i] = bb;
      };
      // compartment of block k for this thread, or -1
      auto index_of = [&](int k) {
        int i = tid + k * NT;
        if (REMAP && k == kl) i = kl * NT + tid - (NT - rem);
        return (i >= k * NT && i < C) ? i : -1;
      };
#pragma unroll 1
      for (int k = 0; k <= kl; k++) {
        const int i = index_of(k);
        if (i < 0) continue;
        In x;
        fetch(i, x);
        if constexpr (sizeof(T) == 4) {
          const int cl = (int)(x.meta >> 16) & CLS_MASK;
          if (cl == SOMA)
            body(std::integral_constant<unsigned, channels_of_row(0)>{}, x, i);
          else if (cl == MAIN)
            body(std::integral_constant<unsigned, channels_of_row(1)>{}, x, i);
          else
            body(std::integral_constant<unsigned, channels_of_row(2)>{}, x, i);
        } else {
          body(std::integral_constant<unsigned, CH_ALL>{}, x, i);
        }
      }
      if constexpr (REMAP) __syncthreads();
      a_valid = false;
      PHASE_ADD(PH_ION, tI);
    };

    // rkcrho, solve_rkc.cu:562-754, carried on the perturbation
    // delta = v - y_n (in D1); D2 holds G = F(y_n + delta) - fn.
    auto rkcrho = [&]() {
      const int itmax = 1000;
      double sprad_old[3] = {0.0, 0.0, 0.0};
      const double sqrtu = sqrt(UROUND), small = 1.0 / 0.125;
      double sy = 0.0, se = 0.0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        D1[k] = (i < C) ? S.ld(P_EDE, i) : T(0);
        double y = (double)hi[k] + (double)lo[k];
        sy += y * y;
        se += (double)D1[k] * (double)D1[k];
      }
      double ynrm = sqrt(wg_sum<NT>(sy, red, slot));
      double vnrm = sqrt(wg_sum<NT>(se, red, slot));
      double dynrm;
      if (ynrm != 0.0 && vnrm != 0.0) {
        dynrm = ynrm * sqrtu;
        const T al = (T)(dynrm / vnrm);
#pragma unroll
        for (int k = 0; k < K; k++) D1[k] = al * D1[k];
      } else if (ynrm != 0.0) {
        dynrm = ynrm * sqrtu;
#pragma unroll
        for (int k = 0; k < K; k++) D1[k] = (T)sqrtu * hi[k];
      } else if (vnrm != 0.0) {
        dynrm = UROUND;
        const T al = (T)(dynrm / vnrm);
#pragma unroll
        for (int k = 0; k < K; k++) D1[k] = al * D1[k] - hi[k];
      } else {
        dynrm = UROUND;
#pragma unroll
        for (int k = 0; k < K; k++) D1[k] = (T)dynrm - hi[k];
      }
      T *cur = L.d0, *nxt = L.d1;
      double sigma = 0.0, sigmal;
      for (int iter = 0; iter < itmax; iter++) {
        count(3, 1.0);
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int i = tid + k * NT;
          if (i < C) cur[i] = D1[k];
        }
        __syncthreads();
        double sq = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int i = tid + k * NT;
          D2[k] = T(0);
          if (i < C) {
            const T G = (A[k] - F0[k]) + lap<T>(L, cur, i, D1[k]) * L.ic[i];
            D2[k] = G;
            sq += (double)G * (double)G;
          }
        }
        const double dfnrm = sqrt(wg_sum<NT>(sq, red, slot));
        T *tmp = cur; cur = nxt; nxt = tmp;
        sigmal = sigma;
        sigma = dfnrm / dynrm;
        sprad_old[0] = sprad_old[1];
        sprad_old[1] = sprad_old[2];
        sprad_old[2] = sprad;
        sprad = 1.2 * sigma;
        bool done = iter >= 1 && fabs(sigma - sigmal) <= fmax(sigma, small) * 0.001;
        if (!done && iter > itmax / 2 && fabs(sprad_old[2] - sprad_old[0]) < 0.00001 &&
            fabs(sprad_old[1] - sprad) < 0.00001) {
          sprad = (sprad + sprad_old[2]) / 2.0;
          done = true;
        }
        if (done) {
#pragma unroll
          for (int k = 0; k < K; k++) {
            const int i = tid + k * NT;
            if (i < C) S.st(P_EDE, i, D1[k]);
          }
          return;
        }
        if (dfnrm != 0.0) {
          const T al = (T)(dynrm / dfnrm);
#pragma unroll
          for (int k = 0; k < K; k++) D1[k] = al * (-D2[k]);   // v = y_n + alpha (fn - F(v))
        } else {
          const int idx = iter % C;
#pragma unroll
          for (int k = 0; k < K; k++)
            if (tid + k * NT == idx) D1[k] = -D1[k];
        }
      }
      idid = 6;
    };

    // rkclow, solve_rkc.cu:908-1096
    auto rkclow = [&]() {
      int loop_num = 0;
      while (true) {
        if (++loop_num > 1000) return;
        if (!a_valid) compute_A();
        if (newspc) {
          PHASE_T0(tR);
          rkcrho();
          if (a.rho_floor) sprad = fmax(sprad, a.rho_floor[cell]);
          jacatt = true;
          PHASE_ADD(PH_RHO, tR);
        }
        PHASE_T0(tS);
        last = false;
        if (1.1 * absh >= fabs(tend - lt)) {
          absh = fabs(tend - lt);
          last = true;
        }
        // (a non-finite h * rho takes the mmax branch instead of an undefined
        // conversion, so the coefficient table is always indexed in range;
        // the error estimate then stops the cell)
        const double mr = sqrt(1.54 * absh * sprad + 1.0);
        int m = (mr < mmax) ? 1 + (int)mr : (int)mmax + 1;
        if (m > mmax) {
          m = (int)mmax;
          absh = (m * m - 1) / (1.54 * sprad);
          last = false;
        }
        double h = absh;
        if (h <= hmin) h = hmin;
        // ---- step(h, m), solve_rkc.cu:815-906, on increments D_j = Y_j - Y_0 ----
        // the coefficients come from the per-m table; only a stage count
        // beyond it runs the recurrence here
        const double *ct = (m <= a.coef_mmax) ? a.coef + (size_t)m * a.coef_mmax * 4 : nullptr;
        double w0 = 0.0, w1 = 0.0, mus1;
        if (ct) {
          mus1 = ct[0];
        } else {
          w0 = 1.0 + 2.0 / (13.0 * m * m);
          const double temp1 = w0 * w0 - 1.0, temp2 = sqrt(temp1);
          const double arg = m * log(w0 + temp2);
          w1 = sinh(arg) * temp1 / (cosh(arg) * m * temp2 - w0 * sinh(arg));
          mus1 = w1 * (1.0 / ((2.0 * w0) * (2.0 * w0)));
        }
        T *cur = L.d0, *nxt = L.d1;
        {
          const T hm = (T)(h * mus1);
#pragma unroll
          for (int k = 0; k < K; k++) {
            D2[k] = T(0);
            D1[k] = hm * F0[k];
            VY[k] = hi[k];   // Y_{m-2} when m == 2
            const int i = tid + k * NT;
            if (i < C) cur[i] = D1[k];
          }
        }
        __syncthreads();
        PHASE_ADD(PH_SETUP, tS);
        PHASE_T0(tL);
        // stage j: D_j = mu D_{j-1} + nu D_{j-2} + h mu~ (F(Y_{j-1}) - a_{j-1} F_0)
        auto stage = [&](int j, T tmu, T tnu, T thm, T taj) {
          const bool lastj = (j == m - 1);
#pragma unroll
          for (int k = 0; k < K; k++) {
            const int i = tid + k * NT;
            if (i < C) {
              const T F = A[k] + lap<T>(L, cur, i, D1[k]) * L.ic[i];
              const T Dn = tmu * D1[k] + tnu * D2[k] + thm * (F - taj * F0[k]);
              if (lastj) VY[k] = hi[k] + (lo[k] + D2[k]);   // Y_{m-2}
              D2[k] = D1[k];
              D1[k] = Dn;
              nxt[i] = Dn;
            }
          }
          __syncthreads();
          T *tmp = cur; cur = nxt; nxt = tmp;
        };
        if (ct) {
          for (int j = 1; j < m; j++) {
            const double *c4 = ct + 4 * j;
            stage(j, (T)c4[0], (T)c4[1], (T)(h * c4[2]), (T)c4[3]);
          }
        } else {
          double bjm1 = 1.0 / ((2.0 * w0) * (2.0 * w0)), bjm2 = bjm1;
          double zjm1 = w0, zjm2 = 1.0, dzjm1 = 1.0, dzjm2 = 0.0, d2zjm1 = 0.0, d2zjm2 = 0.0;
          for (int j = 1; j < m; j++) {
            const double zj = 2.0 * w0 * zjm1 - zjm2;
            const double dzj = 2.0 * w0 * dzjm1 - dzjm2 + 2.0 * zjm1;
            const double d2zj = 2.0 * w0 * d2zjm1 - d2zjm2 + 4.0 * dzjm1;
            const double bj = d2zj / (dzj * dzj);
            const double ajm1 = 1.0 - zjm1 * bjm1;
            const double mu = 2.0 * w0 * bj / bjm1;
            const double nu = -bj / bjm2;
            const double musj = mu * w1 / w0;
            stage(j, (T)mu, (T)nu, (T)(h * musj), (T)ajm1);
            bjm2 = bjm1; bjm1 = bj;
            zjm2 = zjm1; zjm1 = zj;
            dzjm2 = dzjm1; dzjm1 = dzj;
            d2zjm2 = d2zjm1; d2zjm1 = d2zj;
          }
        }
        count(2, m + 1);
        PHASE_ADD(PH_STAGES, tL);
        PHASE_T0(tE);
        // ---- f(Y_m) (into D2) and the error estimate, solve_rkc.cu:58-73 ----
        // (every weight is positive: atol > 0 is enforced at creation)
        double sq = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int i = tid + k * NT;
          D2[k] = T(0);
          if (i < C) {
            const T Fm = A[k] + lap<T>(L, cur, i, D1[k]) * L.ic[i];
            D2[k] = Fm;
            const T yn = hi[k] + lo[k];
            const T y = hi[k] + (lo[k] + D1[k]);
            const T ay = absT(y), ayn = absT(yn);
            const double wt = at + rt * (double)((ay > ayn) ? ay : ayn);
            const double est = 0.8 * (double)(-D1[k]) + 0.4 * h * ((double)F0[k] + (double)Fm);
            const double q = est / wt;
            sq += q * q;
          }
        }
        double err = wg_sum<NT>(sq, red, slot);
        err = sqrt(err / (double)C);
        if (!(err <= 1.0e300)) {   // NaN or infinite state: stop this cell
          idid = NONFINITE;
          PHASE_ADD(PH_ERR, tE);
          return;
        }
        if (err > 1.0) {
          count(1, 1.0);
          absh = 0.8 * absh / ctl_pow<T>(err, 1);
          if (absh < hmin) {
            idid = 4;   // falls through to the accept branch, as the reference does
          } else {
            newspc = !jacatt;
            PHASE_ADD(PH_ERR, tE);
            continue;
          }
        }
        // ---- accept ----
        count(0, 1.0);
        naccpt += 1;
        lt += h;
        gt = lt;
        jacatt = false;
        newspc = false;   // nstsig is never advanced in the reference
        hlast = h;
        ltlast = lt;
#pragma unroll
        for (int k = 0; k < K; k++) {
          two_sum_acc<T>(hi[k], lo[k], D1[k]);
          F0[k] = D2[k];
        }
        a_valid = false;   // A is rebuilt from y_{n+1}, so rounding does not accumulate
        double fac = 10.0, tp1, tp2;
        if (naccpt == 1) {
          tp2 = ctl_pow<T>(err, 1);
          if (0.8 < fac * tp2) fac = 0.8 / tp2;
        } else {
          tp1 = 0.8 * absh * ctl_pow<T>(errold, 1);
          tp2 = fabs(hold) * ctl_pow<T>(err, 2);
          if (tp1 < fac * tp2) fac = tp1 / tp2;
        }
        absh = fmax(0.1, fac) * absh;
        absh = fmax(hmin, fmin(hmax, absh));
        errold = err;
        hold = h;
        idid = last ? 1 : 2;
        PHASE_ADD(PH_ERR, tE);
        return;
      }
    };

    if constexpr (INIT) {
      // ---- RKC_vec_initialize (solve_rkc.cu:756-791): b with i_ext = 0 ----
      ion_and_b(0.0, false, false);
      compute_A();
#pragma unroll
      for (int k = 0; k < K; k++) {
        F0[k] = A[k];
        const int i = tid + k * NT;
        if (i < C) S.st(P_EDE, i, F0[k]);
      }
      rkcrho();
      if (a.rho_floor) sprad = fmax(sprad, a.rho_floor[cell]);
      // set_initial_absh (solve_rkc.cu:489-560)
      double ab = hmax;
      if (sprad * ab > 1.0) ab = 1.0 / sprad;
      ab = fmax(ab, hmin);
      T *cur = L.d0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        D1[k] = (T)ab * F0[k];
        if (i < C) cur[i] = D1[k];
      }
      __syncthreads();
      double sq = 0.0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (i < C) {
          const T F = A[k] + lap<T>(L, cur, i, D1[k]) * L.ic[i];
          const double yn = (double)hi[k] + (double)lo[k];
          const double wt = at + rt * fabs(yn);
          const double d = (double)(F - F0[k]) * (1.0 / wt);
          sq += d * d;
        }
      }
      double est = wg_sum<NT>(sq, red, slot);
      est = ab * sqrt(est / C);
      if (0.1 * ab < hmax * sqrt(est))
        ab = fmax(0.1 * ab / sqrt(est), hmin);
      else
        ab = hmax;
      absh = ab;
      __syncthreads();
    } else {
      for (int s = 0; s < nsteps && status == 0; s++) {
        t_cur = t0 + s * a.interval;
        if (a.syn_on) {
          for (int i = tid; i < C; i += NT) {
            const int og = (int)(S.meta(i) & 0xFFFFu);
            const u32x4 r = philox4x32_10((uint32_t)(cell * COMP + og), (uint32_t)(step0 + s), 0x5359u,
                                          0u, a.syn_k0, a.syn_k1);
            const double u = u01_53(r.v[0], r.v[1]);
            S.st(P_SYN, i, S.ld(P_SYN, i) * a.syn_decay + ((u < a.syn_p) ? a.syn_g_event : T(0)));
          }
        }
        // ---- pkj_solve_by_rkc, solve_rkc.cu:1252-1299 ----
        // A is rebuilt right after every update, so it is dead across the
        // update's loop and holds no registers there
        auto ion_step = [&]() {
          count(4, 1.0);
          ion_and_b(sum_dt, true, true);
          compute_A();
          if (a.refresh_fn) {   // fn = F(y_n) with the new b
#pragma unroll
            for (int k = 0; k < K; k++) F0[k] = A[k];
          }
          sum_dt = 0.0;
        };
        if (sum_dt != 0.0) ion_step();
        while (true) {
          rkclow();
          sum_dt += gt - gt_old;
          gt_old = gt;
          if (sum_dt >= a.integ && idid == 2.0) ion_step();
          if (idid == 2.0) continue;
          if (idid == 1.0) break;
          status = (int)idid;
          break;
        }
        tend += a.interval;
        count(5, 1.0);
        // v_new = y_{n+1} (rkcint at s = 1); record the chosen compartment
        // and mark its upward threshold crossing
        if (a.rec || a.spk) {
#pragma unroll
          for (int k = 0; k < K; k++)
            if (rec_mask >> k & 1u) {
              const double v = (double)hi[k] + (double)lo[k];
              if (a.rec) a.rec[(size_t)s * a.ncell + cell] = v;
              if (a.spk) {
                if (v > a.thr && a.vprev[cell] <= a.thr)
                  atomicOr(&a.spk[(size_t)s * a.spk_words + (cell >> 5)], 1u << (cell & 31));
                a.vprev[cell] = v;
              }
            }
        }
      }
    }

    // ---- write back ----
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = tid + k * NT;
      if (i < C) {
        S.st(P_Y0HI, i, hi[k]);
        S.st(P_Y0LO, i, lo[k]);
        S.st(P_F0, i, F0[k]);
        S.st(P_VY, i, VY[k]);
        a.vnew_out[base + (S.meta(i) & 0xFFFFu)] = (double)hi[k] + (double)lo[k];
      }
    }
    if (tid == 0) {
      ctl[C_ABSH] = absh; ctl[C_SPRAD] = sprad; ctl[C_NACCPT] = naccpt; ctl[C_ERROLD] = errold;
      ctl[C_HOLD] = hold; ctl[C_LT] = lt; ctl[C_TEND] = tend; ctl[C_GT] = gt; ctl[C_GTOLD] = gt_old;
      ctl[C_SUMDT] = sum_dt; ctl[C_MMAX] = mmax; ctl[C_HLAST] = hlast; ctl[C_LTLAST] = ltlast;
      ctl[C_NEWSPC] = newspc; ctl[C_JACATT] = jacatt; ctl[C_LAST] = last; ctl[C_IDID] = idid;
      for (int q = 0; q < 6; q++) ctl[C_STEPS + q] = s_cnt[q];
      ctl[C_STATUS] = status;
#ifdef CB_PHASE_CLOCKS
      for (int q = 0; q < 8; q++) ctl[C_PHASE0 + q] = s_ph[q];
#endif
    }
    __syncthreads();
  }
}

// Occupancy target: 4 waves per SIMD for FP32 (4 cells per CU), 2 for FP64.
// NET: the variant with network synapses (their current enters b at every
// ion update); the plain variant carries none of that code
template <typename T, int NT, int K, bool NET>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2, 8)))
k_cell_run(CellArgs<T> a, int nsteps, double t0, int step0) {
  cell_body<T, NT, K, false, NET>(a, nsteps, t0, step0);
}

template <typename T, int NT, int K>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2, 8)))
k_cell_init(CellArgs<T> a) {
  cell_body<T, NT, K, true, false>(a, 0, 0.0, 0);
}

// Device scratch of cb_pop_run_noalloc: bytes [0, 256) hold the work
// counter of the launch, the rest the recording of rec_comp before its copy
// to the caller's buffer.  A file-scope array, so a call allocates nothing;
// every device the library runs on holds one copy until the process ends,
// shared by all populations and calls on that device (overlapping calls on
// one device race on it).  Host code uses the device address that
// hipGetSymbolAddress returns for the current device, never the name.
constexpr size_t SCRATCH_BYTES = 8u << 20;
constexpr size_t SCRATCH_REC = 256;   // offset of the recording
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};

// the scratch array's address on the current device, looked up once per device
unsigned char *scratch_on_current_device() {
  int dev = 0;
  CB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw Error("device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    CB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[dev] = (unsigned char *)ptr;
  }
  return g_scratch[dev];
}

// clears the work counter with a kernel rather than a memset node, so a
// caller may capture the launch sequence in a graph
__global__ void k_zero_int(int *p) { *p = 0; }

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

// The macro step's synaptic current from its starting y, before the launch
// of that one macro step: per compartment the sum over its synapses of
// val (y - E) in entry order (the current the reference adds at each ion
// update, decayed; solve_rkc.cu:344-367), into its slab plane.  One thread
// per (cell, compartment position); y = hi + lo from the launch-edge planes.
template <typename T>
__global__ void __launch_bounds__(256) k_cell_net(CellArgs<T> a) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= (size_t)a.ncell * a.comp) return;
  const int cell = (int)(q / a.comp), i = (int)(q % a.comp);
  const Slab<T> S(a.slab + (size_t)cell * a.slab_stride, (int)a.slab_stride);
  const int g = cell * a.comp + (int)(S.meta(i) & 0xFFFFu);
  const double y = (double)S.ld(P_Y0HI, i) + (double)S.ld(P_Y0LO, i);
  double acc = 0.0;
  for (int j = a.net.row[g]; j < a.net.row[g + 1]; j++) {
    const int sl = a.net.slot[j];
    const int st = sl >> 30;
    acc += a.net.val[st][sl & 0x3fffffff] * (y - a.net.E[st]);
  }
  S.st(p_nets<T>(), i, (T)acc);
}

template <typename T>
class CellPkj : public Engine {
 public:
  struct Variant {
    int nt, k;
  };

  CellPkj(const PkjHost &H, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn)
      : n_(H.n), C_(H.comp), syn_(syn) {
    using namespace pkj;
    create_stream(&st_);
    CB_HIP(hipEventCreate(&ev0_));
    CB_HIP(hipEventCreate(&ev1_));
    const int n = n_, C = C_;
    const size_t NC = (size_t)n * C;
    nc_ = (int)NC;
    // breadth-first renumbering per cell: soma first, children contiguous
    std::vector<int> perm(NC), inv(NC);   // perm[bfs] = original global, inv[orig] = bfs
    std::vector<std::vector<int>> kids(C);
    for (int c = 0; c < n; c++) {
      for (int i = 0; i < C; i++) kids[i].clear();
      int root = -1;
      for (int i = 0; i < C; i++) {
        int p = H.parent[(size_t)c * C + i];
        if (p < 0) {
          if (root >= 0) throw Error("cell " + std::to_string(c) + " has more than one root");
          root = i;
        } else {
          kids[p - c * C].push_back(i);
        }
      }
      if (root < 0) throw Error("cell " + std::to_string(c) + " has no root");
      std::deque<int> q{root};
      int nb = 0;
      while (!q.empty()) {
        int u = q.front();
        q.pop_front();
        perm[(size_t)c * C + nb] = c * C + u;
        inv[(size_t)c * C + u] = c * C + nb;
        nb++;
        for (int v : kids[u]) q.push_back(v);
      }
      if (nb != C) throw Error("cell " + std::to_string(c) + " is not a single tree");
    }
    std::vector<uint32_t> topo(NC);
    std::vector<T> gpar(NC), invc(NC);
    if (C > PITCH) throw Error("cell too large for the per-cell engine");
    stride_ = (size_t)n_planes<T>() * PITCH * sizeof(T);
    std::vector<unsigned char> slab((size_t)n * stride_, 0);
    auto put_plane = [&](int c, int p, int b, auto v) {
      std::memcpy(slab.data() + (size_t)c * stride_ + (size_t)p * PITCH * sizeof(T) + (size_t)b * sizeof(v), &v,
                  sizeof(v));
    };
    for (int c = 0; c < n; c++) {
      // children count / first child in bfs order
      std::vector<int> fc(C, 0), nch(C, 0);
      for (int b = 0; b < C; b++) {
        int og = perm[(size_t)c * C + b];
        int p = H.parent[og];
        if (p >= 0) {
          int pb = inv[p] - c * C;
          if (nch[pb] == 0) fc[pb] = b;
          nch[pb]++;
        }
      }
      for (int b = 0; b < C; b++) {
        const size_t g = (size_t)c * C + b;
        const int og = perm[g];
        int p = H.parent[og];
        int pb = p >= 0 ? inv[p] - c * C : NO_PARENT;
        if (nch[b] > 3) throw Error("compartment with more than 3 children");
        if (C > 2047) throw Error("cell too large for the per-cell engine");
        topo[g] = (uint32_t)pb | ((uint32_t)fc[b] << 11) | ((uint32_t)nch[b] << 22);
        gpar[g] = (T)H.gpar[og];
        invc[g] = (T)(1.0 / H.cm[og]);
        const int cl = H.compart[og];
        uint32_t flags = (uint32_t)cl;
        if ((long)og % 1599 == 0) flags |= FIXED_REV;   // og is the global index here
        put_plane(c, P_META, b, (uint32_t)(og - c * C) | flags << 16);
        put_plane(c, P_AREA, b, (T)H.area[og]);
        // leak conductance / (1/RM) recovers the spine-corrected area
        put_plane(c, P_LEAK, b, (T)(H.cond[(size_t)g_leak * NC + og] / (1.0 / ((cl == SOMA) ? RM_S : RM_D))));
        // the mixed engine keeps 1 / (2 F shell); the FP64 engine the shell itself
        put_plane(c, P_SHELL, b, (sizeof(T) == 4) ? (T)(1.0 / (2.0 * FARADAY * H.shell[og])) : (T)H.shell[og]);
        put_plane(c, P_Y0HI, b, (T)H.v[og]);
        put_plane(c, P_VOLD, b, (T)H.v[og]);
        put_plane(c, P_VY, b, (T)H.v[og]);
        put_plane(c, P_REV, b, (T)H.rev[og]);
        put_plane(c, P_CA, b, H.ca[og]);
        for (int q = 0; q < NGATE; q++) put_plane(c, P_GATE + q, b, (T)H.gate[(size_t)q * NC + og]);
      }
    }
    perm_ = perm;
    a_.ncell = n;
    a_.comp = C;
    a_.atol = sol.atol;
    a_.rtol = sol.rtol;
    a_.hmin = sol.hmin_or(std::pow(2.0, -10.0));   // H_MIN_PKJ, solve_rkc.cuh:26
    a_.hmax = sol.interval;
    a_.interval = sol.interval;
    a_.integ = sol.ion_interval > 0 ? sol.ion_interval : std::pow(2.0, std::log10(sol.atol));
    a_.refresh_fn = sol.refresh_fn ? 1 : 0;
    a_.rho_floor = nullptr;
    if (!sol.rho_floor.empty()) a_.rho_floor = put(sol.rho_floor);
    a_.coef_mmax = COEF_MMAX;
    a_.coef = put(rkc_stage_table(COEF_MMAX));
    a_.topo = put(topo);
    a_.gpar = put(gpar);
    a_.invc = put(invc);
    a_.slab = put(slab);
    a_.slab_stride = (uint32_t)stride_;
    std::vector<double> ctl((size_t)n * NCTL, 0.0);
    for (int c = 0; c < n; c++) {
      double *q = &ctl[(size_t)c * NCTL];
      q[C_TEND] = sol.interval;
      q[C_LAST] = 1;
      q[C_NEWSPC] = 1;
      q[C_JACATT] = 0;
      q[C_MMAX] = nint_ori(std::sqrt(sol.rtol / (10.0 * UROUND)));
    }
    a_.ctl = put(ctl);
    vnew_ = dalloc<double>(NC);
    a_.vnew_out = vnew_;
    a_.stim_orig = stim.comp;
    a_.stim_t0 = stim.t0;
    a_.stim_t1 = stim.t1;
    std::vector<double> amp = stim.amp;
    amp.resize(n, pkj::CURRENT);
    a_.amp = put(amp);
    a_.syn_on = syn.on ? 1 : 0;
    a_.syn_decay = (T)std::exp(-sol.interval / syn.tau_ms);
    a_.syn_g_event = (T)syn.g_event;
    a_.syn_e = (T)syn.e_syn;
    a_.syn_p = 1.0 - std::exp(-syn.rate_hz * sol.interval * 1e-3);
    a_.syn_k0 = (uint32_t)syn.seed;
    a_.syn_k1 = (uint32_t)(syn.seed >> 32);
    a_.rec_orig = 1599;
    a_.rec = nullptr;
    a_.queue = dalloc<int>(1);
    {
      std::vector<int> id(n);
      for (int c = 0; c < n; c++) id[c] = c;
      order_ = dalloc<int>(n);
      CB_HIP(hipMemcpy(order_, id.data(), sizeof(int) * n, hipMemcpyHostToDevice));
      a_.order = order_;
      cost_seen_.assign(n, 0.0);
    }
    variants_ = variants();
    choose_default();
    scratch_ = scratch_on_current_device();
    launch(1, 0, 0.0);
    check_status();
  }

  ~CellPkj() override {
    (void)hipStreamSynchronize(st_);
    for (void *p : bufs_) (void)hipFree(p);
    (void)hipFree(vnew_);
    (void)hipFree(a_.queue);
    (void)hipFree(order_);
    if (rec_) (void)hipFree(rec_);
    if (spk_) (void)hipFree(spk_);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(st_);
  }

  int nc() const override { return nc_; }
  // granule -> Purkinje synapses (the inhibitory set couples nothing): the
  // CSR by global compartment, read at every macro-step start
  void set_synapses(const SynIn &in) override { a_.net = in; }
  const double *device_v() const override { return vnew_; }
  const double *device_vnew() const override { return vnew_; }
  hipStream_t stream() const override { return st_; }
  int ncell() const override { return n_; }
  Stats stats() const override {
    std::vector<double> ctl((size_t)n_ * NCTL);
    CB_HIP(hipMemcpy(ctl.data(), a_.ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost));
    Stats s;
    for (int c = 0; c < n_; c++) {
      const double *q = &ctl[(size_t)c * NCTL];
      s.steps += (long)q[C_STEPS];
      s.rejects += (long)q[C_REJECTS];
      s.stages += (long)q[C_STAGES];
      s.rho_iters += (long)q[C_RHO];
      s.ion_updates += (long)q[C_IONS];
    }
    s.macro_steps = macro_;
    return s;
  }

  void set_current(double t) override { t_next_ = t; }

  void advance() override { run(1, t_next_, -1, nullptr); }

  void run(int nsteps, double t0, int rec_comp, double *rec_out) override {
    if (nsteps <= 0) return;
    if (rec_out && rec_comp < 0) throw Error("the per-cell engines record one compartment per cell");
    // spikes are found in the kernel on the recorded compartment
    const bool raster = raster_comp_ >= 0;
    if (raster && rec_out && rec_comp != raster_comp_)
      throw Error("the per-cell engines detect spikes on the recorded compartment");
    double *rec = nullptr;
    if (rec_out) {
      if (rec_cap_ < (size_t)nsteps * n_) {
        if (rec_) CB_HIP(hipFree(rec_));
        rec_cap_ = (size_t)nsteps * n_;
        rec_ = dalloc<double>(rec_cap_);
      }
      rec = rec_;
      a_.rec_orig = rec_comp;
    }
    const int words = (n_ + 31) / 32;
    if (raster) {
      if (spk_cap_ < (size_t)nsteps * words) {
        if (spk_) CB_HIP(hipFree(spk_));
        spk_cap_ = (size_t)nsteps * words;
        spk_ = dalloc<uint32_t>(spk_cap_);
      }
      CB_HIP(hipMemsetAsync(spk_, 0, sizeof(uint32_t) * (size_t)nsteps * words, st_));
      a_.rec_orig = raster_comp_;
      a_.thr = raster_thr_;
      a_.vprev = raster_prev();
      a_.spk = spk_;
      a_.spk_words = words;
    }
    a_.rec = rec;
    if (a_.net.on()) {
      // the synaptic current is fixed per macro step from that step's start,
      // and the synapse sets update between macro steps
      if (nsteps != 1) throw Error("with synapses the per-cell engines advance one macro step per call");
      k_cell_net<T><<<(unsigned)(((size_t)n_ * C_ + 255) / 256), 256, 0, st_>>>(a_);
    }
    launch(0, nsteps, t0);
    a_.rec = nullptr;
    a_.spk = nullptr;
    macro_ += nsteps;
    if (rec_out) {
      CB_HIP(hipMemcpyAsync(rec_out, rec, sizeof(double) * (size_t)nsteps * n_, hipMemcpyDeviceToHost, st_));
    }
    std::vector<uint32_t> bits((size_t)(raster ? nsteps : 0) * words);
    if (raster) CB_HIP(hipMemcpyAsync(bits.data(), spk_, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, st_));
    CB_HIP(hipStreamSynchronize(st_));
    if (raster) append_spikes(bits, nsteps, words, t0, a_.interval);
    collect_time();
    check_status();
    t_next_ = t0 + nsteps * a_.interval;
    if (nsteps >= 4) reorder();
  }

  // run() with the work counter and the recording in the device's scratch
  // array: no hipMalloc / hipFree on this path (the population's own arrays
  // were allocated at creation).  nsteps x ncell recorded doubles must fit
  // the scratch after its first 256 bytes.
  void run_noalloc(int nsteps, double t0, int rec_comp, double *rec_out) override {
    if (nsteps <= 0) return;
    const size_t need = rec_out ? sizeof(double) * (size_t)nsteps * n_ : 0;
    if (need > SCRATCH_BYTES - SCRATCH_REC)
      throw Error("allocation-free run: " + std::to_string(nsteps) + " steps x " + std::to_string(n_) +
                  " cells of recording exceed the " + std::to_string((SCRATCH_BYTES - SCRATCH_REC) >> 20) +
                  " MiB scratch");
    int *own_queue = a_.queue;
    a_.queue = (int *)scratch_;
    a_.rec = rec_out ? (double *)(scratch_ + SCRATCH_REC) : nullptr;
    if (rec_out) a_.rec_orig = rec_comp;
    try {
      if (a_.net.on()) {
        if (nsteps != 1) throw Error("with synapses the per-cell engines advance one macro step per call");
        k_cell_net<T><<<(unsigned)(((size_t)n_ * C_ + 255) / 256), 256, 0, st_>>>(a_);
      }
      launch(0, nsteps, t0);
      if (rec_out)
        CB_HIP(hipMemcpyAsync(rec_out, a_.rec, need, hipMemcpyDeviceToHost, st_));
      CB_HIP(hipStreamSynchronize(st_));
    } catch (...) {
      a_.queue = own_queue;
      a_.rec = nullptr;
      throw;
    }
    a_.queue = own_queue;
    a_.rec = nullptr;
    macro_ += nsteps;
    collect_time();
    check_status();
    t_next_ = t0 + nsteps * a_.interval;
    if (nsteps >= 4) reorder();
  }

  // Longest-first queue order for the next launch.  More cells are queued
  // than the chip holds at once (two 512-thread workgroups per CU), and a
  // launch ends with its slowest slot; handing out the cells that cost most
  // in the last run first (cost from the work counters: ~2.4 k cycles per
  // stage, 5.8 k per step attempt, 27 k per ion update) lets the cheap ones
  // fill the tail.  Results do not depend on the order: every cell is
  // advanced by itself.
  void reorder() {
    std::vector<double> ctl((size_t)n_ * NCTL);
    CB_HIP(hipMemcpy(ctl.data(), a_.ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost));
    std::vector<double> cost(n_);
    for (int c = 0; c < n_; c++) {
      const double *q = &ctl[(size_t)c * NCTL];
      const double total = 2.4 * q[C_STAGES] + 5.8 * (q[C_STEPS] + q[C_REJECTS]) + 27.0 * q[C_IONS];
      cost[c] = total - cost_seen_[c];
      cost_seen_[c] = total;
    }
    std::vector<int> ord(n_);
    for (int c = 0; c < n_; c++) ord[c] = c;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    CB_HIP(hipMemcpy(order_, ord.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
  }

  void read(int field, double *dst) override {
    using namespace pkj;
    const size_t NC = nc_;
    if (field == F_VNEW) {
      CB_HIP(hipMemcpy(dst, vnew_, sizeof(double) * NC, hipMemcpyDeviceToHost));
      return;
    }
    // one plane of every cell, in breadth-first order
    auto plane = [&](int p, auto *h) {
      using U = std::remove_pointer_t<decltype(h)>;
      CB_HIP(hipMemcpy2D(h, sizeof(U) * C_, a_.slab + (size_t)p * PITCH * sizeof(T), stride_, sizeof(U) * C_, n_,
                         hipMemcpyDeviceToHost));
    };
    auto fetchT = [&](int p) {
      std::vector<T> h(NC);
      plane(p, h.data());
      for (size_t g = 0; g < NC; g++) dst[perm_[g]] = (double)h[g];
    };
    switch (field) {
      case F_V: fetchT(P_VOLD); return;
      case F_REV: fetchT(P_REV); return;
      case F_Y: fetchT(P_VY); return;
      case F_FN: fetchT(P_F0); return;
      case F_YN: {
        std::vector<T> h(NC), l(NC);
        plane(P_Y0HI, h.data());
        plane(P_Y0LO, l.data());
        for (size_t g = 0; g < NC; g++) dst[perm_[g]] = (double)h[g] + (double)l[g];
        return;
      }
      case F_CTL:
        CB_HIP(hipMemcpy(dst, a_.ctl, sizeof(double) * (size_t)n_ * NCTL, hipMemcpyDeviceToHost));
        return;
      case F_B: fetchT(P_B); return;
      case F_CA: {
        std::vector<double> h(NC);
        plane(P_CA, h.data());
        for (size_t g = 0; g < NC; g++) dst[perm_[g]] = h[g];
        return;
      }
      default:
        if (field >= F_GATE0 && field < F_GATE0 + NGATE) {
          fetchT(P_GATE + (field - F_GATE0));
          return;
        }
    }
    throw Error("field not available from the per-cell engine");
  }

  // Picks the (threads per cell, compartments per thread) mapping: every
  // variant runs the same short trial from the same state and the fastest
  // is kept.  State is restored after each trial.
  int autotune(int trial_steps, double t0) override {
    if (variants_.size() < 2 || trial_steps <= 0) return variants_[cur_].nt;
    if (t0 < 0) t0 = t_next_;
    Snapshot snap = snapshot();
    double best = 1e30;
    int bi = cur_;
    for (size_t v = 0; v < variants_.size(); v++) {
      cur_ = (int)v;
      launch(0, 1, t0);   // warm
      restore(snap);
      CB_HIP(hipStreamSynchronize(st_));
      timed_pending_ = false;
      launch(0, trial_steps, t0);
      CB_HIP(hipStreamSynchronize(st_));
      float ms = 0.0f;
      CB_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
      timed_pending_ = false;
      restore(snap);
      if (ms < best) {
        best = ms;
        bi = (int)v;
      }
      tune_log_.push_back({variants_[v].nt, variants_[v].k, ms * 1e-3});
    }
    cur_ = bi;
    release(snap);
    return variants_[cur_].nt;
  }

  bool select_mapping(int threads) override {
    for (size_t v = 0; v < variants_.size(); v++)
      if (variants_[v].nt == threads && (long)variants_[v].nt * variants_[v].k >= C_) {
        cur_ = (int)v;
        return true;
      }
    return false;
  }

  struct TuneRec {
    int nt, k;
    double sec;
  };
  // for the batched mix engine, which launches this population's cells
  // inside its own kernel: the kernel arguments, the dynamic LDS a cell
  // needs and the device buffers a run changes
  CellArgs<T> &args() { return a_; }
  size_t cell_lds_bytes() const { return lds_bytes(); }
  std::vector<std::pair<void *, size_t>> state_buffers() { return mutable_bufs(); }
  std::vector<TuneRec> tune_log_;
  int chosen_nt() const { return variants_[cur_].nt; }
  std::string mapping() const override {
    std::string m = "one workgroup per cell, " + std::to_string(variants_[cur_].nt) + " threads x " +
                    std::to_string(variants_[cur_].k) + " compartments per thread";
    for (const TuneRec &r : tune_log_)
      m += "; trial " + std::to_string(r.nt) + " x " + std::to_string(r.k) + ": " + std::to_string(r.sec * 1e3) + " ms";
    return m;
  }

 private:
  struct Snapshot {
    std::vector<std::pair<void *, std::pair<void *, size_t>>> saved;
  };

  Snapshot snapshot() {
    Snapshot s;
    for (auto &m : mutable_bufs()) {
      void *p = nullptr;
      CB_HIP(hipMalloc(&p, m.second));
      CB_HIP(hipMemcpyAsync(p, m.first, m.second, hipMemcpyDeviceToDevice, st_));
      s.saved.push_back({m.first, {p, m.second}});
    }
    return s;
  }
  void restore(const Snapshot &s) {
    for (auto &e : s.saved)
      CB_HIP(hipMemcpyAsync(e.first, e.second.first, e.second.second, hipMemcpyDeviceToDevice, st_));
  }
  void release(Snapshot &s) {
    CB_HIP(hipStreamSynchronize(st_));
    for (auto &e : s.saved) (void)hipFree(e.second.first);
    s.saved.clear();
  }
  // everything a run writes: the slabs, the controllers and V at the launch end
  std::vector<std::pair<void *, size_t>> mutable_bufs() {
    return {{a_.slab, stride_ * n_}, {a_.ctl, sizeof(double) * (size_t)n_ * NCTL}, {a_.vnew_out, sizeof(double) * nc_}};
  }

  static int nint_ori(double x) {
    if (x - (double)(int)x < 0.5) return (int)(double)(int)x;
    return (int)(double)(int)(x + 0.9);
  }

  template <typename U>
  U *put(const std::vector<U> &h) {
    U *d = dalloc<U>(h.size());
    h2d(d, h);
    bufs_.push_back((void *)d);
    return d;
  }

  static std::vector<Variant> variants() {
    std::vector<Variant> v;
    if (sizeof(T) == 4) {
      v.push_back({256, 7});
      v.push_back({320, 5});
      v.push_back({512, 4});
    } else {
      v.push_back({256, 7});
      v.push_back({320, 5});
    }
    return v;
  }

  // the kernel's LDS layout (cell_body): FP32 d0 d1 ic bl, the edge
  // conductances (16-byte aligned) and neighbour indices, the reduction
  // slots (16-byte aligned); FP64 d0 d1 gp ic bl tp and the reduction slots
  size_t lds_bytes() const {
    const size_t red = 16 + 2 * 16 * sizeof(double);
    if (sizeof(T) == 4) {
      size_t b = (size_t)C_ * 4 * sizeof(T);
      b = (b + 15) & ~(size_t)15;
      b += (size_t)C_ * (sizeof(float4) + sizeof(uint2));
      return b + red;
    }
    return (size_t)C_ * (5 * sizeof(T) + 4) + red;
  }

  void choose_default() {
    cur_ = 0;
    for (size_t v = 0; v < variants_.size(); v++)
      if (variants_[v].nt == (sizeof(T) == 4 ? 320 : 256)) cur_ = (int)v;
  }

  template <int NT, int K>
  void go(int mode, int nsteps, double t0, int grid) {
    if (mode == 1)
      hipLaunchKernelGGL((k_cell_init<T, NT, K>), dim3(grid), dim3(NT), lds_bytes(), st_, a_);
    else if (a_.net.on())
      hipLaunchKernelGGL((k_cell_run<T, NT, K, true>), dim3(grid), dim3(NT), lds_bytes(), st_, a_, nsteps, t0,
                         (int)macro_);
    else
      hipLaunchKernelGGL((k_cell_run<T, NT, K, false>), dim3(grid), dim3(NT), lds_bytes(), st_, a_, nsteps, t0,
                         (int)macro_);
  }

  void launch(int mode, int nsteps, double t0) {
    const Variant &v = variants_[cur_];
    if ((long)v.nt * v.k < C_) throw Error("per-cell mapping too small for the cell");
    k_zero_int<<<1, 1, 0, st_>>>(a_.queue);
    int dev = 0;
    CB_HIP(hipGetDevice(&dev));
    int ncu = 0;
    CB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = std::max(1, (int)(160 * 1024 / (lds_bytes() + 1024)));
    int grid = std::min(n_, ncu * per_cu);
    if (mode == 0) CB_HIP(hipEventRecord(ev0_, st_));
    switch (v.nt) {
      case 256: go<256, 7>(mode, nsteps, t0, grid); break;
      case 320: go<320, 5>(mode, nsteps, t0, grid); break;
      case 512: go<512, 4>(mode, nsteps, t0, grid); break;
      default: throw Error("unknown mapping");
    }
    CB_HIP(hipGetLastError());
    if (mode == 0) {
      CB_HIP(hipEventRecord(ev1_, st_));
      timed_pending_ = true;
    }
  }

  // called after a stream synchronisation that follows a timed launch
  void collect_time() {
    if (!timed_pending_) return;
    float ms = 0.0f;
    CB_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kt_ms_ += ms;
    kt_n_ += 1;
    timed_pending_ = false;
  }

 public:
  void kernel_time(double *ms, long *launches, bool reset) override {
    *ms = kt_ms_;
    *launches = kt_n_;
    if (reset) {
      kt_ms_ = 0.0;
      kt_n_ = 0;
    }
  }

 private:

  void check_status() {
    std::vector<double> ctl((size_t)n_ * NCTL);
    CB_HIP(hipMemcpyAsync(ctl.data(), a_.ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost, st_));
    CB_HIP(hipStreamSynchronize(st_));
    for (int c = 0; c < n_; c++) {
      int s = (int)ctl[(size_t)c * NCTL + C_STATUS];
      if (s == NONFINITE) throw Error("RKC: non-finite state in cell " + std::to_string(c), s);
      if (s != 0) throw Error("RKC failed in cell " + std::to_string(c) + ": idid = " + std::to_string(s), s);
    }
  }

  int n_, C_, nc_ = 0;
  size_t stride_ = 0;   // slab bytes per cell
  PoissonSpec syn_;
  hipStream_t st_;
  CellArgs<T> a_{};
  std::vector<void *> bufs_;
  double *vnew_ = nullptr;
  double *rec_ = nullptr;
  size_t rec_cap_ = 0;
  uint32_t *spk_ = nullptr;
  size_t spk_cap_ = 0;
  std::vector<int> perm_;
  int *order_ = nullptr;
  unsigned char *scratch_ = nullptr;   // this device's scratch array
  std::vector<double> cost_seen_;
  std::vector<Variant> variants_;
  int cur_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  bool timed_pending_ = false;
  double kt_ms_ = 0.0;
  long kt_n_ = 0;
  long macro_ = 0;
  double t_next_ = 0.0;
};

// ---------------------------------------------------------------------------
// Small cells (granule, Golgi, inferior olive), G lanes of one wavefront per
// cell, 64 / G cells per wave, each with its own FP64 controller: the
// reference's algorithm on a population of one cell (rkclow, step, rkcrho
// and the per-type drivers, solve_rkc.cu:562-1362), the same structure as
// the Purkinje body above.  Lane r of a cell holds compartments r, r + G, ...
// in registers; the cell's stage vector sits in the wave's slice of LDS for
// the cable stencil (the cell's CSR row as a neighbour list); norms are
// butterfly sums over the cell's G lanes.  Each cell takes its own step
// sequence, so control flow diverges between the cells of a wave but never
// inside one cell's lanes: the shuffles and LDS reads of a cell only touch
// lanes on the same path.  Stages and ion updates are FP64 (the models'
// per-compartment functions shared with the population-global engine).

constexpr int SMALL_NB = 4;      // cable neighbours per compartment
constexpr int SMALL_KMAX = 4;    // compartments per lane
constexpr int SMALL_LDS = 256;   // stage-buffer doubles per wave: cells per wave x compartments
enum SmallType { SM_GR = 0, SM_GO = 1, SM_IO = 2 };

struct SmallPop {
  int type, n, comp, G;
  IoState io;
  GoState go;
  GrState gr;
  const int *nb;        // [nc][SMALL_NB] cell-local neighbour index or -1
  const double *nbg;    // [nc][SMALL_NB] conductance to it [mS] (the CSR's off-diagonal)
  const double *invc;   // [nc] 1 / Cm
  double *bb;           // [nc] frozen current (the model state's b)
  double *y0, *f0, *vy, *ede, *vnew;
  double *ctl;          // [n][NCTL]
  const double *gj;     // [n] gap-junction current of compartment 0 over the interval, or null
  double atol, rtol, hmin, hmax, integ, interval;
  int strict;           // 1: ion update once sum_dt > integ (granule, olive); 0: >= (Golgi)
  int refresh_fn;
  int rec_comp;         // recorded / raster compartment
  double *rec;          // [nsteps][n] or null
  double thr;           // raster threshold
  double *vprev;        // [n] previous recorded value
  uint32_t *spk;        // [nsteps][spk_words] or null
  int spk_words;
};

__device__ inline double small_ion(const SmallPop &P, int gi, double vn, double dt, bool do_ion, bool with_iext) {
  if (P.type == SM_GR) return gr_ion_current(P.gr, gi, vn, dt, do_ion, with_iext);
  if (P.type == SM_GO) return go_ion_current(P.go, gi, vn, dt, do_ion, with_iext);
  return io_ion_current(P.io, gi, vn, dt, do_ion, with_iext);
}

// sum over the G lanes of one cell (G a power of two, groups aligned): every
// lane of the group ends with the same value (each level adds a pair in
// either order, and IEEE addition commutes)
__device__ inline double gsum(double x, int G) {
  for (int off = 1; off < G; off <<= 1) x += __shfl_xor(x, off);
  return x;
}

template <bool INIT>
__device__ void small_cells(const SmallPop &P, int c0, int nsteps, double *lds) {
  const int lane = threadIdx.x & 63;
  const int G = P.G, C = P.comp;
  const int slot = lane / G, r = lane - slot * G;
  const int cell = c0 + slot;
  if (cell >= P.n) return;             // the group has no cell (a population's last item)
  const int K = (C + G - 1) / G;       // <= SMALL_KMAX (checked on the host)
  double *cb[2] = {lds + slot * C, lds + SMALL_LDS + slot * C};
  const size_t base = (size_t)cell * C;
  double *ctl = P.ctl + (size_t)cell * NCTL;
  auto comp_of = [&](int k) { return r + k * G; };
  auto valid = [&](int k) { return k < K && r + k * G < C; };

  double y[SMALL_KMAX], F0[SMALL_KMAX], A[SMALL_KMAX], D1[SMALL_KMAX], D2[SMALL_KMAX], VY[SMALL_KMAX],
      bq[SMALL_KMAX];
#pragma unroll
  for (int k = 0; k < SMALL_KMAX; k++) {
    y[k] = F0[k] = VY[k] = bq[k] = A[k] = D1[k] = D2[k] = 0.0;
    if (valid(k)) {
      const size_t gi = base + comp_of(k);
      y[k] = P.y0[gi];
      F0[k] = P.f0[gi];
      VY[k] = P.vy[gi];
      bq[k] = P.bb[gi];
    }
  }
  // sum_nb g (x_nb - x_i) over the compartment's neighbours, x from LDS
  auto lap = [&](const double *x, int k, double xi) {
    const size_t gi = base + comp_of(k);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < SMALL_NB; q++) {
      const int j = P.nb[gi * SMALL_NB + q];
      if (j >= 0) s += P.nbg[gi * SMALL_NB + q] * (x[j] - xi);
    }
    return s;
  };
  auto put = [&](double *buf, const double *v) {
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++)
      if (valid(k)) buf[comp_of(k)] = v[k];
    __builtin_amdgcn_wave_barrier();
  };

  double absh = ctl[C_ABSH], sprad = ctl[C_SPRAD], naccpt = ctl[C_NACCPT], errold = ctl[C_ERROLD],
         hold = ctl[C_HOLD], lt = ctl[C_LT], tend = ctl[C_TEND], gt = ctl[C_GT], gt_old = ctl[C_GTOLD],
         sum_dt = ctl[C_SUMDT], mmax = ctl[C_MMAX];
  bool newspc = ctl[C_NEWSPC] != 0.0, jacatt = ctl[C_JACATT] != 0.0, last = ctl[C_LAST] != 0.0;
  double idid = ctl[C_IDID];
  double cnt[6];
#pragma unroll
  for (int q = 0; q < 6; q++) cnt[q] = ctl[C_STEPS + q];
  int status = (int)ctl[C_STATUS];
  const double at = P.atol, rt = P.rtol, hmin = P.hmin, hmax = P.hmax;
  bool a_valid = false;

  // A = F(y_n) = (-b + L y_n) / C
  auto compute_A = [&]() {
    put(cb[0], y);
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++)
      if (valid(k)) A[k] = (-bq[k] + lap(cb[0], k, y[k])) * P.invc[base + comp_of(k)];
    __builtin_amdgcn_wave_barrier();
    a_valid = true;
  };

  // the type's ion update (when do_ion) and frozen current at V = Y_{m-2};
  // the initial one (!do_ion) has neither stimulus nor junction current: the
  // reference initialises RKC before the output loop's first set_current
  auto ion_and_b = [&](double dt, bool do_ion) {
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++)
      if (valid(k)) {
        const int i = comp_of(k);
        double b = small_ion(P, (int)(base + i), VY[k], dt, do_ion, do_ion);
        if (i == 0 && P.gj && do_ion) b += P.gj[cell];
        P.bb[base + i] = b;
        bq[k] = b;
      }
    a_valid = false;
  };

  // rkcrho, solve_rkc.cu:562-754, on the perturbation delta = v - y_n (in D1)
  auto rkcrho = [&]() {
    const int itmax = 1000;
    double sprad_old[3] = {0.0, 0.0, 0.0};
    const double sqrtu = sqrt(UROUND), small = 1.0 / 0.125;
    double sy = 0.0, se = 0.0;
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++) {
      D1[k] = valid(k) ? P.ede[base + comp_of(k)] : 0.0;
      sy += y[k] * y[k];
      se += D1[k] * D1[k];
    }
    const double ynrm = sqrt(gsum(sy, G)), vnrm = sqrt(gsum(se, G));
    double dynrm;
    if (ynrm != 0.0 && vnrm != 0.0) {
      dynrm = ynrm * sqrtu;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) D1[k] = (dynrm / vnrm) * D1[k];
    } else if (ynrm != 0.0) {
      dynrm = ynrm * sqrtu;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) D1[k] = sqrtu * y[k];
    } else if (vnrm != 0.0) {
      dynrm = UROUND;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) D1[k] = (dynrm / vnrm) * D1[k] - y[k];
    } else {
      dynrm = UROUND;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) D1[k] = dynrm - y[k];
    }
    double sigma = 0.0, sigmal;
    for (int iter = 0; iter < itmax; iter++) {
      cnt[3] += 1.0;
      put(cb[0], D1);
      double sq = 0.0;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) {
        D2[k] = 0.0;
        if (valid(k)) {
          D2[k] = (A[k] - F0[k]) + lap(cb[0], k, D1[k]) * P.invc[base + comp_of(k)];
          sq += D2[k] * D2[k];
        }
      }
      __builtin_amdgcn_wave_barrier();
      const double dfnrm = sqrt(gsum(sq, G));
      sigmal = sigma;
      sigma = dfnrm / dynrm;
      sprad_old[0] = sprad_old[1];
      sprad_old[1] = sprad_old[2];
      sprad_old[2] = sprad;
      sprad = 1.2 * sigma;
      bool done = iter >= 1 && fabs(sigma - sigmal) <= fmax(sigma, small) * 0.001;
      if (!done && iter > itmax / 2 && fabs(sprad_old[2] - sprad_old[0]) < 0.00001 &&
          fabs(sprad_old[1] - sprad) < 0.00001) {
        sprad = (sprad + sprad_old[2]) / 2.0;
        done = true;
      }
      if (done) {
#pragma unroll
        for (int k = 0; k < SMALL_KMAX; k++)
          if (valid(k)) P.ede[base + comp_of(k)] = D1[k];
        return;
      }
      if (dfnrm != 0.0) {
#pragma unroll
        for (int k = 0; k < SMALL_KMAX; k++) D1[k] = (dynrm / dfnrm) * (-D2[k]);
      } else {
        const int idx = iter % C;
#pragma unroll
        for (int k = 0; k < SMALL_KMAX; k++)
          if (valid(k) && comp_of(k) == idx) D1[k] = -D1[k];
      }
    }
    idid = 6;
  };

  // rkclow, solve_rkc.cu:908-1096, on increments D_j = Y_j - Y_0
  auto rkclow = [&]() {
    int loop_num = 0;
    while (true) {
      if (++loop_num > 1000) return;
      if (!a_valid) compute_A();
      if (newspc) {
        rkcrho();
        jacatt = true;
      }
      last = false;
      if (1.1 * absh >= fabs(tend - lt)) {
        absh = fabs(tend - lt);
        last = true;
      }
      const double mr = sqrt(1.54 * absh * sprad + 1.0);
      int m = (mr < mmax) ? 1 + (int)mr : (int)mmax + 1;
      if (m > mmax) {
        m = (int)mmax;
        absh = (m * m - 1) / (1.54 * sprad);
        last = false;
      }
      double h = absh;
      if (h <= hmin) h = hmin;
      const double w0 = 1.0 + 2.0 / (13.0 * m * m);
      const double temp1 = w0 * w0 - 1.0, temp2 = sqrt(temp1);
      const double arg = m * log(w0 + temp2);
      const double w1 = sinh(arg) * temp1 / (cosh(arg) * m * temp2 - w0 * sinh(arg));
      double bjm1 = 1.0 / ((2.0 * w0) * (2.0 * w0)), bjm2 = bjm1;
      const double mus1 = w1 * bjm1;
      int cur = 0;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) {
        D2[k] = 0.0;
        D1[k] = (h * mus1) * F0[k];
        VY[k] = y[k];   // Y_{m-2} when m == 2
      }
      put(cb[cur], D1);
      double zjm1 = w0, zjm2 = 1.0, dzjm1 = 1.0, dzjm2 = 0.0, d2zjm1 = 0.0, d2zjm2 = 0.0;
      for (int j = 1; j < m; j++) {
        const double zj = 2.0 * w0 * zjm1 - zjm2;
        const double dzj = 2.0 * w0 * dzjm1 - dzjm2 + 2.0 * zjm1;
        const double d2zj = 2.0 * w0 * d2zjm1 - d2zjm2 + 4.0 * dzjm1;
        const double bj = d2zj / (dzj * dzj);
        const double ajm1 = 1.0 - zjm1 * bjm1;
        const double mu = 2.0 * w0 * bj / bjm1;
        const double nu = -bj / bjm2;
        const double hmus = h * (mu * w1 / w0);
        double Dn[SMALL_KMAX];
#pragma unroll
        for (int k = 0; k < SMALL_KMAX; k++) {
          Dn[k] = 0.0;
          if (valid(k)) {
            const double F = A[k] + lap(cb[cur], k, D1[k]) * P.invc[base + comp_of(k)];
            Dn[k] = mu * D1[k] + nu * D2[k] + hmus * (F - ajm1 * F0[k]);
            if (j == m - 1) VY[k] = y[k] + D2[k];   // Y_{m-2}
          }
        }
#pragma unroll
        for (int k = 0; k < SMALL_KMAX; k++) {
          D2[k] = D1[k];
          D1[k] = Dn[k];
        }
        put(cb[cur ^ 1], D1);
        cur ^= 1;
        bjm2 = bjm1; bjm1 = bj;
        zjm2 = zjm1; zjm1 = zj;
        dzjm2 = dzjm1; dzjm1 = dzj;
        d2zjm2 = d2zjm1; d2zjm1 = d2zj;
      }
      cnt[2] += m + 1;
      // f(Y_m) (into D2) and the error estimate, solve_rkc.cu:58-73
      double sq = 0.0;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) {
        const double Dm = D1[k];
        D2[k] = 0.0;
        if (valid(k)) {
          const double Fm = A[k] + lap(cb[cur], k, Dm) * P.invc[base + comp_of(k)];
          D2[k] = Fm;
          const double yn = y[k], yy = y[k] + Dm;
          const double wt = at + rt * fmax(fabs(yy), fabs(yn));
          const double est = 0.8 * (yn - yy) + 0.4 * h * (F0[k] + Fm);
          const double q = est / wt;
          sq += q * q;
        }
      }
      __builtin_amdgcn_wave_barrier();
      double err = sqrt(gsum(sq, G) / (double)C);
      if (!(err <= 1.0e300)) {
        idid = NONFINITE;
        return;
      }
      if (err > 1.0) {
        cnt[1] += 1.0;
        absh = 0.8 * absh / pow(err, 1.0 / 3.0);
        if (absh < hmin) {
          idid = 4;   // falls through to the accept branch, as the reference does
        } else {
          newspc = !jacatt;
          continue;
        }
      }
      cnt[0] += 1.0;
      naccpt += 1;
      lt += h;
      gt = lt;
      jacatt = false;
      newspc = false;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++) {
        y[k] = y[k] + D1[k];
        F0[k] = D2[k];
      }
      a_valid = false;
      double fac = 10.0, tp1, tp2;
      if (naccpt == 1) {
        tp2 = pow(err, 1.0 / 3.0);
        if (0.8 < fac * tp2) fac = 0.8 / tp2;
      } else {
        tp1 = 0.8 * absh * pow(errold, 1.0 / 3.0);
        tp2 = fabs(hold) * pow(err, 2.0 / 3.0);
        if (tp1 < fac * tp2) fac = tp1 / tp2;
      }
      absh = fmax(0.1, fac) * absh;
      absh = fmax(hmin, fmin(hmax, absh));
      errold = err;
      hold = h;
      idid = last ? 1 : 2;
      return;
    }
  };

  if constexpr (INIT) {
    // RKC_vec_initialize (solve_rkc.cu:756-791)
    ion_and_b(0.0, false);
    compute_A();
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++) {
      F0[k] = A[k];
      if (valid(k)) P.ede[base + comp_of(k)] = F0[k];
    }
    rkcrho();
    // set_initial_absh (solve_rkc.cu:489-560)
    double ab = hmax;
    if (sprad * ab > 1.0) ab = 1.0 / sprad;
    ab = fmax(ab, hmin);
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++) D1[k] = ab * F0[k];
    put(cb[0], D1);
    double sq = 0.0;
#pragma unroll
    for (int k = 0; k < SMALL_KMAX; k++)
      if (valid(k)) {
        const double F = A[k] + lap(cb[0], k, D1[k]) * P.invc[base + comp_of(k)];
        const double wt = at + rt * fabs(y[k]);
        const double d = (F - F0[k]) * (1.0 / wt);
        sq += d * d;
      }
    __builtin_amdgcn_wave_barrier();
    double est = gsum(sq, G);
    est = ab * sqrt(est / C);
    if (0.1 * ab < hmax * sqrt(est))
      ab = fmax(0.1 * ab / sqrt(est), hmin);
    else
      ab = hmax;
    absh = ab;
  } else {
    for (int s = 0; s < nsteps && status == 0; s++) {
      // the driver of xx_solve_by_rkc (solve_rkc.cu:1100-1362)
      auto ion_step = [&]() {
        cnt[4] += 1.0;
        ion_and_b(sum_dt, true);
        compute_A();
        if (P.refresh_fn) {
#pragma unroll
          for (int k = 0; k < SMALL_KMAX; k++) F0[k] = A[k];
        }
        sum_dt = 0.0;
      };
      if (sum_dt != 0.0) ion_step();
      while (true) {
        rkclow();
        sum_dt += gt - gt_old;
        gt_old = gt;
        if ((P.strict ? sum_dt > P.integ : sum_dt >= P.integ) && idid == 2.0) ion_step();
        if (idid == 2.0) continue;
        if (idid == 1.0) break;
        status = (int)idid;
        break;
      }
      tend += P.interval;
      cnt[5] += 1.0;
#pragma unroll
      for (int k = 0; k < SMALL_KMAX; k++)
        if (valid(k) && comp_of(k) == P.rec_comp) {
          const double v = y[k];
          if (P.rec) P.rec[(size_t)s * P.n + cell] = v;
          if (P.spk) {
            if (v > P.thr && P.vprev[cell] <= P.thr)
              atomicOr(&P.spk[(size_t)s * P.spk_words + (cell >> 5)], 1u << (cell & 31));
            P.vprev[cell] = v;
          }
        }
    }
  }

  // write back
#pragma unroll
  for (int k = 0; k < SMALL_KMAX; k++)
    if (valid(k)) {
      const size_t gi = base + comp_of(k);
      P.y0[gi] = y[k];
      P.f0[gi] = F0[k];
      P.vy[gi] = VY[k];
      P.vnew[gi] = y[k];
    }
  if (r == 0) {
    ctl[C_ABSH] = absh; ctl[C_SPRAD] = sprad; ctl[C_NACCPT] = naccpt; ctl[C_ERROLD] = errold;
    ctl[C_HOLD] = hold; ctl[C_LT] = lt; ctl[C_TEND] = tend; ctl[C_GT] = gt; ctl[C_GTOLD] = gt_old;
    ctl[C_SUMDT] = sum_dt; ctl[C_MMAX] = mmax;
    ctl[C_NEWSPC] = newspc; ctl[C_JACATT] = jacatt; ctl[C_LAST] = last; ctl[C_IDID] = idid;
    for (int q = 0; q < 6; q++) ctl[C_STEPS + q] = cnt[q];
    ctl[C_STATUS] = status;
  }
}

// One persistent launch over a mix of cell types: the Purkinje cells first,
// one workgroup per cell from their own work counter (the costliest first),
// then, wave by wave, the small-cell items (a population and its first
// cell; 64 / G cells each) from a second counter.  INIT runs the RKC
// initialisation instead of nsteps intervals.
struct MixArgs {
  CellArgs<float> pkj;
  int pkj_on;
  const SmallPop *pops;   // device array
  const int *item_pop, *item_cell;
  int nitems;
  int *small_queue;
};

template <int NT, int K, bool INIT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_mix(MixArgs m, int nsteps, double t0, int step0) {
  if (m.pkj_on) cell_body<float, NT, K, INIT, false>(m.pkj, nsteps, t0, step0);
  extern __shared__ __align__(16) unsigned char smem[];
  double *lds = (double *)smem + (size_t)(threadIdx.x >> 6) * 2 * SMALL_LDS;
  while (true) {
    int it = 0;
    if ((threadIdx.x & 63) == 0) it = atomicAdd(m.small_queue, 1);
    it = __shfl(it, 0);
    if (it >= m.nitems) break;
    const SmallPop *P = m.pops + m.item_pop[it];
    small_cells<INIT>(*P, m.item_cell[it], nsteps, lds);
  }
}

// gap-junction current of every inferior-olive cell for the next interval,
// from compartment 0's voltage at the interval's start: for each partner j
// in CSR order, g_ij (0.8 exp(-0.01 dV^2) + 0.2) dV with dV = V_i - V_j (the
// reference's nonlinear junction, gap.cu:149-160, with a per-pair
// conductance in place of G_C)
__global__ void __launch_bounds__(256) k_io_gj(const int *row, const int *col, const double *g, const double *vnew,
                                               int comp, double *gj, int n) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const double vi = vnew[(size_t)c * comp];
  double s = 0.0;
  for (int e = row[c]; e < row[c + 1]; e++) {
    const double dv = vi - vnew[(size_t)col[e] * comp];
    s += g[e] * (0.8 * exp(-0.01 * dv * dv) + 0.2) * dv;
  }
  gj[c] = s;
}

// fills a spike buffer with zeros (a kernel, so it may sit in a graph)
__global__ void k_zero_u32(uint32_t *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Batched engine over an uneven mix of cell types (host side): Purkinje cells
// on the per-cell mixed-precision body, granule / Golgi / olive cells on the
// small-cell lane-group body, all in one kernel per launch.  Olive cells
// exchange gap-junction currents once per interval, so with olive coupling
// every interval is one launch followed by the exchange; the launches of a
// chunk of intervals are captured once in a HIP graph and replayed.
class MixEngine : public Mix {
 public:
  static constexpr int T_GR = 0, T_GO = 1, T_PKJ = 2, T_IO = 3;
  static constexpr int NT = 512, KP = 4;

  explicit MixEngine(const MixSpec &S) : spec_(S) {
    create_stream(&st_);
    CB_HIP(hipEventCreate(&ev0_));
    CB_HIP(hipEventCreate(&ev1_));
    for (int t = 0; t < 4; t++) {
      n_[t] = S.pop[t].n;
      if (n_[t] < 0) throw Error("negative cell count");
      if (n_[t] > 0 && (int)S.pop[t].v0.size() != n_[t]) throw Error("a population needs one initial voltage per cell");
    }
    if (S.interval <= 0.0 || S.interval > 0.125) throw Error("interval must be in (0, 0.125] ms");
    // Purkinje cells: the per-cell mixed engine's state and initialisation
    if (n_[T_PKJ] > 0) {
      std::vector<MorphRow> rows;
      std::string err;
      if (!read_pkj_morph(S.pop[T_PKJ].morph, pkj::COMP, rows, err)) throw Error(err);
      PkjHost H;
      if (!build_pkj_host(rows, n_[T_PKJ], 1, nullptr, nullptr, S.pop[T_PKJ].v0.data(), H, err)) throw Error(err);
      SolverSpec sol;
      sol.atol = S.pkj_atol;
      sol.rtol = S.pkj_rtol;
      sol.interval = S.interval;
      if (S.pkj_stable) {
        sol.ion_interval = std::ldexp(1.0, -8);
        sol.refresh_fn = true;
        std::vector<double> rho;
        cable_spectral_radius(H, rho);
        sol.rho_floor.resize(rho.size());
        for (size_t c = 0; c < rho.size(); c++) sol.rho_floor[c] = 1.2 * rho[c];
      }
      StimSpec stim;
      stim.comp = S.pop[T_PKJ].stim_comp;
      stim.t0 = 0.0;
      stim.t1 = 1e300;
      stim.amp = S.pop[T_PKJ].amp;
      stim.amp.resize(n_[T_PKJ], 0.0);
      pkj_.reset(new CellPkj<float>(H, sol, stim, PoissonSpec()));
      if (!pkj_->select_mapping(NT)) throw Error("no 512-thread Purkinje mapping");
      pkj_->sync();
    }
    for (int t : {T_GR, T_GO, T_IO})
      if (n_[t] > 0) build_small(t);
    if (!S.io_gj.empty() && n_[T_IO] > 0) build_io_gj();
    // device copies of the small populations, the work items, the counter
    small_queue_ = dalloc<int>(1);
    bufs_.push_back(small_queue_);
    pops_dev_ = dalloc<SmallPop>(3);
    bufs_.push_back(pops_dev_);
    upload_pops();
    int dev = 0;
    CB_HIP(hipGetDevice(&dev));
    CB_HIP(hipDeviceGetAttribute(&ncu_, hipDeviceAttributeMultiprocessorCount, dev));
    lds_ = std::max(pkj_ ? pkj_->cell_lds_bytes() : 0, (size_t)(NT / 64) * 2 * SMALL_LDS * sizeof(double));
    // RKC_vec_initialize of every small cell
    if (!items_pop_.empty()) {
      MixArgs a = args(0, nullptr, nullptr);
      a.pkj_on = 0;
      k_zero_int<<<1, 1, 0, st_>>>(small_queue_);
      hipLaunchKernelGGL((k_mix<NT, KP, true>), dim3(grid()), dim3(NT), lds_, st_, a, 0, 0.0, 0);
      CB_HIP(hipGetLastError());
      CB_HIP(hipStreamSynchronize(st_));
      check_status();
    }
  }

  ~MixEngine() {
    (void)hipStreamSynchronize(st_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    for (void *p : bufs_) (void)hipFree(p);
    (void)hipFree(item_pop_dev_);
    (void)hipFree(item_cell_dev_);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(st_);
  }

  int ncell(int t) const override { return n_[t]; }
  int total_cells() const override { return n_[0] + n_[1] + n_[2] + n_[3]; }
  int words() const override { return words_total(); }
  int word_off(int t) const override {
    int w = 0;
    for (int u = 0; u < t; u++) w += (n_[u] + 31) / 32;
    return w;
  }
  int lanes(int t) const override { return t == T_PKJ ? NT : pop_[t].G; }

  // nint intervals; rec_out [nint][total cells] (type order GR, GO, PKJ,
  // IO), spk_out [nint][words()] bits; either may be null
  void run(int nint, double *rec_out, uint32_t *spk_out) override {
    if (nint < 0) throw Error("negative interval count");
    int done = 0;
    while (done < nint) {
      const int k = std::min(CHUNK, nint - done);
      ensure_chunk_buffers();
      launch_chunk(k, true);
      CB_HIP(hipStreamSynchronize(st_));
      collect_time();
      check_status();
      if (rec_out) {
        const int tot = total_cells();
        std::vector<double> tmp;
        for (int t = 0, off = 0; t < 4; off += n_[t], t++) {
          if (!n_[t]) continue;
          tmp.resize((size_t)k * n_[t]);
          CB_HIP(hipMemcpy(tmp.data(), rec_[t], sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
          for (int s = 0; s < k; s++)
            std::memcpy(rec_out + (size_t)(done + s) * tot + off, tmp.data() + (size_t)s * n_[t],
                        sizeof(double) * n_[t]);
        }
      }
      if (spk_out)
        CB_HIP(hipMemcpy(spk_out + (size_t)done * words_total(), spk_, sizeof(uint32_t) * (size_t)k * words_total(),
                         hipMemcpyDeviceToHost));
      intervals_ += k;
      done += k;
    }
  }

  Stats stats(int t) const override {
    Stats s;
    if (t == T_PKJ) {
      if (pkj_) s = pkj_->stats();
      s.macro_steps = intervals_;
      return s;
    }
    if (!n_[t]) return s;
    std::vector<double> ctl((size_t)n_[t] * NCTL);
    CB_HIP(hipMemcpy(ctl.data(), pop_[t].ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < n_[t]; c++) {
      const double *q = &ctl[(size_t)c * NCTL];
      s.steps += (long)q[C_STEPS];
      s.rejects += (long)q[C_REJECTS];
      s.stages += (long)q[C_STAGES];
      s.rho_iters += (long)q[C_RHO];
      s.ion_updates += (long)q[C_IONS];
    }
    s.macro_steps = intervals_;
    return s;
  }

  // Lanes per cell of each small population (cells per wavefront = 64 / G):
  // every candidate runs the same trial from the same state (restored after
  // each), one population at a time, and the fastest is kept.
  void autotune(int trial) override {
    if (trial <= 0) return;
    Snapshot snap = snapshot();
    for (int t : {T_GR, T_IO, T_GO}) {
      if (!n_[t]) continue;
      std::vector<int> cand;
      for (int G = 1; G <= 64; G *= 2)
        if ((pop_[t].comp + G - 1) / G <= SMALL_KMAX && (64 / G) * pop_[t].comp <= SMALL_LDS) cand.push_back(G);
      double best = 1e30;
      int bestG = pop_[t].G;
      for (int G : cand) {
        set_lanes(t, G);
        launch_chunk(std::min(trial, CHUNK), false);   // warm
        restore(snap);
        CB_HIP(hipStreamSynchronize(st_));
        const auto t0 = std::chrono::steady_clock::now();
        launch_chunk(std::min(trial, CHUNK), false);
        CB_HIP(hipStreamSynchronize(st_));
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        restore(snap);
        CB_HIP(hipStreamSynchronize(st_));
        tune_.push_back({t, G, sec});
        if (sec < best) {
          best = sec;
          bestG = G;
        }
      }
      set_lanes(t, bestG);
    }
    release(snap);
    timed_pending_ = false;
  }
  struct TuneRec {
    int type, lanes;
    double sec;
  };
  std::vector<TuneRec> tune_;

  void kernel_time(double *ms, long *launches, bool reset) override {
    *ms = kt_ms_;
    *launches = kt_n_;
    if (reset) {
      kt_ms_ = 0.0;
      kt_n_ = 0;
    }
  }

 private:
  static constexpr int CHUNK = 64;   // intervals per captured graph

  struct Small {
    std::vector<void *> mut;         // device buffers a run changes, with sizes below
    std::vector<size_t> mut_bytes;
  };

  template <typename U>
  U *put(const std::vector<U> &h) {
    U *d = dalloc<U>(h.size());
    h2d(d, h);
    bufs_.push_back((void *)d);
    return d;
  }
  template <typename U>
  U *zeros(size_t n) {
    U *d = dalloc<U>(n);
    CB_HIP(hipMemset(d, 0, sizeof(U) * (n ? n : 1)));
    bufs_.push_back((void *)d);
    return d;
  }
  template <typename U>
  U *mut(const std::vector<U> &h, int t) {
    U *d = put(h);
    small_[t].mut.push_back((void *)d);
    small_[t].mut_bytes.push_back(sizeof(U) * h.size());
    return d;
  }

  // stencil, RKC state and controller records common to the small types
  template <typename Host>
  void common_small(int t, const Host &H, double atol, double rtol, double hmin, bool strict) {
    SmallPop &P = pop_[t];
    const int n = H.n, C = H.comp, nc = H.nc;
    P.n = n;
    P.comp = C;
    std::vector<int> nb((size_t)nc * SMALL_NB, -1);
    std::vector<double> nbg((size_t)nc * SMALL_NB, 0.0), invc(nc);
    for (int gi = 0; gi < nc; gi++) {
      const int cell = gi / C;
      int q = 0;
      for (int e = H.row[gi]; e < H.row[gi + 1]; e++) {
        if (H.col[e] == gi) continue;
        if (H.col[e] / C != cell) throw Error("intercell cable entries are not provided by the batched engine");
        if (q >= SMALL_NB) throw Error("a compartment with more than 4 cable neighbours");
        nb[(size_t)gi * SMALL_NB + q] = H.col[e] - cell * C;
        nbg[(size_t)gi * SMALL_NB + q] = H.val[e];
        q++;
      }
      invc[gi] = 1.0 / H.cm[gi];
    }
    P.nb = put(nb);
    P.nbg = put(nbg);
    P.invc = put(invc);
    P.y0 = mut(H.v, t);
    P.f0 = mut(std::vector<double>(nc, 0.0), t);
    P.vy = mut(H.v, t);
    P.ede = mut(std::vector<double>(nc, 0.0), t);
    P.vnew = mut(H.v, t);
    std::vector<double> ctl((size_t)n * NCTL, 0.0);
    for (int c = 0; c < n; c++) {
      double *q = &ctl[(size_t)c * NCTL];
      q[C_TEND] = spec_.interval;
      q[C_LAST] = 1;
      q[C_NEWSPC] = 1;
      q[C_MMAX] = nint_small(std::sqrt(rtol / (10.0 * UROUND)));
    }
    P.ctl = mut(ctl, t);
    P.atol = atol;
    P.rtol = rtol;
    P.hmin = hmin;
    P.hmax = spec_.interval;
    P.interval = spec_.interval;
    P.integ = std::pow(2.0, std::log10(atol));   // solve_rkc.cu:1141, 1205, 1340
    P.strict = strict ? 1 : 0;
    P.refresh_fn = 0;
    const MixPopSpec &ps = spec_.pop[t];
    if (ps.rec_comp < 0 || ps.rec_comp >= C || ps.stim_comp < 0 || ps.stim_comp >= C)
      throw Error("recorded or stimulated compartment out of range");
    P.rec_comp = ps.rec_comp;
    P.thr = ps.thr;
    P.vprev = mut(std::vector<double>(n, 100.0), t);   // main.cu:162-169
    P.gj = nullptr;
    P.G = 1;
    while ((C + P.G - 1) / P.G > SMALL_KMAX) P.G *= 2;
    if (P.G > 64) throw Error("cell too large for the small-cell mapping (at most 256 compartments)");
    if ((64 / P.G) * C > SMALL_LDS) throw Error("cells too large for the wave's stage buffer");
  }

  std::vector<double> stim_vec(int t, int nc, int C) {
    std::vector<double> iext(nc, 0.0);
    const MixPopSpec &ps = spec_.pop[t];
    for (int c = 0; c < ps.n && c < (int)ps.amp.size(); c++) iext[(size_t)c * C + ps.stim_comp] = ps.amp[c];
    return iext;
  }

  void build_small(int t) {
    const MixPopSpec &ps = spec_.pop[t];
    std::string err;
    SmallPop &P = pop_[t];
    P.type = t == T_GR ? SM_GR : t == T_GO ? SM_GO : SM_IO;
    if (t == T_GR) {
      GrHost H;
      if (!build_gr_host(ps.morph, ps.n, 1, ps.v0.data(), H, err, ps.comp > 0 ? ps.comp : gr::COMP)) throw Error(err);
      common_small(t, H, 1.0e-9, 1.0e-9, std::pow(2.0, -12.0), true);   // TOL_GR, H_MIN_GR (solve_rkc.cuh:24, 28)
      GrState &S = P.gr;
      S = GrState{};
      S.v = mut(H.v, t);
      S.iext = put(stim_vec(t, H.nc, H.comp));
      S.ca = put(H.ca);
      S.compart = put(H.compart);
      S.cond = put(H.cond);
      S.gate = mut(H.gate, t);
      const gr::NaTable tab = gr::na_table();
      S.na = put(std::vector<gr::NaType>(tab.t, tab.t + gr::NTYPE));
      S.b = mut(std::vector<double>(H.nc, 0.0), t);
      S.nc = H.nc;
      P.bb = S.b;
    } else if (t == T_GO) {
      GoHost H;
      if (!build_go_host(ps.morph, ps.n, 1, ps.v0.data(), H, err, go::COMP, false)) throw Error(err);
      common_small(t, H, 1.0e-6, 1.0e-6, std::pow(2.0, -10.0), false);  // TOL_GO, H_MIN_GO
      GoState &S = P.go;
      S = GoState{};
      S.v = mut(H.v, t);
      S.ca = mut(H.ca, t);
      S.ca2 = mut(H.ca2, t);
      S.rev = mut(H.rev, t);
      S.ca_old = mut(H.ca_old, t);
      S.iext = put(stim_vec(t, H.nc, H.comp));
      S.area = put(H.area);
      S.cond = put(H.cond);
      S.gate = mut(H.gate, t);
      S.b = mut(std::vector<double>(H.nc, 0.0), t);
      S.nc = H.nc;
      P.bb = S.b;
    } else {
      IoHost H;
      if (!build_io_host(ps.morph, ps.n, 1, ps.v0.data(), H, err)) throw Error(err);
      common_small(t, H, 1.0e-6, 1.0e-6, std::pow(2.0, -10.0), true);   // TOL_IO, H_MIN_IO
      IoState &S = P.io;
      S = IoState{};
      S.v = mut(H.v, t);
      S.ca = mut(H.ca, t);
      S.iext = put(stim_vec(t, H.nc, H.comp));
      S.cond = put(H.cond);
      S.gate = mut(H.gate, t);
      S.compart = put(H.compart);
      S.b = mut(std::vector<double>(H.nc, 0.0), t);
      S.nc = H.nc;
      S.n = H.n;
      P.bb = S.b;
    }
  }

  // the olive pair list as a CSR by cell, both directions, with each pair's
  // conductance per unit dendritic membrane: g [nS] 1e-6 / (pi d L) of
  // compartment 0 in IO.txt (the model's currents are per cm^2)
  void build_io_gj() {
    std::FILE *f = std::fopen(spec_.pop[T_IO].morph.c_str(), "r");
    if (!f) throw Error("no such file " + spec_.pop[T_IO].morph);
    double area0 = 0.0;
    for (int k = 0; k < io::COMP; k++) {
      int i1, i2, i3;
      double d, len, cm;
      if (std::fscanf(f, "%d %d %lf %lf %lf %d", &i1, &i2, &d, &len, &cm, &i3) != 6) {
        std::fclose(f);
        throw Error("IO_PARAM_FILE_READING_ERROR in " + spec_.pop[T_IO].morph);
      }
      if (i1 == 0) area0 = M_PI * d * 1e-4 * len * 1e-4;
    }
    std::fclose(f);
    if (!(area0 > 0.0)) throw Error("olive compartment 0 has no membrane area");
    std::FILE *g = std::fopen(spec_.io_gj.c_str(), "r");
    if (!g) throw Error("no such file " + spec_.io_gj);
    const int n = n_[T_IO];
    std::vector<std::vector<std::pair<int, double>>> adj(n);
    int a, b;
    double gns;
    while (std::fscanf(g, " %d , %d , %lf", &a, &b, &gns) == 3) {
      if (a < 0 || b < 0 || a >= n || b >= n || a == b) {
        std::fclose(g);
        throw Error("gap-junction pair out of range in " + spec_.io_gj);
      }
      const double gpa = gns * 1e-6 / area0;
      adj[a].push_back({b, gpa});
      adj[b].push_back({a, gpa});
    }
    std::fclose(g);
    std::vector<int> row(n + 1, 0), col;
    std::vector<double> gv;
    for (int c = 0; c < n; c++) {
      for (auto &e : adj[c]) {
        col.push_back(e.first);
        gv.push_back(e.second);
      }
      row[c + 1] = (int)col.size();
    }
    io_row_ = put(row);
    io_col_ = put(col.empty() ? std::vector<int>(1, 0) : col);
    io_g_ = put(gv.empty() ? std::vector<double>(1, 0.0) : gv);
    pop_[T_IO].gj = mut(std::vector<double>(n, 0.0), T_IO);
    io_gj_on_ = true;
  }

  void set_lanes(int t, int G) {
    pop_[t].G = G;
    upload_pops();
    graph_ready_ = false;
  }

  // the costliest items first: Golgi cells, then granule, then olive
  void upload_pops() {
    SmallPop host[3];
    const int order[3] = {T_GO, T_GR, T_IO};
    std::vector<int> ip, ic;
    int slot = 0;
    for (int t : order) {
      if (!n_[t]) continue;
      host[slot] = pop_[t];
      const int W = 64 / pop_[t].G;
      for (int c0 = 0; c0 < n_[t]; c0 += W) {
        ip.push_back(slot);
        ic.push_back(c0);
      }
      pop_slot_[t] = slot++;
    }
    CB_HIP(hipMemcpy(pops_dev_, host, sizeof(SmallPop) * 3, hipMemcpyHostToDevice));
    upload_tables();
    graph_ready_ = false;
    items_pop_ = ip;
    if (item_pop_dev_) {
      CB_HIP(hipStreamSynchronize(st_));
      (void)hipFree(item_pop_dev_);
      (void)hipFree(item_cell_dev_);
    }
    item_pop_dev_ = dalloc<int>(std::max<size_t>(1, ip.size()));
    item_cell_dev_ = dalloc<int>(std::max<size_t>(1, ic.size()));
    if (!ip.empty()) {
      CB_HIP(hipMemcpy(item_pop_dev_, ip.data(), sizeof(int) * ip.size(), hipMemcpyHostToDevice));
      CB_HIP(hipMemcpy(item_cell_dev_, ic.data(), sizeof(int) * ic.size(), hipMemcpyHostToDevice));
    }
  }

  MixArgs args(int s, double *const *rec, uint32_t *spk) {
    MixArgs a;
    a.pkj_on = pkj_ ? 1 : 0;
    if (pkj_) {
      a.pkj = pkj_->args();
      a.pkj.rec = rec ? rec[T_PKJ] + (size_t)s * n_[T_PKJ] : nullptr;
      a.pkj.rec_orig = spec_.pop[T_PKJ].rec_comp;
      a.pkj.thr = spec_.pop[T_PKJ].thr;
      a.pkj.vprev = pkj_vprev_;
      a.pkj.spk = spk ? spk + (size_t)s * words_total() + word_off(T_PKJ) : nullptr;
      a.pkj.spk_words = words_total();
    }
    a.pops = pops_dev_;
    a.item_pop = item_pop_dev_;
    a.item_cell = item_cell_dev_;
    a.nitems = (int)items_pop_.size();
    a.small_queue = small_queue_;
    return a;
  }

  // the recording and spike pointers of one launch: the small populations
  // take theirs from the device copy, so each interval's launch gets its own
  // copy of the population table
  void ensure_chunk_buffers() {
    if (spk_) return;
    for (int t = 0; t < 4; t++)
      if (n_[t]) rec_[t] = zeros<double>((size_t)CHUNK * n_[t]);
    spk_ = zeros<uint32_t>((size_t)CHUNK * words_total());
    if (pkj_) pkj_vprev_ = put(std::vector<double>(n_[T_PKJ], 100.0));
    pops_iv_ = dalloc<SmallPop>((size_t)3 * CHUNK);
    bufs_.push_back(pops_iv_);
    upload_tables();
    graph_ready_ = false;
  }

  int words_total() const { return word_off(4); }

  // the population tables of the CHUNK launches of a chunk, each with its
  // interval's recording and spike pointers (uploaded once, so a captured
  // graph holds kernels only)
  void upload_tables() {
    if (!spk_) return;
    std::vector<SmallPop> tabs((size_t)3 * CHUNK);
    for (int l = 0; l < CHUNK; l++)
      for (int t : {T_GR, T_GO, T_IO}) {
        if (!n_[t]) continue;
        SmallPop P = pop_[t];
        P.rec = rec_[t] + (size_t)l * n_[t];
        P.spk = spk_ + (size_t)l * words_total() + word_off(t);
        P.spk_words = words_total();
        tabs[(size_t)l * 3 + pop_slot_[t]] = P;
      }
    CB_HIP(hipMemcpy(pops_iv_, tabs.data(), sizeof(SmallPop) * tabs.size(), hipMemcpyHostToDevice));
  }

  // one chunk of k intervals: per interval (with olive coupling) the gap
  // currents, the two work counters cleared by a kernel and one mixed
  // launch; without coupling one launch for the whole chunk.  With record,
  // the launches write the chunk's recording and spike buffers.  The mix
  // has no time-dependent input (constant currents, no Poisson drive), so a
  // chunk's launches are the same for every chunk and can be replayed.
  void enqueue(int k, bool record) {
    const bool per = io_gj_on_;
    const int nl = per ? k : 1;
    if (record)
      k_zero_u32<<<(unsigned)(((size_t)k * words_total() + 255) / 256), 256, 0, st_>>>(spk_, (size_t)k * words_total());
    for (int l = 0; l < nl; l++) {
      if (per)
        k_io_gj<<<(unsigned)((n_[T_IO] + 255) / 256), 256, 0, st_>>>(io_row_, io_col_, io_g_, pop_[T_IO].vnew,
                                                                     io::COMP, (double *)pop_[T_IO].gj, n_[T_IO]);
      MixArgs a = args(l, record ? rec_ : nullptr, record ? spk_ : nullptr);
      a.pops = record ? pops_iv_ + (size_t)l * 3 : pops_dev_;
      if (pkj_) k_zero_int<<<1, 1, 0, st_>>>(a.pkj.queue);
      k_zero_int<<<1, 1, 0, st_>>>(small_queue_);
      hipLaunchKernelGGL((k_mix<NT, KP, false>), dim3(grid()), dim3(NT), lds_, st_, a, per ? 1 : k, 0.0, 0);
      CB_HIP(hipGetLastError());
    }
  }

  void launch_chunk(int k, bool record) {
    if (record && k == CHUNK && use_graph_) {
      if (!graph_ready_) {
        if (graph_exec_) CB_HIP(hipGraphExecDestroy(graph_exec_));
        graph_exec_ = nullptr;
        hipGraph_t g = nullptr;
        CB_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        enqueue(k, record);
        CB_HIP(hipStreamEndCapture(st_, &g));
        CB_HIP(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
        CB_HIP(hipGraphDestroy(g));
        graph_ready_ = true;
      }
      CB_HIP(hipEventRecord(ev0_, st_));
      CB_HIP(hipGraphLaunch(graph_exec_, st_));
      CB_HIP(hipEventRecord(ev1_, st_));
    } else {
      CB_HIP(hipEventRecord(ev0_, st_));
      enqueue(k, record);
      CB_HIP(hipEventRecord(ev1_, st_));
    }
    timed_pending_ = true;
  }

 public:
  // false: every chunk is enqueued launch by launch instead of replaying the
  // captured graph (the same kernels and arguments)
  void use_graph(bool on) override { use_graph_ = on; }

 private:
  unsigned grid() const {
    const int per_cu = std::max(1, std::min(2, (int)(160 * 1024 / (lds_ + 1024))));
    return (unsigned)(ncu_ * per_cu);
  }

  void collect_time() {
    if (!timed_pending_) return;
    float ms = 0.0f;
    CB_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kt_ms_ += ms;
    kt_n_ += 1;
    timed_pending_ = false;
  }

  void check_status() {
    if (pkj_) {
      std::vector<double> ctl((size_t)n_[T_PKJ] * NCTL);
      CB_HIP(hipMemcpy(ctl.data(), pkj_->args().ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost));
      for (int c = 0; c < n_[T_PKJ]; c++) {
        const int s = (int)ctl[(size_t)c * NCTL + C_STATUS];
        if (s) throw Error("RKC failed in Purkinje cell " + std::to_string(c) + ": status " + std::to_string(s), s);
      }
    }
    const char *name[4] = {"granule", "Golgi", "Purkinje", "olive"};
    for (int t : {T_GR, T_GO, T_IO}) {
      if (!n_[t]) continue;
      std::vector<double> ctl((size_t)n_[t] * NCTL);
      CB_HIP(hipMemcpy(ctl.data(), pop_[t].ctl, sizeof(double) * ctl.size(), hipMemcpyDeviceToHost));
      for (int c = 0; c < n_[t]; c++) {
        const int s = (int)ctl[(size_t)c * NCTL + C_STATUS];
        if (s == NONFINITE)
          throw Error(std::string("RKC: non-finite state in ") + name[t] + " cell " + std::to_string(c), s);
        if (s) throw Error(std::string("RKC failed in ") + name[t] + " cell " + std::to_string(c) + ": idid = " +
                               std::to_string(s), s);
      }
    }
  }

  struct Snapshot {
    std::vector<std::pair<void *, std::pair<void *, size_t>>> saved;
  };
  std::vector<std::pair<void *, size_t>> state_buffers() {
    std::vector<std::pair<void *, size_t>> v;
    if (pkj_) v = pkj_->state_buffers();
    for (int t : {T_GR, T_GO, T_IO})
      for (size_t q = 0; q < small_[t].mut.size(); q++) v.push_back({small_[t].mut[q], small_[t].mut_bytes[q]});
    return v;
  }
  Snapshot snapshot() {
    Snapshot s;
    for (auto &m : state_buffers()) {
      void *p = nullptr;
      CB_HIP(hipMalloc(&p, m.second));
      CB_HIP(hipMemcpyAsync(p, m.first, m.second, hipMemcpyDeviceToDevice, st_));
      s.saved.push_back({m.first, {p, m.second}});
    }
    return s;
  }
  void restore(const Snapshot &s) {
    for (auto &e : s.saved)
      CB_HIP(hipMemcpyAsync(e.first, e.second.first, e.second.second, hipMemcpyDeviceToDevice, st_));
  }
  void release(Snapshot &s) {
    CB_HIP(hipStreamSynchronize(st_));
    for (auto &e : s.saved) (void)hipFree(e.second.first);
    s.saved.clear();
  }
  static int nint_small(double x) {
    if (x - (double)(int)x < 0.5) return (int)(double)(int)x;
    return (int)(double)(int)(x + 0.9);
  }

  MixSpec spec_;
  int n_[4] = {0, 0, 0, 0};
  std::unique_ptr<CellPkj<float>> pkj_;
  SmallPop pop_[4] = {};
  Small small_[4];
  int pop_slot_[4] = {0, 0, 0, 0};
  SmallPop *pops_dev_ = nullptr, *pops_iv_ = nullptr;
  std::vector<int> items_pop_;
  int *item_pop_dev_ = nullptr, *item_cell_dev_ = nullptr;
  int *small_queue_ = nullptr;
  const int *io_row_ = nullptr, *io_col_ = nullptr;
  const double *io_g_ = nullptr;
  bool io_gj_on_ = false;
  double *rec_[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *spk_ = nullptr;
  double *pkj_vprev_ = nullptr;
  std::vector<void *> bufs_;
  hipStream_t st_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  bool graph_ready_ = false;
  bool use_graph_ = true;
  bool timed_pending_ = false;
  double kt_ms_ = 0.0;
  long kt_n_ = 0;
  long intervals_ = 0;
  int ncu_ = 0;
  size_t lds_ = 0;
};


}  // namespace

Mix *make_mix(const MixSpec &spec) { return new MixEngine(spec); }

Engine *make_cell_pkj(const PkjHost &host, const SolverSpec &sol, const StimSpec &stim, const PoissonSpec &syn,
                      bool mixed) {
  if (mixed) return new CellPkj<float>(host, sol, stim, syn);
  return new CellPkj<double>(host, sol, stim, syn);
}

}  // namespace cbrkc
