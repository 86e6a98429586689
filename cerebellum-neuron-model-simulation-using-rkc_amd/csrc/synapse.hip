// Synapse connection rules (host) and trace updates (device), syn.cu.
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.hpp"
#include "philox.hpp"
#include "synapse.hpp"

namespace cbrkc {
namespace {

constexpr int GR_COMP = 578, GO_COMP = 131, GO_COMP_DEND = 10, PKJ_COMP = 1600, N_SYN_PER_GRPKJ = 25;
constexpr double RKC_DT = 0.125;                       // param.h:12
constexpr double S_STIMULI = 900.0, E_STIMULI = 1000.0;  // param.h:8-9
// syn.cuh:13-55
constexpr double G_MFGR_AMPA = 0.18e-9, G_MFGR_NMDA = 0.025e-9, E_MFGR = 0.0;
constexpr double G_GRGO_AMPA = 45.5e-9, G_GRGO_NMDA = 30.0e-9, E_GRGO = 0.0;
constexpr double G_GOGR_GABA = 0.028e-9, E_GOGR = -82.0;
constexpr double G_GRPKJ_AMPA = 0.7e-9, E_GRPKJ = 0.0;
constexpr double G_MLIPKJ_GABA = 1.0e-9, E_MLIPKJ = -75.0;
enum { ST_MF = 0, ST_GRPKJ_JITTER = 1, ST_MLI = 2 };   // Philox streams

// uniform in (0, 1] from the first word of Philox(slot, step, stream, 0)
__host__ __device__ inline double philox_u(uint32_t k0, uint32_t k1, uint32_t slot, uint32_t step,
                                           uint32_t stream) {
  const u32x4 r = philox4x32_10(slot, step, stream, 0u, k0, k1);
  return ((double)r.v[0] + 1.0) * (1.0 / 4294967296.0);
}

__global__ void __launch_bounds__(256) k_syn_update(SynDev S, const double *__restrict__ pv, double t,
                                                    uint32_t step) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= S.n) return;
  switch (S.kind) {
    case SYN_MFGR: {   // mfgr_update, syn.cu:63-102
      double fr_mf = 5.0;
      (S_STIMULI <= t && t < E_STIMULI) ? (fr_mf = 60.0) : (fr_mf = 5.0);
      const double f = philox_u(S.k0, S.k1, (uint32_t)id, step, ST_MF);
      const int firing_flag = (fr_mf * RKC_DT * 0.001 > f) ? 1 : 0;
      const double a = S.x0[id] * S.d0 + firing_flag;
      S.x0[id] = a;
      const double nm = a * S.d1 + firing_flag;   // fed from the updated AMPA trace (syn.cu:96)
      S.x1[id] = nm;
      S.val[id] = S.weight[id] * (G_MFGR_AMPA * a + G_MFGR_NMDA * nm);
      break;
    }
    case SYN_GRGO: {   // grgo_update, syn.cu:263-288
      const double pre_comp_v = pv[S.pre[id]];
      const int firing_flag = (pre_comp_v > 0.0 && S.old_v[id] < 0.0) ? 1 : 0;
      const double a = S.x0[id] * S.d0 + firing_flag;
      const double n1 = S.x1[id] * S.d1 + firing_flag;
      const double n2 = S.x2[id] * S.d2 + firing_flag;
      S.x0[id] = a;
      S.x1[id] = n1;
      S.x2[id] = n2;
      S.val[id] = S.weight[id] * (G_GRGO_AMPA * a + G_GRGO_NMDA * (0.33 * n1 + 0.67 * n2));
      S.old_v[id] = pre_comp_v;
      break;
    }
    case SYN_GOGR: {   // gogr_update, syn.cu:412-437
      const double pre_comp_v = pv[S.pre[id]];
      const int firing_flag = (pre_comp_v > 0.0 && S.old_v[id] < 0.0) ? 1 : 0;
      const double a = S.x0[id] * S.d0 + firing_flag;
      const double b = S.x1[id] * S.d1 + firing_flag;
      S.x0[id] = a;
      S.x1[id] = b;
      S.val[id] = S.weight[id] * G_GOGR_GABA * (0.43 * a + 0.57 * b);
      S.old_v[id] = pre_comp_v;
      break;
    }
    case SYN_GRPKJ: {  // grpkj_update, syn.cu:646-668
      const double pre_comp_v = pv[S.pre[id]];
      const int firing_flag = (pre_comp_v > 0.0 && S.old_v[id] < 0.0) ? 1 : 0;
      const double a = S.x0[id] * S.d0 + firing_flag;
      S.x0[id] = a;
      S.val[id] = a * S.weight[id] * G_GRPKJ_AMPA;
      S.old_v[id] = pre_comp_v;
      break;
    }
    case SYN_MLIPKJ: { // mlipkj_update, syn.cu:761-788 (the drivers never call it)
      const double fr_mli = 30.0;
      const double f = philox_u(S.k0, S.k1, (uint32_t)id, step, ST_MLI);
      const int firing_flag = (fr_mli * 0.125 * 0.001 > f) ? 1 : 0;
      const double a = S.x0[id] * 0.98757780049 + firing_flag;
      S.x0[id] = a;
      S.val[id] = S.weight[id] * (G_MLIPKJ_GABA * a);
      break;
    }
  }
}

__global__ void __launch_bounds__(256) k_syn_macro(SynIn S, const double *__restrict__ y, double *__restrict__ s,
                                                   int nc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nc) s[i] = syn_current(S, i, y[i]);
}

double gr_x(int i, int nx) { return (int)(i % nx) * 16.0; }
double gr_y(int i, int nx) { return (int)(i / nx) * 16.0; }
double c32_x(int i, int nx) { return (int)(i % nx) * 32.0 + 8.0; }
double c32_y(int i, int nx) { return (int)(i / nx) * 32.0 + 8.0; }

}  // namespace

void build_syn(int kind, int nx_a, int ny_a, int nx_b, int ny_b, const double *x1600, uint64_t seed,
               SynHost &out) {
  out = SynHost();
  out.kind = kind;
  const int n_a = nx_a * ny_a, n_b = nx_b * ny_b;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  auto push = [&](int pre, int post, double w) {
    out.pre.push_back(pre);
    out.post.push_back(post);
    out.weight.push_back(w);
  };
  switch (kind) {
    case SYN_MFGR:   // mfgr_create / mfgr_initialize (syn.cu:42-128)
      out.E = E_MFGR;
      out.decay[0] = std::exp(-RKC_DT / 1.2);
      out.decay[1] = std::exp(-RKC_DT / 52.0);
      for (int id = 0; id < n_a; id++)
        for (int j = 0; j < 4; j++) push(id, j + id * GR_COMP, 40000.0 / 4.0);
      break;
    case SYN_GRGO: {  // grgo_create / grgo_initialize (syn.cu:144-261); a = GR, b = GO
      out.E = E_GRGO;
      out.decay[0] = std::exp(-RKC_DT / 1.5);
      out.decay[1] = std::exp(-RKC_DT / 31.0);
      out.decay[2] = std::exp(-RKC_DT / 170.0);
      if (n_a == 0 || n_b == 0) break;
      std::vector<int> num_go(n_b, 0);
      std::vector<std::vector<int>> label(n_a);
      for (int i_gr = 0; i_gr < n_a; i_gr++)
        for (int i_go = 0; i_go < n_b; i_go++)
          if (std::fabs(c32_x(i_go, nx_b) - gr_x(i_gr, nx_a)) < 16.0 &&
              std::fabs(c32_y(i_go, nx_b) - gr_y(i_gr, nx_a)) < 220.0) {
            label[i_gr].push_back(i_go);
            num_go[i_go]++;
          }
      for (int l_n_gr = 0; l_n_gr < n_a; l_n_gr++)
        for (int l_n_go : label[l_n_gr]) {
          const int iy_gr = l_n_gr / nx_a, iy_go = l_n_go / nx_b;
          const float diff_grgo = 8.0 + iy_go * 32.0 - iy_gr * 16.0;
          const int gr_ax_comp = 250 + 77 - (int)(diff_grgo / 10.0);
          push(gr_ax_comp + GR_COMP * l_n_gr, 0 + GO_COMP * l_n_go, 64.0 / 4.0 / (num_go[l_n_go] * 1.0 / 4.0));
        }
      break;
    }
    case SYN_GOGR: {  // gogr_create / gogr_initialize3 (syn.cu:292-410); a = GO, b = GR
      out.E = E_GOGR;
      out.decay[0] = std::exp(-RKC_DT / 7.0);
      out.decay[1] = std::exp(-RKC_DT / 59.0);
      if (n_a == 0 || n_b == 0) break;
      std::vector<int> num_gr(n_b, 0);
      std::vector<std::vector<int>> label(n_a);
      for (int i_go = 0; i_go < n_a; i_go++)
        for (int i_gr = 0; i_gr < n_b; i_gr++)
          if (std::fabs(c32_x(i_go, nx_a) - gr_x(i_gr, nx_b)) < 32.0 &&
              std::fabs(c32_y(i_go, nx_a) - gr_y(i_gr, nx_b)) < 32.0) {
            label[i_go].push_back(i_gr);
            num_gr[i_gr]++;
          }
      for (int l_n_go = 0; l_n_go < n_a; l_n_go++)
        for (int l_n_gr : label[l_n_go])
          push((GO_COMP_DEND * 3) + GO_COMP * l_n_go, 16 + GR_COMP * l_n_gr,
               160000.0 * 4.0 / (num_gr[l_n_gr] * 4.0));
      break;
    }
    case SYN_GRPKJ: {  // grpkj_create / grpkj_initialize (syn.cu:440-644); a = GR, b = PKJ
      out.E = E_GRPKJ;
      out.decay[0] = std::exp(-RKC_DT / 6.0);
      if (n_a == 0 || n_b == 0) break;
      std::vector<int> num_pkj(n_b, 0);
      std::vector<std::vector<int>> label(n_a);
      for (int i_gr = 0; i_gr < n_a; i_gr++)
        for (int i_pkj = 0; i_pkj < n_b; i_pkj++)
          if (std::fabs(c32_x(i_pkj, nx_b) - gr_x(i_gr, nx_a)) < 120.0 &&
              std::fabs(c32_y(i_pkj, nx_b) - gr_y(i_gr, nx_a)) < 220.0) {
            label[i_gr].push_back(i_pkj);
            num_pkj[i_pkj]++;
          }
      int pair = 0;
      for (int l_n_gr = 0; l_n_gr < n_a; l_n_gr++)
        for (int l_n_pkj : label[l_n_gr]) {
          const int ix_gr = l_n_gr % nx_a, ix_pkj = l_n_pkj % nx_b;
          const int iy_gr = l_n_gr / nx_a, iy_pkj = l_n_pkj / nx_b;
          const double diff_x = ix_gr * 16.0 - (8.0 + ix_pkj * 32.0);
          const double diff_y = iy_gr * 16.0 - (8.0 + iy_pkj * 32.0);
          const int gr_ax_comp = 250 + 77 + (int)(diff_y / 10.0);
          int pkj_d_comp = -1;
          double min_x = 1000000.0;   // not reset between a pair's synapses (syn.cu:479)
          const int guarantee = 1;
          for (int i_cp = 0; i_cp < N_SYN_PER_GRPKJ; i_cp++) {
            const int k = pair * N_SYN_PER_GRPKJ + i_cp;
            const double f = philox_u(k0, k1, (uint32_t)k, 0u, ST_GRPKJ_JITTER) - 0.5;
            for (int i_comp = 0; i_comp < PKJ_COMP; i_comp++)
              if (std::fabs(x1600[i_comp] - diff_x + 10.0 * f) < min_x) {
                min_x = std::fabs(x1600[i_comp] - diff_x + 10.0 * f);
                pkj_d_comp = i_comp;
              }
            push(gr_ax_comp + GR_COMP * l_n_gr, pkj_d_comp + PKJ_COMP * l_n_pkj,
                 1000000.0 * 0.3 / (512.0 * num_pkj[l_n_pkj] * 1.0 * guarantee / 10.0));
          }
          pair++;
        }
      break;
    }
    case SYN_MLIPKJ:  // mlipkj_create / mlipkj_initialize (syn.cu:666-738); a = PKJ, b = GR
      out.E = E_MLIPKJ;
      if (n_a == 0 || n_b == 0) break;
      for (int id = 0; id < n_a; id++) push(id, 1599 + id * PKJ_COMP, 0.0 / (1.0));
      break;
    default:
      throw Error("unknown synapse kind");
  }
  out.n = (int)out.post.size();
}

void syn_update(const SynDev &s, const double *pre_v, double t, uint32_t step, hipStream_t st) {
  if (s.n == 0) return;
  k_syn_update<<<(s.n + 255) / 256, 256, 0, st>>>(s, pre_v, t, step);
  CB_HIP(hipGetLastError());
}

void syn_macro_current(const SynIn &S, const double *y, double *s, int nc, hipStream_t st) {
  k_syn_macro<<<(nc + 255) / 256, 256, 0, st>>>(S, y, s, nc);
  CB_HIP(hipGetLastError());
}

}  // namespace cbrkc
