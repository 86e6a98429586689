// Seeded synthetic inputs in the reference's file formats (host):
//
//   * granule chains: Gr.txt rows (idx parent diam len Ra Cm type + 10
//     densities, gr.cu:29-58), `comp` compartments per cell in one block per
//     cell, an unbranched chain whose compartment k takes Ra, Cm, type and
//     densities from the first row of the template file with type types[k]
//     and its diameter and length from that row scaled by U(0.8, 1.25);
//   * Golgi trees: Go_131.txt rows (idx parent diam len Ra Cm type + 13
//     densities, go.cu:30-63), one block per cell: the template's topology,
//     types and densities with every diameter and length scaled by
//     U(0.8, 1.25);
//   * inferior-olive gap junctions: "pre,post,g_nS" lines, each cell joined
//     to k distinct random partners (pairs listed once, pre < post), g from
//     U(g_min, g_max) nS.
//
// Every draw comes from Philox4x32-10 with key (seed, stream) and counter
// (cell, draw, 0, 0), so a cell's geometry does not depend on how many cells
// are written.
#include <algorithm>
#include <cstdio>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "philox.hpp"

namespace cbrkc {
namespace {

double u01(uint64_t seed, uint32_t stream, uint32_t cell, uint32_t k) {
  const u32x4 r = philox4x32_10(cell, k, 0u, 0u, (uint32_t)seed ^ stream, (uint32_t)(seed >> 32));
  return u01_53(r.v[0], r.v[1]);
}

struct Row {
  int idx, parent, type;
  double diam, len, ra, cm;
  std::vector<double> dens;
};

bool read_rows(const std::string &path, int ndens, std::vector<Row> &rows, std::string &err) {
  std::FILE *f = std::fopen(path.c_str(), "r");
  if (!f) {
    err = "no such file " + path;
    return false;
  }
  rows.clear();
  while (true) {
    Row r;
    r.dens.assign(ndens, 0.0);
    int got = std::fscanf(f, "%d %d %lf %lf %lf %lf %d", &r.idx, &r.parent, &r.diam, &r.len, &r.ra, &r.cm, &r.type);
    if (got == EOF) break;
    bool ok = got == 7;
    for (int q = 0; q < ndens && ok; q++) ok = std::fscanf(f, "%lf", &r.dens[q]) == 1;
    if (!ok) {
      std::fclose(f);
      err = "malformed row in " + path;
      return false;
    }
    rows.push_back(r);
  }
  std::fclose(f);
  if (rows.empty()) {
    err = "empty file " + path;
    return false;
  }
  return true;
}

void write_row(std::FILE *f, const Row &r) {
  std::fprintf(f, "%d\t%d\t%.6f\t%.6f\t%g\t%g\t%d", r.idx, r.parent, r.diam, r.len, r.ra, r.cm, r.type);
  for (double d : r.dens) std::fprintf(f, "\t%g", d);
  std::fprintf(f, "\n");
}

constexpr uint32_t STREAM_GR = 0x67726368u;   // "grch"
constexpr uint32_t STREAM_GO = 0x676f7472u;   // "gotr"
constexpr uint32_t STREAM_IO = 0x696f676au;   // "iogj"

}  // namespace

bool synth_gr_chains(const std::string &path, const std::string &template_path, uint64_t seed, int ncell,
                     const std::vector<int> &types, std::string &err) {
  std::vector<Row> tpl;
  if (!read_rows(template_path, 10, tpl, err)) return false;
  const int comp = (int)types.size();
  std::vector<const Row *> proto(comp, nullptr);
  for (int k = 0; k < comp; k++) {
    for (const Row &r : tpl)
      if (r.type == types[k]) {
        proto[k] = &r;
        break;
      }
    if (!proto[k]) {
      err = "no row of type " + std::to_string(types[k]) + " in " + template_path;
      return false;
    }
  }
  std::FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot write " + path;
    return false;
  }
  for (int c = 0; c < ncell; c++)
    for (int k = 0; k < comp; k++) {
      Row r = *proto[k];
      r.idx = k;
      r.parent = (k + 1 < comp) ? k + 1 : -1;   // the chain's root is its last compartment
      r.diam *= 0.8 + 0.45 * u01(seed, STREAM_GR, (uint32_t)c, (uint32_t)(2 * k));
      r.len *= 0.8 + 0.45 * u01(seed, STREAM_GR, (uint32_t)c, (uint32_t)(2 * k + 1));
      write_row(f, r);
    }
  std::fclose(f);
  return true;
}

bool synth_go_trees(const std::string &path, const std::string &template_path, uint64_t seed, int ncell,
                    std::string &err) {
  std::vector<Row> tpl;
  if (!read_rows(template_path, 13, tpl, err)) return false;
  std::FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot write " + path;
    return false;
  }
  for (int c = 0; c < ncell; c++)
    for (size_t k = 0; k < tpl.size(); k++) {
      Row r = tpl[k];
      r.diam *= 0.8 + 0.45 * u01(seed, STREAM_GO, (uint32_t)c, (uint32_t)(2 * k));
      r.len *= 0.8 + 0.45 * u01(seed, STREAM_GO, (uint32_t)c, (uint32_t)(2 * k + 1));
      write_row(f, r);
    }
  std::fclose(f);
  return true;
}

bool synth_io_gj(const std::string &path, uint64_t seed, int ncell, int k, double g_min, double g_max,
                 std::string &err) {
  if (ncell < 2 || k < 1 || k >= ncell) {
    err = "need 1 <= k < cells";
    return false;
  }
  std::set<std::pair<int, int>> edges;
  for (int c = 0; c < ncell; c++) {
    std::set<int> mine;
    for (uint32_t d = 0; (int)mine.size() < k; d++) {
      int j = (int)(u01(seed, STREAM_IO, (uint32_t)c, d) * (ncell - 1));
      if (j >= c) j++;   // anyone but the cell itself
      if (j >= ncell) j = ncell - 1;
      if (j == c) continue;
      mine.insert(j);
    }
    for (int j : mine) edges.insert({std::min(c, j), std::max(c, j)});
  }
  std::FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    err = "cannot write " + path;
    return false;
  }
  uint32_t e = 0;
  for (const auto &pq : edges) {
    const double g = g_min + (g_max - g_min) * u01(seed, STREAM_IO ^ 0x1u, e++, 0u);
    std::fprintf(f, "%d,%d,%.17g\n", pq.first, pq.second, g);
  }
  std::fclose(f);
  return true;
}

}  // namespace cbrkc
