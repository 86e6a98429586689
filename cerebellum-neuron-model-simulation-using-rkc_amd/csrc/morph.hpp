// Host-side population builder: morphology loading, synthetic tree
// generation, per-compartment parameters and the cable matrix.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cbrkc {

// One line of a Pkj.txt-format morphology: idx parent x y z diam type
// (pkj.cu:30-34).  Lengths are micrometres; parent -1 marks the root.
struct MorphRow {
  int idx, parent, type;
  double x, y, z, d;
};

// Reads every whitespace-separated 7-column row of a Pkj.txt-format file.
// Returns false (with a message in err) when the file is missing or its row
// count is not a multiple of comp.
bool read_pkj_morph(const std::string &path, int comp, std::vector<MorphRow> &rows, std::string &err);

// Reads the 26 "pre, post" pairs of Pkj_GJ.csv (pkj_solve.cu:86-97).
bool read_pkj_gj(const std::string &path, std::vector<int> &pre, std::vector<int> &post, std::string &err);

// Seeded synthetic Purkinje tree in Pkj.txt layout: soma (type 0) is the last
// index and every parent has a higher index than its children.  Soma sphere
// radius soma_r_um, binary branching, segment length U(lmin, lmax) um,
// diameter tapering linearly with depth from d0 to d1 um, scaled by dscale.
void synth_pkj_tree(uint64_t seed, int comp, double soma_r_um, double lmin, double lmax, double d0,
                    double d1, double dscale, std::vector<MorphRow> &rows);

bool write_pkj_morph(const std::string &path, const std::vector<MorphRow> &rows, std::string &err);

// Structure-of-arrays description of a Purkinje population on the host: the
// reference's neuron_t fields (param.h:27-35) plus the cable CSR
// (pkj_solve.cu:28-254) in RKC sign convention.
struct PkjHost {
  int n = 0, nx = 0, ny = 0, comp = 0, nc = 0;
  std::vector<double> cm, area, shell, iext, v, ca, rev;
  std::vector<int8_t> compart;
  std::vector<double> cond;   // [NCOND][nc]
  std::vector<double> gate;   // [NGATE][nc]
  std::vector<int> parent;    // global index of the parent compartment, -1 for roots
  std::vector<double> gpar;   // axial conductance to the parent [mS]
  // CSR with stored diagonal: off-diagonal +g, diagonal -sum(g) (RKC sign)
  std::vector<int> row, col;
  std::vector<double> val;
  int n_gj = 0;
};

// Builds the population from a morphology (one block replicated, or one
// block per cell), optional gap-junction list (only used when nx > 1) and
// per-cell initial voltages (nullptr: draw -68 + 2 (U - 0.5) from libc rand()
// per cell in order, as pkj_ion.cu:367-369 does).
bool build_pkj_host(const std::vector<MorphRow> &rows, int nx, int ny, const std::vector<int> *gj_pre,
                    const std::vector<int> *gj_post, const double *v0, PkjHost &out, std::string &err);

// Inferior-olive population on the host: per-area densities and Cm from
// IO.txt (io.cu:3-101), the asymmetric three-compartment coupling in CSR with
// the RKC sign (io_solve.cu:2-150), initial state (io_ion.cu:163-194) and the
// gap-junction partners of every cell (gap.cu:4-133).
struct IoHost {
  int n = 0, nx = 0, ny = 0, comp = 3, nc = 0;
  std::vector<double> cm, iext, v, ca;
  std::vector<int8_t> compart;
  std::vector<double> cond;   // [NCOND][nc]
  std::vector<double> gate;   // [NGATE][nc]
  std::vector<int> row, col;
  std::vector<double> val;
  std::vector<int> nbr;       // [n][4] partner cell or -1
};

// Reads IO.txt (idx parent diam len Cm type, 3 rows) and builds nx * ny cells;
// v0 per cell, or nullptr for -60 + 10 (U - 0.5) from libc rand() per cell in
// order (io_ion.cu:168-170).
bool build_io_host(const std::string &param_path, int nx, int ny, const double *v0, IoHost &out, std::string &err);

// Golgi-cell population on the host: Go_131.txt (idx parent diam len Ra Cm
// type + 13 channel densities, go.cu:30-63), conductances and Cm as density x
// cylinder area, the cable CSR with the grid gap junctions between the cells'
// compartments 0 (go_solve.cu:4-215, RKC sign) and the initial state
// (go_ion.cu:590-634).
struct GoHost {
  int n = 0, nx = 0, ny = 0, comp = 131, nc = 0, n_gj = 0;
  std::vector<double> cm, area, iext, v, ca, ca2, rev, ca_old;
  std::vector<double> cond;   // [NCOND][nc]
  std::vector<double> gate;   // [NGATE][nc]
  std::vector<int> row, col;
  std::vector<double> val;
};

// Reads a Go_131.txt-format file and builds nx * ny cells; v0 per cell, or
// nullptr for -60 + 2 (U - 0.5) from libc rand() per cell in order
// (go_ion.cu:597-600).
// comp: compartments per cell (the reference's 131 by default); the file
// holds one block of comp rows for all cells or one block per cell;
// grid_gj = false leaves out the grid gap junctions.
bool build_go_host(const std::string &param_path, int nx, int ny, const double *v0, GoHost &out, std::string &err,
                   int comp = 131, bool grid_gj = true);

// Granule-cell population on the host: Gr.txt (idx parent diam len Ra Cm type
// + 10 channel densities, gr.cu:29-58), conductances and Cm as density x
// cylinder area, the cable CSR (gr_solve.cu:2-139, RKC sign) and the initial
// state (gr_ion.cu:635-672).
struct GrHost {
  int n = 0, nx = 0, ny = 0, comp = 578, nc = 0;
  std::vector<double> cm, iext, v, ca;
  std::vector<int8_t> compart;
  std::vector<double> cond;   // [NCOND][nc]
  std::vector<double> gate;   // [NGATE][nc]
  std::vector<int> row, col;
  std::vector<double> val;
};

// Reads a Gr.txt-format file and builds nx * ny cells; v0 per cell, or
// nullptr for -80 + 2 ((rand() + 1) / (RAND_MAX + 2) - 0.5) from libc rand()
// per cell in order (gr_ion.cu:642-644).
// comp: compartments per cell (the reference's 578 by default); the file
// holds one block of comp rows for all cells or one block per cell.
bool build_gr_host(const std::string &param_path, int nx, int ny, const double *v0, GrHost &out, std::string &err,
                   int comp = 578);

// Spectral radius of each cell's cable operator C^-1 L (the RKC Jacobian
// with the ionic current frozen), from the CSR rows of the cell without
// intercell entries: Lanczos with full reorthogonalisation on the symmetric
// form C^-1/2 L C^-1/2 (80 steps), largest Ritz value.
void cable_spectral_radius(const PkjHost &P, std::vector<double> &rho);

}  // namespace cbrkc
