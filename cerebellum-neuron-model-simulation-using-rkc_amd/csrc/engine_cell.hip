// placeholder, replaced by the per-cell engine
#include "engine.hpp"
namespace cbrkc {
Engine *make_cell_pkj(const PkjHost &, const SolverSpec &, const StimSpec &, const PoissonSpec &, bool) {
  throw Error("per-cell engine not built");
}
}
