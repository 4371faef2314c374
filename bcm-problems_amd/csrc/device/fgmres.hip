// Flexible GMRES and two-sided Lanczos drivers (filled in below).
#include "solver.hpp"

namespace bcm {

void solve_fgmres(Matrix&, Factor&, Context&, const SolveOptions&, const double*, double*, SolveStats&) {
  throw Error(kErrArg, "FGMRES driver not built yet");
}

void solve_lanczos(Matrix&, Factor&, Context&, const SolveOptions&, const double*, double*, SolveStats&) {
  throw Error(kErrArg, "Lanczos driver not built yet");
}

}  // namespace bcm
