// solver_f64.hip -- the FP64 instantiation of the one-grid solver (the
// reference's arithmetic, for parity).
#include "solver.hpp"

namespace aether {
std::unique_ptr<SolverBase> make_solver_f64(int nx, int ny, int nz, double h, hipStream_t st) {
    return std::make_unique<Solver<double>>(nx, ny, nz, h, st);
}
}  // namespace aether
