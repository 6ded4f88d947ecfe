// solver_f32.hip -- the FP32 (and FP32 + FP64-refinement) instantiation of
// the one-grid solver, compiled on its own so that the library's units build
// in parallel.
#include "solver.hpp"

namespace aether {
std::unique_ptr<SolverBase> make_solver_f32(int nx, int ny, int nz, double h, hipStream_t st,
                                            bool ir64) {
    return std::make_unique<Solver<float>>(nx, ny, nz, h, st, ir64);
}
}  // namespace aether
