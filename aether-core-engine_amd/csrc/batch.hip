// batch.hip -- batched grids (FP32 and FP64 instantiations).
#include "batch.hpp"

namespace aether {
std::unique_ptr<BatchBase> make_batch(const aether_fluid_instance_desc *desc, int count,
                                      int precision, const aether_fluid_params &shared,
                                      hipStream_t st) {
    if (precision == AETHER_FP64) return std::make_unique<Batch<double>>(desc, count, shared, st);
    return std::make_unique<Batch<float>>(desc, count, shared, st);
}
}  // namespace aether
