// kernels_mixed_f32.hip -- mixed-length batch kernel, FP32 storage.
#include "kernels_mixed_impl.hpp"

namespace plf {

hipError_t launch_mixed_f32(int level, const MixedArgs &a, int grid, size_t lds,
                            hipStream_t s) {
  return launch_mixed_t<float>(level, a, grid, lds, s);
}
int mixed_blocks_per_cu_f32(int level, size_t lds) {
  return mixed_blocks_per_cu_t<float>(level, lds);
}

}  // namespace plf
