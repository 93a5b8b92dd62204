// kernels_mixed_f64.hip -- mixed-length batch kernel, FP64 storage.
#include "kernels_mixed_impl.hpp"

namespace plf {

hipError_t launch_mixed_f64(int level, const MixedArgs &a, int grid, size_t lds,
                            hipStream_t s) {
  return launch_mixed_t<double>(level, a, grid, lds, s);
}
int mixed_blocks_per_cu_f64(int level, size_t lds) {
  return mixed_blocks_per_cu_t<double>(level, lds);
}

}  // namespace plf
