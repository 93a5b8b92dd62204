// kernels_mixed_h.hip -- mixed-length batch kernel, FP16 and BF16
// storage (FP32 arithmetic), plus the precision dispatch.
#include "kernels_mixed_impl.hpp"
#include "registry.hpp"

namespace plf {

hipError_t launch_mixed_f32(int level, const MixedArgs &a, int grid, size_t lds,
                            hipStream_t s);
hipError_t launch_mixed_f64(int level, const MixedArgs &a, int grid, size_t lds,
                            hipStream_t s);
int mixed_blocks_per_cu_f32(int level, size_t lds);
int mixed_blocks_per_cu_f64(int level, size_t lds);

hipError_t launch_mixed(int prec, int level, const MixedArgs &a, int grid, size_t lds,
                        hipStream_t s) {
  switch (prec) {
    case P_F32: return launch_mixed_f32(level, a, grid, lds, s);
    case P_F64: return launch_mixed_f64(level, a, grid, lds, s);
    case P_F16: return launch_mixed_t<half_pair>(level, a, grid, lds, s);
    case P_BF16: return launch_mixed_t<bf16_pair>(level, a, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

int mixed_blocks_per_cu(int prec, int level, size_t lds) {
  switch (prec) {
    case P_F32: return mixed_blocks_per_cu_f32(level, lds);
    case P_F64: return mixed_blocks_per_cu_f64(level, lds);
    case P_F16: return mixed_blocks_per_cu_t<half_pair>(level, lds);
    case P_BF16: return mixed_blocks_per_cu_t<bf16_pair>(level, lds);
    default: return 0;
  }
}

}  // namespace plf
