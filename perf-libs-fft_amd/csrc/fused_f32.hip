// fused_f32.hip -- FP32 fused Stockham kernel configurations.
#include "fused_launch.hpp"

namespace plf {

void register_fused_f32(std::vector<FusedEntry> &r) {
  //        storage  prec   TPT  TPB PAD  radix chain
  PLF_FUSED(float, P_F32,    1, 256, 5, 2);
  PLF_FUSED(float, P_F32,    1, 256, 5, 4);
  PLF_FUSED(float, P_F32,    1, 256, 5, 8);
  PLF_FUSED(float, P_F32,    4,  64, 5, 4, 4);
  PLF_FUSED(float, P_F32,    1, 256, 5, 16);
  PLF_FUSED(float, P_F32,    4,  64, 5, 8, 4);
  PLF_FUSED(float, P_F32,    8,  32, 5, 8, 8);
  PLF_FUSED(float, P_F32,    4,  64, 5, 16, 4);
  PLF_FUSED(float, P_F32,    8,  32, 5, 16, 8);
  PLF_FUSED(float, P_F32,   16,  16, 5, 16, 16);
  PLF_FUSED(float, P_F32,   32,   8, 5, 16, 16, 2);
  PLF_FUSED(float, P_F32,   64,   4, 5, 8, 8, 8);
  PLF_FUSED(float, P_F32,   64,   4, 5, 16, 16, 4);
  PLF_FUSED(float, P_F32,   32,   8, 5, 32, 32);
  PLF_FUSED(float, P_F32,  128,   2, 5, 16, 16, 8);
  PLF_FUSED(float, P_F32,  256,   1, 5, 16, 16, 16);
  PLF_FUSED(float, P_F32,  128,   2, 5, 16, 16, 16);
  PLF_FUSED(float, P_F32,  512,   1, 5, 16, 16, 16, 2);
  PLF_FUSED(float, P_F32,  256,   1, 5, 16, 16, 32);
  PLF_FUSED(float, P_F32, 1024,   1, 5, 16, 16, 16, 4);
  PLF_FUSED(float, P_F32,  512,   1, 5, 16, 16, 16, 4);
}

}  // namespace plf
