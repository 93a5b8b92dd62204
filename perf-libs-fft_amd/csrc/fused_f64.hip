// fused_f64.hip -- FP64 fused Stockham kernel configurations.
#include "fused_launch.hpp"

namespace plf {

void register_fused_f64(std::vector<FusedEntry> &r) {
  //        storage  prec   TPT  TPB PAD  radix chain
  PLF_FUSED(double, P_F64,   1, 256, 5, 2);
  PLF_FUSED(double, P_F64,   1, 256, 5, 4);
  PLF_FUSED(double, P_F64,   1, 256, 5, 8);
  PLF_FUSED(double, P_F64,   4,  64, 5, 4, 4);
  PLF_FUSED(double, P_F64,   4,  64, 5, 8, 4);
  PLF_FUSED(double, P_F64,   8,  32, 5, 8, 8);
  PLF_FUSED(double, P_F64,  16,  16, 5, 8, 4, 4);
  PLF_FUSED(double, P_F64,  32,   8, 5, 8, 8, 4);
  PLF_FUSED(double, P_F64,  64,   4, 5, 8, 8, 8);
  PLF_FUSED(double, P_F64, 128,   2, 5, 8, 8, 4, 4);
  PLF_FUSED(double, P_F64, 256,   1, 5, 8, 8, 8, 4);
  PLF_FUSED(double, P_F64, 512,   1, 5, 8, 8, 8, 8);
  PLF_FUSED(double, P_F64, 1024,  1, 5, 8, 8, 8, 8, 2);
}

}  // namespace plf
