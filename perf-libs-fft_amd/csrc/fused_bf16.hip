// fused_bf16.hip -- bf16 storage, FP32 arithmetic, fused Stockham kernels.
#include "fused_launch.hpp"

namespace plf {

void register_fused_bf16(std::vector<FusedEntry> &r) {
  //        storage     prec    TPT  TPB PAD  radix chain
  PLF_FUSED(bf16_pair, P_BF16,    1, 256, 5, 2);
  PLF_FUSED(bf16_pair, P_BF16,    1, 256, 5, 4);
  PLF_FUSED(bf16_pair, P_BF16,    1, 256, 5, 8);
  PLF_FUSED(bf16_pair, P_BF16,    1, 256, 5, 16);
  PLF_FUSED(bf16_pair, P_BF16,    4,  64, 5, 8, 4);
  PLF_FUSED(bf16_pair, P_BF16,    8,  32, 5, 8, 8);
  PLF_FUSED(bf16_pair, P_BF16,    8,  32, 5, 16, 8);
  PLF_FUSED(bf16_pair, P_BF16,   16,  16, 5, 16, 16);
  PLF_FUSED(bf16_pair, P_BF16,   64,   4, 5, 8, 8, 8);
  PLF_FUSED(bf16_pair, P_BF16,   64,   4, 5, 16, 16, 4);
  PLF_FUSED(bf16_pair, P_BF16,  128,   2, 5, 16, 16, 8);
  PLF_FUSED(bf16_pair, P_BF16,  256,   1, 5, 16, 16, 16);
  PLF_FUSED(bf16_pair, P_BF16,  512,   1, 5, 16, 16, 16, 2);
  PLF_FUSED(bf16_pair, P_BF16, 1024,   1, 5, 16, 16, 16, 4);
}

}  // namespace plf
