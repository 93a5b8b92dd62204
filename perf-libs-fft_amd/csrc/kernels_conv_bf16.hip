// kernels_conv_bf16.hip -- bf16 Bluestein / Rader kernel configurations.
#include "kernels_conv.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_conv_bf16(std::vector<KEntry> &r) {
    using S = bf16_pair;
    constexpr int P = P_BF16;
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P,  64,   4, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 256,   1, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 1024,  1, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,   64,   4, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,  256,   1, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P, 1024,   1, 16);
}

}  // namespace plf
