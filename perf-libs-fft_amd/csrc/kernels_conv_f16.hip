// kernels_conv_f16.hip -- f16 Bluestein / Rader kernel configurations.
#include "kernels_conv.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_conv_f16(std::vector<KEntry> &r) {
    using S = half_pair;
    constexpr int P = P_F16;
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P,  64,   4, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 256,   1, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 1024,  1, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,   64,   4, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,  256,   1, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P, 1024,   1, 16);
}

}  // namespace plf
