// kernels_conv_f64.hip -- f64 Bluestein / Rader kernel configurations.
#include "kernels_conv.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_conv_f64(std::vector<KEntry> &r) {
    using S = double;
    constexpr int P = P_F64;
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,    8,  32, 8, 8);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,   16,  16, 8, 4, 4);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,   32,   8, 8, 8, 4);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,   64,   4, 8, 8, 8);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,  128,   2, 8, 8, 4, 4);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,  256,   1, 8, 8, 8, 4);
    PLF_STATIC(k_bluestein, KK_BLUE, S, P,  512,   1, 8, 8, 8, 8);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P,  64,   4, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 256,   1, 16);
    PLF_DYNAMIC(k_bluestein, KK_BLUE, S, P, 1024,  1, 8);
    // compile-time chains for the Rader lengths of n = 97 (96) and 1009 (1008)
    // 96 on 16 threads x 6 points per member (11 % faster than 8 x 12);
    // 1008 as 16 x 9 x 7 (radix 9 = 3 x 3 in registers: one LDS exchange
    // fewer than 16 x 3 x 3 x 7, 124 instead of 164 registers in FP32,
    // 9 % faster; 4 % in FP64); 1008 on 128 threads measured 5-11 % slower
    PLF_STATIC(k_rader, KK_RADER, S, P,   16,  16, 4, 8, 3);
    PLF_STATIC(k_rader, KK_RADER, S, P,   64,   4, 16, 9, 7);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,   16,  16, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,   64,   4, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P,  256,   1, 16);
    PLF_DYNAMIC(k_rader, KK_RADER, S, P, 1024,   1, 8);
}

}  // namespace plf
