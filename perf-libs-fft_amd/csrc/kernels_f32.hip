// kernels_f32.hip -- FP32 kernel configurations.
#include "launch.hpp"

namespace plf {

void register_kernels_f32(std::vector<KEntry> &r) {
  using S = float;
  constexpr int P = P_F32;
  // batched C2C, compile-time chains (first entry per n is the default)
  //                          TPT  TPB  radix chain
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    4,  64, 4, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 16);
  // prime n <= 23: one thread per member, the whole odd-prime butterfly in
  // registers (the runtime-chain kernel spends 16 threads on its one
  // butterfly)
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 3);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 5);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 7);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 11);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 13);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 17);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 19);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 23);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    4,  64, 8, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    8,  32, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    4,  64, 16, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    8,  32, 16, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   16,  16, 16, 16);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   32,   8, 16, 16, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   64,   4, 8, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   64,   4, 16, 16, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   32,   8, 32, 32);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  128,   2, 16, 16, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  256,   1, 16, 16, 16);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  128,   2, 16, 16, 16);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  512,   1, 16, 16, 16, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  256,   1, 16, 16, 32);
  PLF_STATIC(k_c2c, KK_C2C, S, P, 1024,   1, 16, 16, 16, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  512,   1, 16, 16, 16, 4);
  // batched C2C, runtime chains
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   16,  16, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   64,   4, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,  256,   1, 16);
  // 15015 = 3.5.7.11.13 in one LDS pass on a compile-time chain
  PLF_STATIC(k_c2c, KK_C2C, S, P, 1024,   1, 3, 5, 7, 11, 13);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P, 1024,   1, 16);
}

}  // namespace plf
