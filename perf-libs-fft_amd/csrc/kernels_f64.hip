// kernels_f64.hip -- FP64 kernel configurations.
#include "launch.hpp"

namespace plf {

void register_kernels_f64(std::vector<KEntry> &r) {
  using S = double;
  constexpr int P = P_F64;
  //                          TPT  TPB  radix chain
  // small odd primes with packed members: LDS-staged global access
  PLF_STAGED(S, P, 3);
  PLF_STAGED(S, P, 5);
  PLF_STAGED(S, P, 7);
  PLF_STAGED(S, P, 11);
  PLF_STAGED(S, P, 13);
  PLF_STAGED(S, P, 17);
  PLF_STAGED(S, P, 19);
  PLF_STAGED(S, P, 23);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 8);
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
  PLF_STATIC(k_c2c, KK_C2C, S, P,    4,  64, 4, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    4,  64, 8, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    8,  32, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   16,  16, 8, 4, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   64,   4, 8, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  128,   2, 8, 8, 4, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  256,   1, 8, 8, 8, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  512,   1, 8, 8, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P, 1024,   1, 8, 8, 8, 8, 2);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   16,  16, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   64,   4, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,  256,   1, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P, 1024,   1, 8);
  // four-step passes: 8 columns of 16 B = 128 B runs
  PLF_STATIC(k_col, KK_COL, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_col, KK_COL, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_col, KK_COL, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_col, KK_COL, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_col, KK_COL, S, P,  128,   8, 8, 8, 4, 4);
  PLF_STATIC(k_col, KK_COL, S, P,  256,   4, 8, 8, 8, 4);
  PLF_STATIC(k_row, KK_ROW, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_row, KK_ROW, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_row, KK_ROW, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_row, KK_ROW, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_row, KK_ROW, S, P,  128,   8, 8, 8, 4, 4);
  PLF_STATIC(k_row, KK_ROW, S, P,  256,   4, 8, 8, 8, 4);
  // 15015 = 105 x 143: compile-time sides 3.5.7 and 11.13 (16 threads x 16
  // columns / rows per block)
  PLF_STATIC(k_col, KK_COL, S, P,   16,  16, 3, 5, 7);
  PLF_STATIC(k_row, KK_ROW, S, P,   16,  16, 11, 13);
  PLF_DYNAMIC(k_col, KK_COL, S, P,   64,   8, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,   64,   8, 16);
  PLF_DYNAMIC(k_col, KK_COL, S, P,  128,   4, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,  128,   4, 16);
}

}  // namespace plf
