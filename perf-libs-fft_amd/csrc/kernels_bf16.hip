// kernels_bf16.hip -- bf16 storage with FP32 arithmetic.
#include "kernels_xlane.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_bf16(std::vector<KEntry> &r) {
  using S = bf16_pair;
  constexpr int P = P_BF16;
  //                          TPT  TPB  radix chain
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,    1, 256, 8);
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
  PLF_STATIC(k_c2c, KK_C2C, S, P,    8,  32, 16, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   16,  16, 16, 16);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   64,   4, 8, 8, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,   64,   4, 16, 16, 4);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  128,   2, 16, 16, 8);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  256,   1, 16, 16, 16);
  PLF_STATIC(k_c2c, KK_C2C, S, P,  512,   1, 16, 16, 16, 2);
  PLF_STATIC(k_c2c, KK_C2C, S, P, 1024,   1, 16, 16, 16, 4);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   16,  16, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,   64,   4, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P,  256,   1, 16);
  PLF_DYNAMIC(k_c2c, KK_C2C, S, P, 1024,   1, 16);
  PLF_DYNAMIC(k_col, KK_COL, S, P,   64,  16, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,   64,  16, 16);
  PLF_DYNAMIC(k_col, KK_COL, S, P,  256,   4, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,  256,   4, 16);
  // unit stride: non-temporal loads and stores, no stride arithmetic
  //              TPT  TPB PAD PAIR   radix chain
  PLF_XLANE(S, P, 16, 2, 2, 2, 2);  // cross-lane, packed members
  PLF_XLANE(S, P, 32, 2, 2, 2, 2, 2);
  // small odd primes with packed members: LDS-staged global access
  PLF_STAGED(S, P, 3);
  PLF_STAGED(S, P, 5);
  PLF_STAGED(S, P, 7);
  PLF_STAGED(S, P, 11);
  PLF_STAGED(S, P, 13);
  PLF_STAGED(S, P, 17);
  PLF_STAGED(S, P, 19);
  PLF_STAGED(S, P, 23);
  PLF_UNIT(S, P,  16,  16, 4, false, 16, 16);
  PLF_UNIT(S, P,  64,   4, 4, false, 16, 16, 4);
  // 4096 and 16384: the split (real, then imaginary) LDS exchange first
  // (the ESTIMATE choice): half the bytes per point leave LDS occupancy as
  // the bound, and the split exchange fits more blocks per CU (4096: 0.39 of
  // the HBM peak against 0.29 with the plain exchange); 16384 at 64
  // registers (8 waves per SIMD, two blocks per CU)
  PLF_UNIT_SPLIT(S, P, 256, 1, 5, 16, 16, 16);
  PLF_UNIT_SPLIT_OCC(S, P, 1024, 1, 5, 8, 16, 16, 16, 4);
  PLF_UNIT_SPLIT(S, P, 1024, 1, 5, 16, 16, 16, 4);
  PLF_UNIT(S, P, 256,   1, 4, false, 16, 16, 16);
  PLF_UNIT(S, P, 1024,  1, 4, false, 16, 16, 16, 4);
}

}  // namespace plf
