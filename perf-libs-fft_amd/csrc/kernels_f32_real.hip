// kernels_f32_real.hip -- FP32 real-data (R2C / C2R) kernel configurations.
#include "kernels_real.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f32_real(std::vector<KEntry> &r) {
  {
    using S = float;
    constexpr int P = P_F32;
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 4, 64, 8, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 8, 32, 8, 8);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 8, 32, 16, 8);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 16, 16, 16, 16);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 32, 8, 16, 16, 2);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 64, 4, 16, 16, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 128, 2, 16, 16, 8);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 256, 1, 16, 16, 16);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 512, 1, 16, 16, 16, 2);
    // n = 4096: the split (real, then imaginary) LDS exchange halves the LDS
    // per member, so more blocks share a CU; measured 6% faster than the
    // plain exchange (0.283 vs 0.300 ms for 32768 transforms)
    PLF_STATIC1_SPLIT(k_c2r_even, KK_C2R_E, S, P, 128, 2, 5, 16, 16, 8);
    // n = 65536: the half-length 32768-point transform held by 1024 threads
    // (32 points each) on the split exchange (132 KB of LDS), one HBM pass
    // instead of the two of the four-step passes
    PLF_STATIC1_SPLIT(k_r2c_even_split, KK_R2C_E, S, P, 1024, 1, 5, 16, 16, 16, 8);
    PLF_STATIC1_SPLIT(k_c2r_even_split, KK_C2R_E, S, P, 1024, 1, 5, 16, 16, 16, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 4, 64, 8, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 8, 32, 8, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 8, 32, 16, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 16, 16, 16, 16);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 32, 8, 16, 16, 2);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 64, 4, 16, 16, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 128, 2, 16, 16, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 256, 1, 16, 16, 16);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 512, 1, 16, 16, 16, 2);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 1024, 1, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 1024, 1, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 1024, 1, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 1024, 1, 16);
  }
}

}  // namespace plf
