// kernels_f64_real.hip -- FP64 real-data (R2C / C2R) kernel configurations.
#include "kernels_real.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f64_real(std::vector<KEntry> &r) {
  {
    using S = double;
    constexpr int P = P_F64;
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 4, 64, 8, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 8, 32, 8, 8);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 16, 16, 8, 4, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 32, 8, 8, 8, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 64, 4, 8, 8, 8);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 128, 2, 8, 8, 4, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 256, 1, 8, 8, 8, 4);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 512, 1, 8, 8, 8, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 4, 64, 8, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 8, 32, 8, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 16, 16, 8, 4, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 32, 8, 8, 8, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 64, 4, 8, 8, 8);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 128, 2, 8, 8, 4, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 256, 1, 8, 8, 8, 4);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 512, 1, 8, 8, 8, 8);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_r2c_even, KK_R2C_E, S, P, 1024, 1, 8);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_c2r_even, KK_C2R_E, S, P, 1024, 1, 8);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_r2c_full, KK_R2C_F, S, P, 1024, 1, 8);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 16, 16, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 64, 4, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 256, 1, 16);
    PLF_DYNAMIC1(k_c2r_full, KK_C2R_F, S, P, 1024, 1, 8);
  }
}

}  // namespace plf
