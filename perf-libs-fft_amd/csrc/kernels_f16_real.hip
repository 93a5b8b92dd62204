// kernels_f16_real.hip -- f16 storage real-data (R2C / C2R) kernels.
#include "kernels_real.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f16_real(std::vector<KEntry> &r) {
    using S = half_pair;
    constexpr int P = P_F16;
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 16, 16, 16, 16);
    PLF_STATIC1(k_r2c_even, KK_R2C_E, S, P, 256, 1, 16, 16, 16);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 16, 16, 16, 16);
    PLF_STATIC1(k_c2r_even, KK_C2R_E, S, P, 256, 1, 16, 16, 16);
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

}  // namespace plf
