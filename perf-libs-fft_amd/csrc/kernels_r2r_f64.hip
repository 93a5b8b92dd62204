// kernels_r2r_f64.hip -- DCT-II / DST-II and DCT-III / DST-III kernel configurations.
#include "kernels_real.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_r2r_f64(std::vector<KEntry> &r) {
  {
    using S = double;
    constexpr int P = P_F64;
    PLF_STATIC(k_dct2_even, KK_DCT2_E, S, P, 64, 4, 8, 8, 8);
    PLF_STATIC(k_dct2_even, KK_DCT2_E, S, P, 256, 1, 8, 8, 8, 4);
    PLF_STATIC(k_dct3_even, KK_DCT3_E, S, P, 64, 4, 8, 8, 8);
    PLF_STATIC(k_dct3_even, KK_DCT3_E, S, P, 256, 1, 8, 8, 8, 4);
    PLF_DYNAMIC(k_dct2_even, KK_DCT2_E, S, P, 16, 16, 16);
    PLF_DYNAMIC(k_dct2_even, KK_DCT2_E, S, P, 64, 4, 16);
    PLF_DYNAMIC(k_dct2_even, KK_DCT2_E, S, P, 256, 1, 16);
    PLF_DYNAMIC(k_dct2_even, KK_DCT2_E, S, P, 1024, 1, 8);
    PLF_DYNAMIC(k_dct3_even, KK_DCT3_E, S, P, 16, 16, 16);
    PLF_DYNAMIC(k_dct3_even, KK_DCT3_E, S, P, 64, 4, 16);
    PLF_DYNAMIC(k_dct3_even, KK_DCT3_E, S, P, 256, 1, 16);
    PLF_DYNAMIC(k_dct3_even, KK_DCT3_E, S, P, 1024, 1, 8);
    PLF_DYNAMIC(k_dct2_full, KK_DCT2_F, S, P, 16, 16, 16);
    PLF_DYNAMIC(k_dct2_full, KK_DCT2_F, S, P, 64, 4, 16);
    PLF_DYNAMIC(k_dct2_full, KK_DCT2_F, S, P, 256, 1, 16);
    PLF_DYNAMIC(k_dct2_full, KK_DCT2_F, S, P, 1024, 1, 8);
    PLF_DYNAMIC(k_dct3_full, KK_DCT3_F, S, P, 16, 16, 16);
    PLF_DYNAMIC(k_dct3_full, KK_DCT3_F, S, P, 64, 4, 16);
    PLF_DYNAMIC(k_dct3_full, KK_DCT3_F, S, P, 256, 1, 16);
    PLF_DYNAMIC(k_dct3_full, KK_DCT3_F, S, P, 1024, 1, 8);
  }
}

}  // namespace plf
