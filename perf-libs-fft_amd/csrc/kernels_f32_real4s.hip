// kernels_f32_real4s.hip -- four-step R2C / C2R passes (kernels_real4s.hpp)
// for FP32, FP16 and BF16 storage: the column and row configurations of the
// complex four-step passes (kernels_f32_4s.hip).
#include "kernels_real4s.hpp"
#include "launch.hpp"

namespace plf {

#define PLF_REAL4S_STATIC(S, P, TPT, TPB, ...)                                  \
  do {                                                                        \
    PLF_STATIC1(k_col_r2c, KK_COL_R2C, S, P, TPT, TPB, __VA_ARGS__);          \
    PLF_STATIC1(k_row_r2c, KK_ROW_R2C, S, P, TPT, TPB, __VA_ARGS__);          \
    PLF_STATIC1(k_col_c2r, KK_COL_C2R, S, P, TPT, TPB, __VA_ARGS__);          \
    PLF_STATIC1(k_row_c2r, KK_ROW_C2R, S, P, TPT, TPB, __VA_ARGS__);          \
  } while (0)
#define PLF_REAL4S_DYNAMIC(S, P, TPT, TPB, PTS)                                 \
  do {                                                                        \
    PLF_DYNAMIC1(k_col_r2c, KK_COL_R2C, S, P, TPT, TPB, PTS);                 \
    PLF_DYNAMIC1(k_row_r2c, KK_ROW_R2C, S, P, TPT, TPB, PTS);                 \
    PLF_DYNAMIC1(k_col_c2r, KK_COL_C2R, S, P, TPT, TPB, PTS);                 \
    PLF_DYNAMIC1(k_row_c2r, KK_ROW_C2R, S, P, TPT, TPB, PTS);                 \
  } while (0)

void register_kernels_f32_real4s(std::vector<KEntry> &r) {
  {
    using S = float;
    constexpr int P = P_F32;
    //                       TPT  TPB  radix chain
    PLF_REAL4S_STATIC(S, P,    8,  16, 8, 8);
    PLF_REAL4S_STATIC(S, P,    8,  16, 16, 8);
    PLF_REAL4S_STATIC(S, P,   16,  32, 16, 16);
    PLF_REAL4S_STATIC(S, P,   32,  16, 16, 16, 2);
    PLF_REAL4S_STATIC(S, P,   64,  16, 16, 16, 4);
    PLF_REAL4S_STATIC(S, P,  128,   8, 16, 16, 8);
    PLF_REAL4S_STATIC(S, P,  256,   4, 16, 16, 16);
    PLF_REAL4S_DYNAMIC(S, P,  64,  16, 16);
    PLF_REAL4S_DYNAMIC(S, P, 256,   4, 16);
  }
  {
    using S = half_pair;
    constexpr int P = P_F16;
    PLF_REAL4S_DYNAMIC(S, P,  64,  16, 16);
    PLF_REAL4S_DYNAMIC(S, P, 256,   4, 16);
  }
  {
    using S = bf16_pair;
    constexpr int P = P_BF16;
    PLF_REAL4S_DYNAMIC(S, P,  64,  16, 16);
    PLF_REAL4S_DYNAMIC(S, P, 256,   4, 16);
  }
}

}  // namespace plf
