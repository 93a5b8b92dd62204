// kernels_f64_real4s.hip -- four-step R2C / C2R passes (kernels_real4s.hpp)
// for FP64: the column and row configurations of the complex FP64 four-step
// passes (kernels_f64.hip).
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

void register_kernels_f64_real4s(std::vector<KEntry> &r) {
  using S = double;
  constexpr int P = P_F64;
  //                       TPT  TPB  radix chain
  PLF_REAL4S_STATIC(S, P,    8,   8, 8, 8);
  PLF_REAL4S_STATIC(S, P,   16,   8, 8, 4, 4);
  PLF_REAL4S_STATIC(S, P,   32,   8, 8, 8, 4);
  PLF_REAL4S_STATIC(S, P,   64,   8, 8, 8, 8);
  PLF_REAL4S_STATIC(S, P,  128,   8, 8, 8, 4, 4);
  PLF_REAL4S_STATIC(S, P,  256,   4, 8, 8, 8, 4);
  PLF_REAL4S_DYNAMIC(S, P,  64,   8, 16);
  PLF_REAL4S_DYNAMIC(S, P, 128,   4, 16);
}

}  // namespace plf
