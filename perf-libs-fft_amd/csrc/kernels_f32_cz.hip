// kernels_f32_cz.hip -- FP32 three-pass chirp-z kernels (the column and row
// engines of the four-step passes, sides 128..1024).
#include "kernels_cz.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f32_cz(std::vector<KEntry> &r) {
  using S = float;
  constexpr int P = P_F32;
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   64,  16, 16, 16, 4);
}

}  // namespace plf
