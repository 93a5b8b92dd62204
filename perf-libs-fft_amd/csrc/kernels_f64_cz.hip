// kernels_f64_cz.hip -- FP64 three-pass chirp-z and Rader kernels (the
// column and row engines of the four-step passes, sides 64..512).
#include "kernels_cz.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f64_cz(std::vector<KEntry> &r) {
  using S = double;
  constexpr int P = P_F64;
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   64,   8, 8, 8, 8);
  // Rader's permuted column passes, the same column sides
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,    8,   8, 8, 8);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   16,   8, 8, 4, 4);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   32,   8, 8, 8, 4);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   64,   8, 8, 8, 8);
  // row sides with a factor 3, for M = 3 * 2^k
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   16,   8, 8, 4, 3);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   32,   8, 8, 8, 3);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,   4, 8, 8, 2, 3);
}

}  // namespace plf
