// kernels_f32_cz.hip -- three-pass chirp-z and Rader kernels with FP32 arithmetic
// (f32, f16 and bf16 storage; the column and row engines of the four-step
// passes, sides 64..1024).
#include "kernels_cz.hpp"
#include "launch.hpp"

namespace plf {

// sides 64..1024 with FP32 arithmetic for storage S
template <typename S, int P>
void register_cz_f32_compute(std::vector<KEntry> &r) {
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,    8,  16, 16, 8);
  // 256: 32 threads x 8 points (42 VGPRs) measured 3 % faster than 16 x 16;
  // the chirp-out, Rader and row passes measured 2-3 % slower that way
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   32,  16, 8, 8, 4);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_cz_col_in, KK_CZ_COL_IN, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   16,  32, 16, 16);
  // 512: 64 threads x 8 points (70 VGPRs, 7 waves per SIMD) measured 2 %
  // faster than 32 x 16 (16, 16, 2)
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,   8, 8, 8, 8);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_cz_col_out, KK_CZ_COL_OUT, S, P,   64,  16, 16, 16, 4);
  // Rader's permuted column passes, the same column sides
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_rd_col_in, KK_RD_COL_IN, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_rd_col_out, KK_RD_COL_OUT, S, P,   64,  16, 16, 16, 4);
  // row sides with a factor 3, for M = 3 * 2^k
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,    8,  16, 8, 4, 3);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   16,  16, 16, 4, 3);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   32,   8, 16, 8, 3);
  PLF_STATIC(k_cz_row_mid, KK_CZ_ROW_MID, S, P,   64,   4, 16, 16, 3);
}

void register_kernels_f32_cz(std::vector<KEntry> &r) {
  register_cz_f32_compute<float, P_F32>(r);
  register_cz_f32_compute<half_pair, P_F16>(r);
  register_cz_f32_compute<bf16_pair, P_BF16>(r);
}

}  // namespace plf
