// kernels_f32_4s.hip -- FP32 four-step column/row pass configurations.
#include "launch.hpp"

namespace plf {

void register_kernels_f32_4s(std::vector<KEntry> &r) {
  using S = float;
  constexpr int P = P_F32;
  // four-step passes (TPB = columns/rows per block). The first entry of a
  // length is the one used: 32-wide tiles (256-byte row segments) measured
  // 7% faster than 16-wide ones at 256 x 256; 64-wide ones slower.
  PLF_STATIC(k_col, KK_COL, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_row, KK_ROW, S, P,   16,  32, 16, 16);
  PLF_STATIC(k_col, KK_COL, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_col, KK_COL, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_col, KK_COL, S, P,   16,  16, 16, 16);
  PLF_STATIC(k_col, KK_COL, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_col, KK_COL, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_col, KK_COL, S, P,  128,   8, 16, 16, 8);
  PLF_STATIC(k_col, KK_COL, S, P,  256,   4, 16, 16, 16);
  PLF_STATIC(k_row, KK_ROW, S, P,    8,  16, 8, 8);
  PLF_STATIC(k_row, KK_ROW, S, P,    8,  16, 16, 8);
  PLF_STATIC(k_row, KK_ROW, S, P,   16,  16, 16, 16);
  PLF_STATIC(k_row, KK_ROW, S, P,   32,  16, 16, 16, 2);
  PLF_STATIC(k_row, KK_ROW, S, P,   64,  16, 16, 16, 4);
  PLF_STATIC(k_row, KK_ROW, S, P,  128,   8, 16, 16, 8);
  PLF_STATIC(k_row, KK_ROW, S, P,  256,   4, 16, 16, 16);
  PLF_DYNAMIC(k_col, KK_COL, S, P,   64,  16, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,   64,  16, 16);
  PLF_DYNAMIC(k_col, KK_COL, S, P,  256,   4, 16);
  PLF_DYNAMIC(k_row, KK_ROW, S, P,  256,   4, 16);
}

}  // namespace plf
