// kernels_f32_unit.hip -- FP32 unit-stride streaming C2C configurations
// (first entry per length is the default; the autotuner may pick another).
#include "kernels_xlane.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_f32_unit(std::vector<KEntry> &r) {
  using S = float;
  constexpr int P = P_F32;
  // n = 16, 32 with packed members: butterflies across lanes (DPP,
  // ds_swizzle), no LDS exchange. At 2^27 points: n = 16 0.35 ms (LDS
  // kernel 0.84), n = 32 0.36 ms (0.47); n = 64 measured 0.41 ms against
  // 0.39 for the LDS radix-8x8 kernel, so it is not registered.
  PLF_XLANE(S, P, 16, 2, 2, 2, 2);
  PLF_XLANE(S, P, 32, 2, 2, 2, 2, 2);
  // small odd primes with packed members: LDS-staged global access
  PLF_STAGED(S, P, 3);
  PLF_STAGED(S, P, 5);
  PLF_STAGED(S, P, 7);
  PLF_STAGED(S, P, 11);
  PLF_STAGED(S, P, 13);
  PLF_STAGED(S, P, 17);
  PLF_STAGED(S, P, 19);
  PLF_STAGED(S, P, 23);
  //              TPT  TPB PAD PAIR   radix chain
  PLF_UNIT(S, P,  16,  16, 4, false, 16, 16);
  PLF_UNIT(S, P,   8,  32, 4, true,  16, 16);
  PLF_UNIT(S, P,  64,   4, 4, false, 16, 16, 4);
  PLF_UNIT(S, P,  32,   8, 4, true,  16, 16, 4);
  PLF_UNIT(S, P, 128,   2, 4, true,  8, 8, 8, 4);
  // 4096: the first entry is the ESTIMATE choice and the incumbent of a timed
  // race; at 2^27 points the p5/p4 plain-exchange kernels measured 0.70 of
  // the HBM peak, the split exchange 0.68, the pipelined ones 0.61-0.65 and
  // the 8-point-butterfly and paired ones 0.51-0.54
  PLF_UNIT(S, P, 256,   1, 5, false, 16, 16, 16);
  PLF_UNIT(S, P, 256,   1, 4, false, 16, 16, 16);
  PLF_UNIT(S, P, 128,   2, 4, true,  16, 16, 16);
  PLF_UNIT(S, P, 256,   1, 4, true,  8, 8, 8, 8);
  PLF_UNIT(S, P, 256,   2, 4, true,  8, 8, 8, 8);
  // 16384: the split exchange at 64 registers (two blocks per CU) first:
  // 0.63-0.64 of the HBM peak against 0.62 (one block per CU) and 0.50 for
  // the plain exchange
  PLF_UNIT_SPLIT_OCC(S, P, 1024, 1, 5, 8, 4, 16, 16, 16);
  PLF_UNIT(S, P, 1024,  1, 4, false, 16, 16, 16, 4);
  PLF_UNIT(S, P, 512,   1, 4, true,  16, 16, 16, 4);
  // software-pipelined persistent blocks (next member's loads in flight);
  // measured within a few percent of the plain 4096 kernel, slower for
  // shorter and longer lengths, so only these remain as tuning candidates
  PLF_PIPE(S, P, 256, 4, 4, 16, 16, 16);
  PLF_PIPE(S, P, 256, 4, 3, 16, 16, 16);
  // split (real, then imaginary) LDS exchange: two 16384-point blocks per CU
  PLF_UNIT_SPLIT(S, P, 1024, 1, 5, 16, 16, 16, 4);
  PLF_UNIT_SPLIT(S, P, 1024, 1, 5, 4, 16, 16, 16);
  PLF_UNIT_SPLIT(S, P, 512, 1, 5, 16, 16, 8, 4);
  PLF_UNIT_SPLIT(S, P, 256, 1, 5, 16, 16, 16);
  PLF_UNIT_SPLIT(S, P, 1024, 1, 5, 16, 16, 16, 8);
  // two 16384-point blocks per CU: 64 KB of LDS and 64 registers each
  // (a few registers spill); measured 4240-4420 GB/s against 4000 for the
  // one-block-per-CU kernel. The same cap at 4096 and 8192 points was
  // measured 6-9% slower than their uncapped kernels.
  PLF_UNIT_SPLIT_OCC(S, P, 1024, 1, 5, 8, 16, 16, 16, 4);
}

}  // namespace plf
