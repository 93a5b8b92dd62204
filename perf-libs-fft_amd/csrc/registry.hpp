// registry.hpp -- host-side table of compiled kernel configurations.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernel_args.hpp"

namespace plf {

// Storage precisions of the device kernels (not the public enum values).
enum Prec : int { P_F32 = 0, P_F64 = 1, P_F16 = 2, P_BF16 = 3 };

inline int prec_real_bytes(int p) { return p == P_F64 ? 8 : (p == P_F32 ? 4 : 2); }
inline int prec_compute_bytes(int p) { return p == P_F64 ? 8 : 4; }
inline const char *prec_name(int p) {
  return p == P_F32 ? "f32" : p == P_F64 ? "f64" : p == P_F16 ? "f16" : "bf16";
}

// Kernel families (see kernels.hpp).
enum KKind : int {
  KK_C2C = 0,
  KK_COL = 1,
  KK_ROW = 2,
  KK_R2C_E = 3,  // R2C, even n, engine length n/2
  KK_C2R_E = 4,  // C2R, even n, engine length n/2
  KK_R2C_F = 5,  // R2C, full-length engine
  KK_C2R_F = 6,  // C2R, full-length engine
  KK_BLUE = 7,   // fused Bluestein, engine length M >= 2n-2
  KK_RADER = 8,  // fused Rader, engine length n-1
  KK_DCT2_E = 9,   // DCT-II (bwd slot: DST-II), even n, engine length n/2
  KK_DCT2_F = 10,  // DCT-II / DST-II, full-length engine
  KK_DCT3_E = 11,  // DCT-III (bwd slot: DST-III), even n, engine length n/2
  KK_DCT3_F = 12,  // DCT-III / DST-III, full-length engine
  KK_C2C_UNIT = 13,  // batched C2C, unit element stride, streaming access
  KK_COL_R2C = 14,   // R2C of even n, half length M = m1*m2: four-step column pass
  KK_ROW_R2C = 15,   //   row pass with the fused post-twiddle
  KK_COL_C2R = 16,   // C2R of even n: column pass with the fused pre-twiddle
  KK_ROW_C2R = 17,   //   row pass storing real pairs
  KK_CZ_COL_IN = 18,   // chirp-z in three passes: chirp-in column pass
  KK_CZ_ROW_MID = 19,  //   row pass: DFT, x B, first inverse DFT
  KK_CZ_COL_OUT = 20,  //   chirp-out column pass
  KK_RD_COL_IN = 21,   // Rader in three passes: permuted-input column pass
  KK_RD_COL_OUT = 22,  //   permuted-output column pass (rows: KK_CZ_ROW_MID)
};
inline const char *kkind_name(int k) {
  static const char *names[] = {"c2c",       "col",       "row",       "r2c_half", "c2r_half",
                                "r2c_full",  "c2r_full",  "bluestein", "rader",    "dct2_half",
                                "dct2_full", "dct3_half", "dct3_full", "c2c_unit", "col_r2c",
                                "row_r2c",   "col_c2r",   "row_c2r",  "cz_col_in", "cz_row_mid",
                                "cz_col_out", "rd_col_in", "rd_col_out"};
  return (k >= 0 && k <= 22) ? names[k] : "?";
}

using KLaunch = hipError_t (*)(const KArgs &, int64_t grid, size_t lds, hipStream_t);

// One compiled kernel. n > 0: compile-time chain of length n; n == 0:
// runtime chain (dynamic engine) with at most pts points per thread.
struct KEntry {
  int kind;
  int prec;
  bool bwd;
  int n;
  std::vector<int> chain;
  int tpt, tpb, pts;
  bool uses_lds;
  int pad = 5;  // LDS padding shift of the engine
  int lds_scalars = 2;  // LDS scalars per element: 2, or 1 for the split exchange
  int persist = 0;      // > 0: persistent launch with at most this many blocks per CU
  int waves = 0;        // > 0: compiled for at least this many waves per SIMD
  bool xlane = false;   // cross-lane kernel: packed members only, lane twiddle table
  bool mfma = false;    // first stage on the matrix cores
  bool packed = false;  // members contiguous at distance n only
  bool staged = false;  // one member per thread, global access staged through LDS
  KLaunch launch;
  std::string name() const;
  // LDS bytes of one block for engine length len
  size_t lds_bytes(int64_t len) const {
    if (!uses_lds) return 0;
    return (size_t)tpb * (size_t)(len + (len >> pad) + 1) * lds_scalars * prec_compute_bytes(prec);
  }
  // can the dynamic engine run this radix chain within pts points/thread
  // (and with the butterflies this configuration compiles)?
  bool fits(int64_t len, const std::vector<int> &ch) const;
  // dynamic entries: radix 17/19/23 butterflies present (<= 256-thread blocks)
  bool big_primes() const { return n == 0 && tpt <= 256 && tpb <= 16 && tpt * tpb <= 256; }
};

const std::vector<KEntry> &kernel_registry();
void register_kernels_f32(std::vector<KEntry> &r);
void register_kernels_f32_4s(std::vector<KEntry> &r);
void register_kernels_f32_unit(std::vector<KEntry> &r);
void register_kernels_f32_real(std::vector<KEntry> &r);
void register_kernels_f64_real(std::vector<KEntry> &r);
void register_kernels_f32_real4s(std::vector<KEntry> &r);
void register_kernels_f64_real4s(std::vector<KEntry> &r);
void register_kernels_conv(std::vector<KEntry> &r);
void register_kernels_r2r(std::vector<KEntry> &r);
void register_kernels_conv_f32(std::vector<KEntry> &r);
void register_kernels_conv_f64(std::vector<KEntry> &r);
void register_kernels_conv_f16(std::vector<KEntry> &r);
void register_kernels_conv_bf16(std::vector<KEntry> &r);
void register_kernels_f16_real(std::vector<KEntry> &r);
void register_kernels_bf16_real(std::vector<KEntry> &r);

// Element-wise chirp-z helpers for lengths beyond one LDS tile
// (kernels_conv.hip); `prec` selects the storage type.
hipError_t launch_chirp_in(int prec, bool bwd, const KArgs &a, hipStream_t s);
hipError_t launch_pointwise_swap(int prec, const KArgs &a, hipStream_t s);
hipError_t launch_chirp_out(int prec, bool bwd, const KArgs &a, hipStream_t s);
// Real <-> complex staging (kernels_conv.hip): which = 0 widen, 1 narrow,
// 2 hermitian, 3 realpart, 4 dct pre, 5 dct post, 6 r2r pre, 7 r2r post
// (the other R2R kinds); `total` threads.
hipError_t launch_real_stage(int which, int prec, const KArgs &a, int64_t total, hipStream_t s);
void register_kernels_f64(std::vector<KEntry> &r);
void register_kernels_f32_cz(std::vector<KEntry> &r);
void register_kernels_f64_cz(std::vector<KEntry> &r);
void register_kernels_f16(std::vector<KEntry> &r);
void register_kernels_bf16(std::vector<KEntry> &r);

// Radices with a butterfly in the dynamic engine.
inline bool dynamic_radix_ok(int r) {
  switch (r) {
    case 2: case 3: case 4: case 5: case 7: case 8: case 11: case 13: case 16:
    case 17: case 19: case 23:
      return true;
  }
  return false;
}

// Fixed-point (Q0.15 / Q0.7) kernels: thread tiers of k_fixed, largest n
// each handles in one pass through LDS.
struct FixedTier {
  int tpt, tpb, pts;
  int64_t max_n;
};
constexpr int kNumFixedTiers = 5;
// launch_fixed tier of the n = 4096 kernel on a compile-time chain (unit
// strides, members at distance idist / odist)
constexpr int kFixedTier4096 = 5;
extern const FixedTier kFixedTiers[kNumFixedTiers];
hipError_t launch_fixed(int bits, int tier, const KArgs &a, int64_t grid, size_t lds,
                        hipStream_t s);
// which: 0 widen real -> complex, 1 bins 0..n/2 out, 2 Hermitian extension,
// 3 real parts out
hipError_t launch_fixed_elem(int bits, int which, const KArgs &a, int64_t total, hipStream_t s);

}  // namespace plf
