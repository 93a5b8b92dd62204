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

using FusedLaunch = hipError_t (*)(const FusedArgs &, int64_t grid, hipStream_t);

// One compiled fused Stockham kernel: length n = product of `chain`, TPT
// threads per transform, TPB transforms per block.
struct FusedEntry {
  int n;
  int prec;
  bool bwd;
  int tpt, tpb, padshift;
  std::vector<int> chain;
  size_t lds_bytes;
  int block;
  FusedLaunch launch;
  std::string name() const;
};

const std::vector<FusedEntry> &fused_registry();

void register_fused_f32(std::vector<FusedEntry> &r);
void register_fused_f64(std::vector<FusedEntry> &r);
void register_fused_f16(std::vector<FusedEntry> &r);
void register_fused_bf16(std::vector<FusedEntry> &r);

}  // namespace plf
