// kernel_args.hpp -- POD argument block passed by value to every kernel.
#pragma once

#include <stdint.h>

namespace plf {

// Runtime radix chain for the dynamic engine (stage s has radix[s],
// L[s] = product of the earlier radices, its twiddles start at twoff[s]).
constexpr int kMaxStages = 24;
struct GenericDesc {
  int n;
  int ns;
  int radix[kMaxStages];
  int L[kMaxStages];
  int twoff[kMaxStages];
};

// One argument block for all kernels; each kernel documents the fields it
// reads. Strides and distances count logical elements.
struct KArgs {
  const void *in;
  void *out;
  const void *tw;            // Stockham twiddles (compute precision)
  const GenericDesc *desc;   // dynamic-engine chain (device memory) or null
  const void *aux0;          // kernel-specific tables (four-step twiddles, chirps, ...)
  const void *aux1;
  const void *aux2;          // per-member records of a gathered launch, or null
  void *aux3;                // writable per-member slots (Rader: x_0 and A_0), or null
  int64_t count;            // transforms (or columns / rows) in this launch
  int64_t istride, idist;
  int64_t ostride, odist;
  int64_t n;                 // length handled by the engine
  int64_t n1, n2;            // four-step split N = n1 * n2
  int64_t bigdist;           // four-step: workspace distance between big transforms
  int shift;                 // four-step twiddle split: W^e = hi[e >> shift] * lo[e & mask]
  int shift2;                // second split table (real four-step: W_n)
  int flags;
};

// One member of a gathered three-pass chirp-z launch (kernels_cz.hpp):
// offsets in elements of the input / output storage type, the member's
// workspace (compute complex) and its chirp and B tables in aux1.
struct CzMember {
  int64_t n;
  int64_t in_off, out_off, ws_off;
  int64_t chirp_off, B_off;
};

// One transform of a gathered one-block Bluestein launch (kernels_conv.hpp):
// n points read at in_off with unit stride and written at out_off with
// stride ostride (elements of the kernel's storage type), with the chirp
// and B tables at chirp_off and B_off in aux1 and aux0.
struct BlueMember {
  int64_t n;
  int64_t in_off, out_off, ostride;
  int64_t chirp_off, B_off;
};

// One block of a gathered fixed-point launch (kernels_fixed.hpp): `count`
// members from sorted member index `member` on, all of one length, whose
// descriptor is desc-th in the launch's array and whose stage twiddles start
// tw_off table entries into the launch's table. One length per block keeps
// every thread of a block on the same stage sequence (the same barriers).
struct FxBlock {
  int32_t desc, count;
  int64_t tw_off, member;
};

}  // namespace plf
