// mixed_args.hpp -- host/device records of the mixed-length batch kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernel_args.hpp"

namespace plf {

// Per-member description (device memory).
struct MixedXf {
  int n;         // transform length
  int kind;      // 0 direct, 1 four-step, 2 Bluestein in one block, 3 chirp-z (grouped,
                 // no tiles), 4 four-step whose rows are Bluestein in one block
  int len;       // engine length of a one-block member (n or M); M for chirp-z; kind 4:
                 // the rows' padded length
  int n1, n2;    // four-step split of n (kinds 1, 4)
  int d0, d1;    // engine descriptors: one-block (d0) or column (d0) / row (d1)
  int shift;     // W_N split: N = n1*n2, W^e = hi[e >> shift] * lo[e & mask]
  int lanes0, lanes1;  // tile width of the column / row passes (16 or 8; columns of
                       // Bluestein-row members also 64 or 256)
  int64_t in_off, out_off;  // complex elements
  int64_t ws_off;           // complex elements of workspace (kind 1: n)
  int64_t tw4_off;          // split twiddles in aux: lo (2^shift), then hi
  int64_t chirp_off;        // chirp c_k (n entries) in aux, at the residue of in_off mod 16
                            // (kind 4: the rows' chirp, n2 entries)
  int64_t B_off;            // B = DFT_M(b)/M (M entries) in aux; chirp-z: row-major n1 x n2
};

enum MixedTaskType : int {
  MT_ONE = 0,  // whole member in one pass through LDS
  MT_COL = 1,  // four-step column tile
  MT_ROW = 2,  // four-step row tile
};

struct MixedTask {
  int type;
  int xf;     // member
  int first;  // first column / row of the tile
  int end;    // column tiles: one past the last column
};

struct MixedArgs {
  const void *in;
  void *out;
  void *ws;                // compute-precision complex workspace
  const MixedXf *xf;
  const MixedTask *tasks;  // this launch's tasks, by decreasing predicted cost
  int64_t ntasks;
  const GenericDesc *descs;
  const int64_t *desc_tw;  // twiddle offset (complex entries) of each descriptor
  const void *tw;          // all engine twiddles
  const void *aux;         // split twiddles, chirps, B tables
  int bwd;                 // 1: backward transform (exponent sign +)
};

// Launch the mixed-batch kernel of dependency level `level` (0, 1) for
// storage precision `prec` (Prec values).
hipError_t launch_mixed(int prec, int level, const MixedArgs &a, int grid, size_t lds,
                        hipStream_t s);
// Resident blocks per CU of that kernel with `lds` bytes of LDS.
int mixed_blocks_per_cu(int prec, int level, size_t lds);

}  // namespace plf
