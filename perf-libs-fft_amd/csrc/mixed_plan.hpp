// mixed_plan.hpp -- plan of a batch of C2C transforms of different lengths,
// run as at most two tile launches (one per dependency level), one launch
// per group of one-block Bluestein transforms and three launches per group
// of chirp-z members.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "mixed_args.hpp"
#include "plan.hpp"

namespace plf {

class MixedPlan {
 public:
  // Member i transforms n[i] complex elements read at in_off[i] and written
  // at out_off[i] (complex elements of the storage type). A null offset
  // array packs the members back to back. dir: -1 forward, +1 backward.
  // fixed_bits 15 / 7: Q0.15 / Q0.7 storage (power-of-two n <= 32768, one
  // launch per kernel tier); 0: floating point (prec).
  static std::unique_ptr<MixedPlan> build(int prec, int fixed_bits, int dir, int64_t count,
                                          const int64_t *n, const int64_t *in_off,
                                          const int64_t *out_off, std::string *why);

  // Launch the tile kernel on `s` once per non-empty dependency level (at
  // most two launches), then one launch per Bluestein group and three per
  // chirp-z group. `ws` holds workspace_bytes() of device memory.
  hipError_t enqueue(const void *in, void *out, void *ws, hipStream_t s) const;
  size_t workspace_bytes() const { return data_bytes_; }
  std::string describe() const { return summary_; }
  int prec() const { return prec_; }
  int64_t count() const { return count_; }
  int64_t in_span() const { return in_span_; }    // complex elements
  int64_t out_span() const { return out_span_; }  // complex elements
  int grid() const { return grid_; }
  size_t lds_bytes() const { return lds_; }

 private:
  int prec_ = P_F32;
  int bwd_ = 0;
  int64_t count_ = 0;
  int64_t in_span_ = 0, out_span_ = 0;
  int grid_ = 0;  // resident blocks (largest grid of a launch)
  int level_grid_[2] = {0, 0};  // resident blocks of each level's kernel
  int64_t level_first_[2] = {0, 0}, level_count_[2] = {0, 0};
  size_t lds_ = 0, data_bytes_ = 0;
  DevMem xf_, tasks_, descs_, desc_tw_, tw_, aux_;
  // chirp-z members on the three-pass compiled kernels (kernels_cz.hpp),
  // one group per padded length M: three launches each
  struct CzGroup {
    Choice cin, cmid, cout;
    EngineData ein, emid, eout;
    int64_t n1 = 0, n2 = 0, count = 0, split_off = 0;
    int shift = 0;
    DevMem members;  // CzMember records
  };
  std::vector<std::unique_ptr<CzGroup>> cz_groups_;
  // one-block Bluestein transforms on the compiled fused kernel, one launch
  // per (padded length M, source): whole members read from the input, or
  // the rows of Cooley-Tukey members read from the workspace
  struct BlueGroup {
    Choice c;
    EngineData eng;
    int64_t M = 0, count = 0;
    bool from_ws = false;
    DevMem members;  // BlueMember records
  };
  std::vector<std::unique_ptr<BlueGroup>> blue_groups_;
  std::vector<std::unique_ptr<Node>> fixed_groups_;  // fixed point: one gathered node
  std::string summary_;
};

}  // namespace plf
