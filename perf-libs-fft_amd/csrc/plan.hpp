// plan.hpp -- GPU plan tree for the 1D transform path.
//
// A plan is a tree of Nodes; each Node enqueues its kernels on a HIP stream.
// Device resources (twiddle tables, chirps) are owned by the nodes and live
// on the device that was current at plan creation. Workspace memory is passed
// in by the caller so that one plan can run concurrently on several streams.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "registry.hpp"

namespace plf {

enum Kind : int { K_C2C = 0, K_R2C = 1, K_C2R = 2, K_R2R = 3 };

// The validated problem, in the reference's terms (src/planner/problem.hpp:
// 22-58): kind, precision, n, strides, batch, distances, direction.
struct Problem {
  int kind = K_C2C;
  int dir = -1;       // -1 forward, +1 backward (C2C)
  int r2r_kind = 0;   // plfft_r2r_kind_t for K_R2R
  int prec = P_F32;   // storage precision (floating types)
  int fixed_bits = 0; // 15 or 7 for Q15 / Q7 storage, 0 otherwise
  int64_t n = 1, howmany = 1;
  int64_t istride = 1, idist = 0, ostride = 1, odist = 0;
  bool may_alias = false;
  int rigor = 1;
  // 1: plan the adjoint of the transform described above (input/output
  // layouts already swapped by the caller)
  int adjoint = 0;
  // 1: single-kernel or four-step plans only (the padded inner transform of
  // a chirp-z plan), so that chirp-z never nests
  int no_conv = 0;
};

// Owning device allocation.
struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem();
  hipError_t alloc(size_t b);
  hipError_t upload(const void *host, size_t b);
};

class Node {
 public:
  virtual ~Node() = default;
  // Enqueue the transform of `in` into `out` on `s`. `work` holds at least
  // workspace_bytes() of device memory (may be null when that is 0).
  virtual hipError_t enqueue(const void *in, void *out, void *work,
                             hipStream_t s) const = 0;
  virtual size_t workspace_bytes() const { return 0; }
  // Enqueue with only `cap` bytes of workspace. Nodes that stream their
  // batch through the workspace in chunks shrink the chunk to fit; the
  // others need workspace_bytes() <= cap. hipErrorInvalidValue (nothing
  // launched) when it cannot fit.
  virtual hipError_t enqueue_capped(const void *in, void *out, void *work, size_t cap,
                                    hipStream_t s) const {
    if (workspace_bytes() > cap) return hipErrorInvalidValue;
    return enqueue(in, out, work, s);
  }
  // Smallest `cap` that enqueue_capped accepts.
  virtual size_t min_workspace_bytes() const { return workspace_bytes(); }
  virtual std::string describe() const = 0;
};

// Thrown-free failure reporting for planning.
struct PlanError {
  std::string msg;
};

// Builds the node for a validated problem; returns null (and a reason) when
// no GPU implementation exists.
std::unique_ptr<Node> build_node(const Problem &p, std::string *why);
std::unique_ptr<Node> build_c2c_node(const Problem &p, std::string *why);
std::unique_ptr<Node> build_real_node(const Problem &p, std::string *why);
std::unique_ptr<Node> build_r2r_node(const Problem &p, std::string *why);
// Q0.15 / Q0.7 problems (p.fixed_bits != 0): power-of-two n only.
std::unique_ptr<Node> build_fixed_node(const Problem &p, std::string *why);
// Fixed-point C2C members of power-of-two lengths n[i] at explicit offsets
// (complex elements, unit element stride): one gathered launch per kernel
// tier; *launches receives their number.
std::unique_ptr<Node> build_fixed_mixed_node(int bits, bool bwd, const std::vector<int64_t> &n,
                                             const std::vector<int64_t> &in_off,
                                             const std::vector<int64_t> &out_off,
                                             int *launches, std::string *why);

// ---- planner internals shared by the node builders ----------------------
// A kernel choice for one engine length: the registry entry plus the chain.
struct Choice {
  const KEntry *e = nullptr;
  std::vector<int> chain;
};
// Compile-time chain first (registry order = preference), else the dynamic
// entry with the fewest threads that fits; e == null if none.
// packed: members contiguous at distance n with unit stride (cross-lane
// kernels need it)
Choice choose(int kind, int64_t len, int prec, bool bwd, int need_tpb_div = 0,
              bool packed = false);
std::vector<int> dynamic_chain(int64_t n);
std::string chain_str(const Choice &c);
// Upload a FP64 table (interleaved re/im) in the compute precision.
hipError_t upload_table(DevMem &dst, const std::vector<double> &v, int prec);

// Device-side engine data for a chain: Stockham twiddles and, for the
// dynamic engine, its descriptor.
struct EngineData {
  DevMem tw, desc;
  hipError_t init(const Choice &c, int64_t len);
  const GenericDesc *d() const { return static_cast<const GenericDesc *>(desc.p); }
};

// Autotuner (autotune.cpp): index of the candidate to use among entries for
// one (kind, n, precision, direction); times them under MEASURE or above,
// caches the winner in memory and in the wisdom file. -1 when the
// PLFFT_KERNEL_FILTER diagnostic excludes every candidate.
int tune_pick(int kind, int64_t n, int prec, bool bwd, const std::vector<Choice> &cands);
// Starts the planning clock of the calling thread with its rigor.
void set_planning_rigor(int rigor);
int planning_rigor();
// The reference's planning budget in seconds and decision margin per rigor
// (src/plfft.cpp:149-184); plfft_rigor_t values.
double planner_budget_secs(int rigor, int64_t n);
double planner_margin(int rigor);
// Wisdom key of a kernel choice: family, length, precision, direction and
// the device (architecture name / CU count).
std::string wisdom_key(int kind, int64_t n, int prec, bool bwd);

// W_n^k (forward sign) for k in [0, count), interleaved re/im, FP64.
std::vector<double> root_table(int64_t n, int64_t count);
// Split table of W_n^e for e in [0, count): lo = W_n^r for r < 2^shift, then
// hi = W_n^{q 2^shift} for q <= (count-1) >> shift, so that W_n^e =
// hi[e >> shift] * lo[e & (2^shift - 1)]. Forward sign, interleaved, FP64.
std::vector<double> split_root_table(int64_t n, int64_t count, int *shift);
// Big transforms per column/row pass of a four-step plan of length n (the
// workspace holds that many; about 1 GiB unless PLFFT_FOURSTEP_CHUNK_MB says).
int64_t fourstep_chunk(int64_t n, int prec, int64_t batch);

// Three-pass chirp-z (kernels_cz.hpp) for length n: the smallest padded
// length M >= 2n-2 of the form 2^k or 3 * 2^k with compiled kernels for a
// split M = n1 * n2 (n1 a power of two, n1 <= n2, the most balanced).
struct Cz3Choice {
  int64_t M = 0, n1 = 0;
  Choice cin, cmid, cout;
};
bool cz3_choice(int64_t n, int prec, bool bwd, Cz3Choice *out);

// Twiddle table of a Stockham radix chain, FORWARD sign, FP64:
// for stage s >= 1 with radix R and L = product of earlier radices,
// tw[off_s + (m-1)*L + j] = exp(-2*pi*i*j*m/(L*R)), m in [1,R), j in [0,L).
std::vector<double> stockham_twiddles(const std::vector<int> &chain);
// Lane twiddles of the cross-lane kernels (n = 16, 32, 64): row r for the
// decimation-in-frequency stage of half-size h = n / 2^(r+1) (h >= 2), entry
// j = W_{2h}^{j mod h} where bit h of j is set, else 1.
std::vector<double> xlane_twiddles(int64_t n);
// Operands of the 16-point matrix-core kernel: the real 32 x 32 DFT_16
// matrix [[Wr, -Wi], [Wi, Wr]] in the A layout of v_mfma_f32_32x32x16_f16
// (lane l holds row l % 32, K columns 8 (l / 32) .. + 7), for the real and
// imaginary K halves, as f16 high and low parts: 4 x 64 x 8 f16 bit patterns.
std::vector<uint16_t> mfma_dft16_table();

// Bluestein chirp c_k = exp(-i pi (k^2 mod 2n)/n), k < n, interleaved FP64.
std::vector<double> chirp_table(int64_t n);
// In-place forward DFTs of `count` contiguous length-M rows (interleaved
// FP64) computed on the current device with this library's FP64 plans.
hipError_t dft_f64_gpu(std::vector<double> &v, int64_t M, int64_t count);

// exp(sign*2*pi*i*num/den) with num reduced exactly mod den, long double
// evaluation rounded once to double.
void unit_root(int64_t num, int64_t den, double *re, double *im);

// Last error text for the calling thread (plan creation diagnostics).
void set_last_error(const std::string &s);
const char *last_error();

}  // namespace plf
