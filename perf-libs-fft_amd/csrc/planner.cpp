// planner.cpp -- the GPU planner.
//
// Replaces the reference's strategy race (src/planner/planner.cpp:49-182,
// 286-372) with a choice among compiled GPU kernels:
//   * a fused LDS-resident Stockham kernel when one exists for (n, precision);
// further node kinds are added alongside (four-step, Bluestein, Rader, real
// and real-to-real wrappers).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>

#include "plan.hpp"

namespace plf {

// ---------------------------------------------------------------------------
// device memory helper
// ---------------------------------------------------------------------------
DevMem::~DevMem() {
  if (p) (void)hipFree(p);
}
hipError_t DevMem::alloc(size_t b) {
  bytes = b;
  if (b == 0) return hipSuccess;
  return hipMalloc(&p, b);
}
hipError_t DevMem::upload(const void *host, size_t b) {
  hipError_t e = alloc(b);
  if (e != hipSuccess) return e;
  if (b == 0) return hipSuccess;
  return hipMemcpy(p, host, b, hipMemcpyHostToDevice);
}

// ---------------------------------------------------------------------------
// registry
// ---------------------------------------------------------------------------
std::string FusedEntry::name() const {
  std::ostringstream o;
  o << "fused_" << prec_name(prec) << (bwd ? "_bwd" : "_fwd") << "_n" << n << "_r";
  for (size_t i = 0; i < chain.size(); ++i) o << (i ? "x" : "") << chain[i];
  o << "_t" << tpt << "_b" << tpb;
  return o.str();
}

const std::vector<FusedEntry> &fused_registry() {
  static const std::vector<FusedEntry> reg = [] {
    std::vector<FusedEntry> r;
    register_fused_f32(r);
    register_fused_f64(r);
    register_fused_f16(r);
    register_fused_bf16(r);
    return r;
  }();
  return reg;
}

static std::vector<const FusedEntry *> fused_candidates(int64_t n, int prec, bool bwd) {
  std::vector<const FusedEntry *> out;
  for (const auto &e : fused_registry())
    if (e.n == n && e.prec == prec && e.bwd == bwd) out.push_back(&e);
  return out;
}

// Upload a FP64 twiddle table converted to the compute precision.
static hipError_t upload_twiddles(DevMem &dst, const std::vector<double> &tw, int prec) {
  if (prec_compute_bytes(prec) == 8) return dst.upload(tw.data(), tw.size() * sizeof(double));
  std::vector<float> f(tw.size());
  for (size_t i = 0; i < tw.size(); ++i) f[i] = (float)tw[i];
  return dst.upload(f.data(), f.size() * sizeof(float));
}

// ---------------------------------------------------------------------------
// fused C2C node
// ---------------------------------------------------------------------------
class FusedC2CNode : public Node {
 public:
  FusedC2CNode(const FusedEntry *e, int64_t batch, int64_t is, int64_t id,
               int64_t os, int64_t od)
      : e_(e), batch_(batch), is_(is), id_(id), os_(os), od_(od) {}
  hipError_t init() { return upload_twiddles(tw_, stockham_twiddles(e_->chain), e_->prec); }
  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    if (batch_ <= 0) return hipSuccess;
    FusedArgs a{in, out, tw_.p, batch_, is_, id_, os_, od_};
    const int64_t grid = (batch_ + e_->tpb - 1) / e_->tpb;
    return e_->launch(a, grid, s);
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << e_->name() << " batch=" << batch_ << " is=" << is_ << " id=" << id_
      << " os=" << os_ << " od=" << od_ << ")";
    return o.str();
  }

 private:
  const FusedEntry *e_;
  DevMem tw_;
  int64_t batch_, is_, id_, os_, od_;
};

static std::unique_ptr<Node> build_c2c(const Problem &p, std::string *why) {
  const bool bwd = p.dir > 0;
  auto cands = fused_candidates(p.n, p.prec, bwd);
  if (!cands.empty()) {
    auto node = std::make_unique<FusedC2CNode>(cands.front(), p.howmany, p.istride,
                                               p.howmany > 1 ? p.idist : 0, p.ostride,
                                               p.howmany > 1 ? p.odist : 0);
    hipError_t e = node->init();
    if (e != hipSuccess) {
      *why = std::string("twiddle upload failed: ") + hipGetErrorString(e);
      return nullptr;
    }
    return node;
  }
  *why = "no GPU kernel for C2C n=" + std::to_string(p.n) + " " + prec_name(p.prec);
  return nullptr;
}

std::unique_ptr<Node> build_node(const Problem &p, std::string *why) {
  switch (p.kind) {
    case K_C2C:
      return build_c2c(p, why);
    default:
      *why = "transform kind not implemented on the GPU yet";
      return nullptr;
  }
}

}  // namespace plf
