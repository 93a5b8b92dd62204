// real_plan.cpp -- R2C / C2R plan nodes (docs/plfft.md:111-134).
//
// Even n runs the half-length packed kernels; odd n (or an even n whose
// half length has no kernel) runs the full-length real-input kernels.
#include <cstdint>
#include <sstream>

#include "plan.hpp"

namespace plf {

std::vector<double> root_table(int64_t n, int64_t count) {
  std::vector<double> t(2 * (size_t)count);
  for (int64_t k = 0; k < count; ++k) {
    double re, im;
    unit_root(k, n, &re, &im);
    t[2 * k] = re;
    t[2 * k + 1] = -im;
  }
  return t;
}

class RealNode : public Node {
 public:
  RealNode(bool r2c, bool half, Choice c, int64_t n, int prec, int64_t batch, int64_t is,
           int64_t id, int64_t os, int64_t od)
      : r2c_(r2c), half_(half), c_(std::move(c)), n_(n), prec_(prec), batch_(batch), is_(is),
        id_(id), os_(os), od_(od) {}
  int64_t len() const { return half_ ? n_ / 2 : n_; }
  hipError_t init() {
    hipError_t err = eng_.init(c_, len());
    if (err != hipSuccess || !half_) return err;
    return upload_table(aux_, root_table(n_, n_ / 2 + 1), prec_);
  }
  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    if (batch_ <= 0) return hipSuccess;
    const size_t rb = (size_t)prec_real_bytes(prec_);
    KArgs a{};
    a.in = in;
    a.out = out;
    a.tw = eng_.tw.p;
    a.desc = eng_.d();
    a.aux0 = aux_.p;
    a.count = batch_;
    a.istride = is_;
    a.idist = id_;
    a.ostride = os_;
    a.odist = od_;
    a.n = len();
    if (half_) {
      // real side read/written as complex pairs: unit stride, even distance,
      // pair-aligned base
      const void *rp = r2c_ ? in : out;
      const int64_t rs = r2c_ ? is_ : os_, rd = r2c_ ? id_ : od_;
      const bool ok = rs == 1 && (batch_ == 1 || rd % 2 == 0) &&
                      ((uintptr_t)rp % (2 * rb)) == 0;
      a.flags = ok ? 1 : 0;
    }
    const int64_t grid = (batch_ + c_.e->tpb - 1) / c_.e->tpb;
    return c_.e->launch(a, grid, c_.e->lds_bytes(len()), s);
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (r2c_ ? "r2c" : "c2r") << " n=" << n_ << " " << chain_str(c_)
      << " batch=" << batch_ << " is=" << is_ << " id=" << id_ << " os=" << os_
      << " od=" << od_ << ")";
    return o.str();
  }

 private:
  bool r2c_, half_;
  Choice c_;
  int64_t n_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_;
  EngineData eng_;
  DevMem aux_;
};

std::unique_ptr<Node> build_real_node(const Problem &p, std::string *why) {
  const bool r2c = p.kind == K_R2C;
  const int64_t id = p.howmany > 1 ? p.idist : 0;
  const int64_t od = p.howmany > 1 ? p.odist : 0;
  Choice c;
  bool half = false;
  if (p.n % 2 == 0) {
    c = choose(r2c ? KK_R2C_E : KK_C2R_E, p.n / 2, p.prec, false);
    half = c.e != nullptr;
  }
  if (!c.e) c = choose(r2c ? KK_R2C_F : KK_C2R_F, p.n, p.prec, false);
  if (!c.e) {
    *why = std::string("no GPU kernel for ") + (r2c ? "R2C" : "C2R") + " n=" +
           std::to_string(p.n) + " " + prec_name(p.prec);
    return nullptr;
  }
  auto node = std::make_unique<RealNode>(r2c, half, c, p.n, p.prec, p.howmany, p.istride, id,
                                         p.ostride, od);
  hipError_t err = node->init();
  if (err != hipSuccess) {
    *why = std::string("device upload failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  return node;
}

}  // namespace plf
