// planner.cpp -- the GPU planner for complex transforms.
//
// Replaces the reference's recursive strategy race (src/planner/planner.cpp:
// 49-182, 286-372) with a choice among compiled GPU kernels, in order:
//   1. one fused LDS kernel over the batch: a compile-time radix chain for
//      the length if one is compiled, else the runtime-chain engine
//      (lengths whose prime factors are <= 23 and that fit in LDS);
//   2. four-step: N = n1*n2, column pass + twiddle into a workspace, row
//      pass with a transposed store (two HBM round trips, batch chunked so
//      the workspace stays small);
//   3. Rader / Bluestein for lengths with larger prime factors
//      (conv_plan.cpp).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
std::string KEntry::name() const {
  std::ostringstream o;
  o << kkind_name(kind) << "_" << prec_name(prec) << (bwd ? "_bwd" : "_fwd");
  if (n) {
    o << "_n" << n << "_r";
    for (size_t i = 0; i < chain.size(); ++i) o << (i ? "x" : "") << chain[i];
  } else {
    o << "_dyn";
  }
  o << "_t" << tpt << "_b" << tpb;
  if (kind == KK_C2C_UNIT) o << "_p" << pad << (pts ? "_pair" : "");
  if (lds_scalars == 1) o << "_split";
  if (persist) o << "_pipe" << persist;
  if (waves) o << "_w" << waves;
  if (xlane) o << "_xlane";
  if (mfma) o << "_mfma";
  if (staged) o << "_staged";
  return o.str();
}

bool KEntry::fits(int64_t len, const std::vector<int> &ch) const {
  if (n) return n == len;
  for (int r : ch) {
    if (!dynamic_radix_ok(r)) return false;
    if (r >= 17 && !big_primes()) return false;
    const int64_t kmax = (pts + r - 1) / r;
    if (len / r > kmax * tpt) return false;
  }
  return lds_bytes(len) <= 160 * 1024;
}

const std::vector<KEntry> &kernel_registry() {
  static const std::vector<KEntry> reg = [] {
    std::vector<KEntry> r;
    register_kernels_f32(r);
    register_kernels_f32_4s(r);
    register_kernels_f32_unit(r);
    register_kernels_f32_real(r);
    register_kernels_f64(r);
    register_kernels_f32_cz(r);
    register_kernels_f64_cz(r);
    register_kernels_f64_real(r);
    register_kernels_f32_real4s(r);
    register_kernels_f64_real4s(r);
    register_kernels_f16_real(r);
    register_kernels_bf16_real(r);
    register_kernels_conv(r);
    register_kernels_r2r(r);
    register_kernels_f16(r);
    register_kernels_bf16(r);
    return r;
  }();
  return reg;
}

// Radix chain over the dynamic engine's butterflies: powers of two as 16s
// (remainder 2/4/8 first), then the odd primes. Empty if another prime
// divides n.
std::vector<int> dynamic_chain(int64_t n) {
  std::vector<int> c;
  int64_t m = n;
  int e = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++e;
  }
  if (e % 4) c.push_back(1 << (e % 4));
  for (int i = 0; i < e / 4; ++i) c.push_back(16);
  static const int odd[] = {3, 5, 7, 11, 13, 17, 19, 23};
  for (int p : odd)
    while (m % p == 0) {
      c.push_back(p);
      m /= p;
    }
  if (m != 1) c.clear();
  return c;
}

// Compile-time chain first (registry order = preference), else the dynamic
// entry with the fewest threads that fits.
Choice choose(int kind, int64_t len, int prec, bool bwd, int need_tpb_div, bool packed) {
  Choice c;
  std::vector<Choice> stat;
  for (const auto &e : kernel_registry()) {
    if (e.kind != kind || e.prec != prec || e.bwd != bwd || e.n != len) continue;
    if ((e.xlane || e.packed) && !packed) continue;
    if (need_tpb_div && need_tpb_div % e.tpb) continue;
    if (e.lds_bytes(len) > 160 * 1024) continue;
    Choice s;
    s.e = &e;
    s.chain = e.chain;
    stat.push_back(s);
  }
  if (!stat.empty()) {
    const int pick = tune_pick(kind, len, prec, bwd, stat);
    if (pick >= 0) return stat[pick];
    if (kind == KK_C2C_UNIT) return c;  // filtered out: the planner falls back
  }
  std::vector<int> ch = len == 1 ? std::vector<int>{} : dynamic_chain(len);
  if (len != 1 && ch.empty()) return c;
  for (const auto &e : kernel_registry()) {
    if (e.kind != kind || e.prec != prec || e.bwd != bwd || e.n != 0) continue;
    if (need_tpb_div && need_tpb_div % e.tpb) continue;
    if (!e.fits(len, ch)) continue;
    if (!c.e || e.tpt < c.e->tpt) {
      c.e = &e;
      c.chain = ch;
    }
  }
  return c;
}

// Upload a FP64 table converted to the compute precision.
hipError_t upload_table(DevMem &dst, const std::vector<double> &v, int prec) {
  if (prec_compute_bytes(prec) == 8) return dst.upload(v.data(), v.size() * sizeof(double));
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
  return dst.upload(f.data(), f.size() * sizeof(float));
}

hipError_t EngineData::init(const Choice &c, int64_t len) {
    if (c.e->xlane) return upload_table(tw, xlane_twiddles(len), c.e->prec);
    if (c.e->mfma && c.e->packed) {
      const std::vector<uint16_t> t = mfma_dft16_table();
      return tw.upload(t.data(), t.size() * sizeof(uint16_t));
    }
    hipError_t err = upload_table(tw, stockham_twiddles(c.chain), c.e->prec);
    if (err != hipSuccess || c.e->n) return err;
    GenericDesc d{};
    d.n = (int)len;
    d.ns = (int)c.chain.size();
    int64_t L = 1, off = 0;
    for (int s = 0; s < d.ns; ++s) {
      d.radix[s] = c.chain[s];
      d.L[s] = (int)L;
      d.twoff[s] = (int)off;
      if (s > 0) off += L * (c.chain[s] - 1);
      L *= c.chain[s];
    }
    return desc.upload(&d, sizeof(d));
}

std::string chain_str(const Choice &c) {
  std::ostringstream o;
  o << c.e->name();
  if (!c.e->n) {
    o << "[";
    for (size_t i = 0; i < c.chain.size(); ++i) o << (i ? "x" : "") << c.chain[i];
    o << "]";
  }
  return o.str();
}

// ---------------------------------------------------------------------------
// batched single-kernel C2C
// ---------------------------------------------------------------------------
class BatchC2CNode : public Node {
 public:
  BatchC2CNode(Choice c, int64_t n, int64_t batch, int64_t is, int64_t id, int64_t os,
               int64_t od)
      : c_(std::move(c)), n_(n), batch_(batch), is_(is), id_(id), os_(os), od_(od) {}
  hipError_t init() { return eng_.init(c_, n_); }
  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    if (batch_ <= 0) return hipSuccess;
    KArgs a{};
    a.in = in;
    a.out = out;
    a.tw = eng_.tw.p;
    a.desc = eng_.d();
    a.count = batch_;
    a.istride = is_;
    a.idist = id_;
    a.ostride = os_;
    a.odist = od_;
    a.n = n_;
    const int64_t grid = (batch_ + c_.e->tpb - 1) / c_.e->tpb;
    return c_.e->launch(a, grid, c_.e->lds_bytes(n_), s);
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << chain_str(c_) << " n=" << n_ << " batch=" << batch_ << " is=" << is_
      << " id=" << id_ << " os=" << os_ << " od=" << od_ << ")";
    return o.str();
  }

 private:
  Choice c_;
  int64_t n_;
  EngineData eng_;
  int64_t batch_, is_, id_, os_, od_;
};

// ---------------------------------------------------------------------------
// four-step C2C: X[k1 + n1*k2] = sum_n2 W_n2^{n2 k2} W_N^{n2 k1}
//                                 sum_n1 x[n2 + n2_len*n1] W_n1^{n1 k1}
// ---------------------------------------------------------------------------
class FourStepNode : public Node {
 public:
  FourStepNode(Choice col, Choice row, int64_t n1, int64_t n2, int prec, int64_t batch,
               int64_t is, int64_t id, int64_t os, int64_t od, int64_t chunk)
      : col_(std::move(col)), row_(std::move(row)), n1_(n1), n2_(n2), prec_(prec),
        batch_(batch), is_(is), id_(id), os_(os), od_(od), chunk_(chunk) {}

  hipError_t init() {
    hipError_t err = ecol_.init(col_, n1_);
    if (err == hipSuccess) err = erow_.init(row_, n2_);
    if (err != hipSuccess) return err;
    const int64_t n = n1_ * n2_;
    int lg = 0;
    while ((int64_t(1) << lg) < n) ++lg;
    shift_ = (lg + 1) / 2;
    const int64_t nlo = int64_t(1) << shift_;
    const int64_t nhi = ((n - 1) >> shift_) + 1;
    std::vector<double> lo(2 * nlo), hi(2 * nhi);
    for (int64_t r = 0; r < nlo; ++r) {
      double re, im;
      unit_root(r, n, &re, &im);
      lo[2 * r] = re;
      lo[2 * r + 1] = -im;
    }
    for (int64_t q = 0; q < nhi; ++q) {
      double re, im;
      unit_root(q * nlo, n, &re, &im);
      hi[2 * q] = re;
      hi[2 * q + 1] = -im;
    }
    err = upload_table(lo_, lo, prec_);
    if (err == hipSuccess) err = upload_table(hi_, hi, prec_);
    return err;
  }
  size_t per_bytes() const { return (size_t)(n1_ * n2_) * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return (size_t)chunk_ * per_bytes(); }
  size_t min_workspace_bytes() const override { return per_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    return run(in, out, work, chunk_, s);
  }
  hipError_t enqueue_capped(const void *in, void *out, void *work, size_t cap,
                            hipStream_t s) const override {
    const int64_t fit = (int64_t)(cap / per_bytes());
    if (fit < 1) return hipErrorInvalidValue;
    return run(in, out, work, std::min(chunk_, fit), s);
  }
  // column pass then row pass over `chunk` big transforms at a time
  hipError_t run(const void *in, void *out, void *work, int64_t chunk, hipStream_t s) const {
    const int64_t n = n1_ * n2_;
    const size_t ie = 2 * (size_t)prec_real_bytes(prec_);
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk) {
      const int64_t cb = std::min(chunk, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * ie;
      a.out = work;
      a.tw = ecol_.tw.p;
      a.desc = ecol_.d();
      a.aux0 = lo_.p;
      a.aux1 = hi_.p;
      a.count = cb * n2_;
      a.istride = is_;
      a.idist = id_;
      a.n = n1_;
      a.n1 = n1_;
      a.n2 = n2_;
      a.bigdist = n;
      a.shift = shift_;
      hipError_t err = col_.e->launch(a, (a.count + col_.e->tpb - 1) / col_.e->tpb,
                                      col_.e->lds_bytes(n1_), s);
      if (err != hipSuccess) return err;
      KArgs b{};
      b.in = work;
      b.out = static_cast<char *>(out) + (size_t)(b0 * od_) * ie;
      b.tw = erow_.tw.p;
      b.desc = erow_.d();
      b.count = cb * n1_;
      b.ostride = os_;
      b.odist = od_;
      b.n = n2_;
      b.n1 = n1_;
      b.n2 = n2_;
      b.bigdist = n;
      err = row_.e->launch(b, (b.count + row_.e->tpb - 1) / row_.e->tpb,
                           row_.e->lds_bytes(n2_), s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(fourstep n=" << n1_ * n2_ << " = " << n1_ << "x" << n2_ << " chunk=" << chunk_
      << " batch=" << batch_ << " " << chain_str(col_) << " " << chain_str(row_) << ")";
    return o.str();
  }

 private:
  Choice col_, row_;
  int64_t n1_, n2_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  EngineData ecol_, erow_;
  DevMem lo_, hi_;
  int shift_ = 0;
};

// Chunk of big transforms per column/row pair: keeps the workspace near
// `mb` MiB (env PLFFT_FOURSTEP_CHUNK_MB overrides the default of 1024). On
// MI355X the intermediate is not served from the Infinity Cache at any chunk
// size measured (32 MiB .. 1 GiB); fewer, larger launches win (65536-point
// batch of 1 GiB: 2.22 TB/s at 128 MiB, 2.48 TB/s at 1 GiB).
int64_t fourstep_chunk(int64_t n, int prec, int64_t batch) {
  int64_t mb = 1024;
  if (const char *s = std::getenv("PLFFT_FOURSTEP_CHUNK_MB")) mb = std::max<int64_t>(1, atoll(s));
  const int64_t per = n * 2 * prec_compute_bytes(prec);
  int64_t c = std::max<int64_t>(1, (mb << 20) / per);
  return std::min(c, std::max<int64_t>(1, batch));
}

std::unique_ptr<Node> build_bluestein(const Problem &p, std::string *why);

std::unique_ptr<Node> build_c2c_node(const Problem &p, std::string *why) {
  const bool bwd = p.dir > 0;
  const int64_t id = p.howmany > 1 ? p.idist : 0;
  const int64_t od = p.howmany > 1 ? p.odist : 0;
  // 1. one fused kernel (the streaming unit-stride variant when it applies)
  Choice c;
  const bool packed = p.istride == 1 && p.ostride == 1 &&
                      (p.howmany <= 1 || (p.idist == p.n && p.odist == p.n));
  if (p.istride == 1 && p.ostride == 1) c = choose(KK_C2C_UNIT, p.n, p.prec, bwd, 0, packed);
  if (!c.e) c = choose(KK_C2C, p.n, p.prec, bwd);
  if (c.e) {
    auto node = std::make_unique<BatchC2CNode>(c, p.n, p.howmany, p.istride, id, p.ostride, od);
    hipError_t err = node->init();
    if (err != hipSuccess) {
      *why = std::string("device upload failed: ") + hipGetErrorString(err);
      return nullptr;
    }
    return node;
  }
  // 2. four-step over the most balanced split with kernels for both sides
  int64_t best1 = 0;
  Choice bc, br;
  for (int64_t n1 = 2; n1 * n1 <= p.n * 4 && n1 < p.n; ++n1) {
    if (p.n % n1) continue;
    const int64_t n2 = p.n / n1;
    Choice cc = choose(KK_COL, n1, p.prec, bwd);
    if (!cc.e) continue;
    Choice rr = choose(KK_ROW, n2, p.prec, bwd);
    if (!rr.e) continue;
    const double bal = std::fabs(std::log((double)n1 / (double)n2));
    const bool stat = cc.e->n && rr.e->n;
    const bool best_stat = bc.e && bc.e->n && br.e->n;
    const double best_bal =
        best1 ? std::fabs(std::log((double)best1 / (double)(p.n / best1))) : 1e30;
    if (!best1 || (stat && !best_stat) || (stat == best_stat && bal < best_bal)) {
      best1 = n1;
      bc = cc;
      br = rr;
    }
  }
  if (best1) {
    const int64_t chunk = fourstep_chunk(p.n, p.prec, p.howmany);
    auto node = std::make_unique<FourStepNode>(bc, br, best1, p.n / best1, p.prec, p.howmany,
                                               p.istride, id, p.ostride, od, chunk);
    hipError_t err = node->init();
    if (err != hipSuccess) {
      *why = std::string("device upload failed: ") + hipGetErrorString(err);
      return nullptr;
    }
    return node;
  }
  // 3. chirp-z / Rader
  if (p.no_conv) {
    *why = "no single-kernel or four-step plan for the padded length " + std::to_string(p.n) +
           " of a chirp-z transform (" + prec_name(p.prec) + " four-step covers up to " +
           (prec_compute_bytes(p.prec) == 8 ? "2^22" : "2^24") + " points)";
    return nullptr;
  }
  return build_bluestein(p, why);
}

std::unique_ptr<Node> build_node(const Problem &p, std::string *why) {
  if (p.fixed_bits) return build_fixed_node(p, why);
  switch (p.kind) {
    case K_C2C:
      return build_c2c_node(p, why);
    case K_R2C:
    case K_C2R:
      return build_real_node(p, why);
    case K_R2R:
      return build_r2r_node(p, why);
    default:
      *why = "transform kind not implemented on the GPU yet";
      return nullptr;
  }
}


}  // namespace plf
