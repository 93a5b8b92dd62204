// real_plan.cpp -- R2C / C2R plan nodes (docs/plfft.md:111-134).
//
// Even n runs the half-length packed kernels; odd n (or an even n whose
// half length has no kernel) runs the full-length real-input kernels.
#include <algorithm>
#include <cmath>
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

std::vector<double> split_root_table(int64_t n, int64_t count, int *shift) {
  int lg = 0;
  while ((int64_t(1) << lg) < count) ++lg;
  *shift = (lg + 1) / 2;
  const int64_t nlo = int64_t(1) << *shift;
  const int64_t nhi = ((count - 1) >> *shift) + 1;
  std::vector<double> t(2 * (size_t)(nlo + nhi));
  for (int64_t r = 0; r < nlo; ++r) unit_root(r, n, &t[2 * r], &t[2 * r + 1]);
  for (int64_t q = 0; q < nhi; ++q)
    unit_root(q * nlo, n, &t[2 * (nlo + q)], &t[2 * (nlo + q) + 1]);
  for (size_t i = 1; i < t.size(); i += 2) t[i] = -t[i];  // forward sign
  return t;
}

class RealNode : public Node {
 public:
  RealNode(bool r2c, bool half, Choice c, int64_t n, int prec, int64_t batch, int64_t is,
           int64_t id, int64_t os, int64_t od, int extra_flags)
      : r2c_(r2c), half_(half), c_(std::move(c)), n_(n), prec_(prec), batch_(batch), is_(is),
        id_(id), os_(os), od_(od), extra_(extra_flags) {}
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
    a.flags |= extra_;
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
  int extra_;
  EngineData eng_;
  DevMem aux_;
};

// Real transform through a complex plan of the same length (lengths with no
// real kernel: large primes, or n beyond one LDS tile). R2C: widen to
// (x, 0), forward C2C, keep bins 0..n/2. C2R: Hermitian extension (swapped),
// forward C2C (= swapped inverse), real part.
class RealViaComplexNode : public Node {
 public:
  RealViaComplexNode(bool r2c, int64_t n, int prec, int64_t batch, int64_t is, int64_t id,
                     int64_t os, int64_t od, int64_t chunk)
      : r2c_(r2c), n_(n), prec_(prec), batch_(batch), is_(is), id_(id), os_(os), od_(od),
        chunk_(chunk) {}
  hipError_t init(std::string *why) {
    Problem q;
    q.kind = K_C2C;
    q.dir = -1;
    q.prec = prec_compute_bytes(prec_) == 8 ? P_F64 : P_F32;
    q.n = n_;
    q.howmany = chunk_;
    q.istride = q.ostride = 1;
    q.idist = q.odist = n_;
    q.may_alias = true;
    inner_ = build_c2c_node(q, why);
    return inner_ ? hipSuccess : hipErrorInvalidValue;
  }
  size_t buf_bytes() const { return (size_t)chunk_ * (size_t)n_ * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return buf_bytes() + inner_->workspace_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    char *buf = static_cast<char *>(work);
    void *iws = buf + buf_bytes();
    const size_t rb = (size_t)prec_real_bytes(prec_);
    const size_t ib = r2c_ ? rb : 2 * rb, ob = r2c_ ? 2 * rb : rb;
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk_) {
      const int64_t cb = std::min(chunk_, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * ib;
      a.out = buf;
      a.count = cb;
      a.istride = is_;
      a.idist = id_;
      a.n = n_;
      a.flags = flags_;
      hipError_t err = launch_real_stage(r2c_ ? 0 : 2, prec_, a, cb * n_, s);
      if (err == hipSuccess) err = inner_->enqueue(buf, buf, iws, s);
      if (err != hipSuccess) return err;
      KArgs c{};
      c.in = buf;
      c.out = static_cast<char *>(out) + (size_t)(b0 * od_) * ob;
      c.count = cb;
      c.ostride = os_;
      c.odist = od_;
      c.n = n_;
      c.flags = flags_;
      err = launch_real_stage(r2c_ ? 1 : 3, prec_, c, cb * (r2c_ ? n_ / 2 + 1 : n_), s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (r2c_ ? "r2c" : "c2r") << "_via_c2c n=" << n_ << " chunk=" << chunk_
      << " batch=" << batch_ << " " << inner_->describe() << ")";
    return o.str();
  }
  int flags_ = 0;

 private:
  bool r2c_;
  int64_t n_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  std::unique_ptr<Node> inner_;
};

// R2C / C2R of even n = 2M through the four-step passes of
// kernels_real4s.hpp: M = m1 * m2 (m1 even), pass A over columns, pass B
// over rows (R2C: pairs of rows k1, m1 - k1 per slot), `chunk` members per
// pass through a workspace of chunk * M complex values.
class RealFourStepNode : public Node {
 public:
  RealFourStepNode(bool r2c, Choice col, Choice row, int64_t m1, int64_t m2, int prec,
                   int64_t batch, int64_t is, int64_t id, int64_t os, int64_t od, int64_t chunk,
                   int extra_flags)
      : r2c_(r2c), col_(std::move(col)), row_(std::move(row)), m1_(m1), m2_(m2), prec_(prec),
        batch_(batch), is_(is), id_(id), os_(os), od_(od), chunk_(chunk), extra_(extra_flags) {}
  hipError_t init() {
    hipError_t err = ecol_.init(col_, m1_);
    if (err == hipSuccess) err = erow_.init(row_, m2_);
    if (err != hipSuccess) return err;
    const int64_t M = m1_ * m2_;
    err = upload_table(wm_, split_root_table(M, M, &shift_m_), prec_);
    if (err == hipSuccess) err = upload_table(wn_, split_root_table(2 * M, M, &shift_n_), prec_);
    return err;
  }
  size_t per_bytes() const { return (size_t)(m1_ * m2_) * 2 * prec_compute_bytes(prec_); }
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
  hipError_t run(const void *in, void *out, void *work, int64_t chunk, hipStream_t s) const {
    if (batch_ <= 0) return hipSuccess;
    const int64_t M = m1_ * m2_;
    const size_t rb = (size_t)prec_real_bytes(prec_);
    const size_t ib = r2c_ ? rb : 2 * rb, ob = r2c_ ? 2 * rb : rb;
    // the real side as pairs: unit stride, even distance, pair-aligned base
    const void *rp = r2c_ ? in : out;
    const int64_t rs = r2c_ ? is_ : os_, rd = r2c_ ? id_ : od_;
    const int pair = (rs == 1 && (batch_ == 1 || rd % 2 == 0) &&
                      ((uintptr_t)rp % (2 * rb)) == 0) ? 1 : 0;
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk) {
      const int64_t cb = std::min(chunk, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * ib;
      a.out = work;
      a.tw = ecol_.tw.p;
      a.desc = ecol_.d();
      a.aux0 = wm_.p;
      a.aux1 = wn_.p;
      a.count = cb * m2_;
      a.istride = is_;
      a.idist = id_;
      a.n = m1_;
      a.n1 = m1_;
      a.n2 = m2_;
      a.bigdist = M;
      a.shift = shift_m_;
      a.shift2 = shift_n_;
      a.flags = pair | extra_;
      hipError_t err = col_.e->launch(a, (a.count + col_.e->tpb - 1) / col_.e->tpb,
                                      col_.e->lds_bytes(m1_), s);
      if (err != hipSuccess) return err;
      KArgs b{};
      b.in = work;
      b.out = static_cast<char *>(out) + (size_t)(b0 * od_) * ob;
      b.tw = erow_.tw.p;
      b.desc = erow_.d();
      b.aux1 = wn_.p;
      b.ostride = os_;
      b.odist = od_;
      b.n = m2_;
      b.n1 = m1_;
      b.n2 = m2_;
      b.bigdist = M;
      b.shift2 = shift_n_;
      b.flags = pair | extra_;
      int64_t grid;
      const int tpb = row_.e->tpb;
      if (r2c_) {
        b.count = cb;
        grid = cb * ((m1_ / 2 + tpb / 2 - 1) / (tpb / 2));
      } else {
        b.count = cb * m1_;
        grid = (b.count + tpb - 1) / tpb;
      }
      err = row_.e->launch(b, grid, row_.e->lds_bytes(m2_), s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (r2c_ ? "r2c" : "c2r") << (extra_ & 2 ? "_adjoint" : "") << "_fourstep n="
      << 2 * m1_ * m2_ << " half=" << m1_ << "x" << m2_ << " chunk=" << chunk_
      << " batch=" << batch_ << " is=" << is_ << " id=" << id_ << " os=" << os_
      << " od=" << od_ << " " << chain_str(col_) << " " << chain_str(row_) << ")";
    return o.str();
  }

 private:
  bool r2c_;
  Choice col_, row_;
  int64_t m1_, m2_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  int extra_;
  EngineData ecol_, erow_;
  DevMem wm_, wn_;
  int shift_m_ = 0, shift_n_ = 0;
};

// The four-step split M = m1 * m2 (m1 even, kernels for both sides) with
// compile-time chains preferred, then the most balanced; m1 == 0 if none.
static void real_fourstep_split(bool r2c, int64_t M, int prec, int64_t *m1_out, Choice *cc,
                                Choice *rc) {
  *m1_out = 0;
  const int kc = r2c ? KK_COL_R2C : KK_COL_C2R, kr = r2c ? KK_ROW_R2C : KK_ROW_C2R;
  for (int64_t m1 = 2; m1 < M && m1 * m1 <= M * 4; m1 += 2) {
    if (M % m1) continue;
    const int64_t m2 = M / m1;
    Choice c = choose(kc, m1, prec, false);
    if (!c.e) continue;
    Choice r = choose(kr, m2, prec, false);
    if (!r.e) continue;
    const double bal = std::fabs(std::log((double)m1 / (double)m2));
    const bool stat = c.e->n && r.e->n;
    const bool best_stat = *m1_out && cc->e->n && rc->e->n;
    const double best_bal =
        *m1_out ? std::fabs(std::log((double)*m1_out / (double)(M / *m1_out))) : 1e30;
    if (!*m1_out || (stat && !best_stat) || (stat == best_stat && bal < best_bal)) {
      *m1_out = m1;
      *cc = c;
      *rc = r;
    }
  }
}

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
  if (!c.e && p.n % 2 == 0 && p.n >= 8) {
    // packed half-length transform on the four-step passes
    int64_t m1 = 0;
    Choice cc, rc;
    real_fourstep_split(r2c, p.n / 2, p.prec, &m1, &cc, &rc);
    if (m1) {
      const int64_t M = p.n / 2;
      auto node = std::make_unique<RealFourStepNode>(
          r2c, cc, rc, m1, M / m1, p.prec, p.howmany, p.istride, id, p.ostride, od,
          fourstep_chunk(M, p.prec, p.howmany), p.adjoint ? 2 : 0);
      hipError_t err = node->init();
      if (err != hipSuccess) {
        *why = std::string("device upload failed: ") + hipGetErrorString(err);
        return nullptr;
      }
      return node;
    }
  }
  if (!c.e) c = choose(r2c ? KK_R2C_F : KK_C2R_F, p.n, p.prec, false);
  if (!c.e) {
    const int64_t per = p.n * 2 * prec_compute_bytes(p.prec);
    const int64_t chunk =
        std::max<int64_t>(1, std::min<int64_t>(p.howmany, (int64_t(64) << 20) / per));
    auto node = std::make_unique<RealViaComplexNode>(r2c, p.n, p.prec, p.howmany, p.istride, id,
                                                     p.ostride, od, chunk);
    node->flags_ = p.adjoint ? 2 : 0;
    if (node->init(why) != hipSuccess) return nullptr;
    return node;
  }
  auto node = std::make_unique<RealNode>(r2c, half, c, p.n, p.prec, p.howmany, p.istride, id,
                                         p.ostride, od, p.adjoint ? 2 : 0);
  hipError_t err = node->init();
  if (err != hipSuccess) {
    *why = std::string("device upload failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  return node;
}

// ---------------------------------------------------------------------------
// DCT-II / DST-II / DCT-III / DST-III (r2r kinds 1, 5, 2, 6)
// ---------------------------------------------------------------------------
class DctNode : public Node {
 public:
  DctNode(bool two, bool dst, bool half, Choice c, int64_t n, int prec, int64_t batch,
          int64_t is, int64_t id, int64_t os, int64_t od, int flags)
      : two_(two), dst_(dst), half_(half), c_(std::move(c)), n_(n), prec_(prec), batch_(batch),
        is_(is), id_(id), os_(os), od_(od), flags_(flags) {}
  int64_t len() const { return half_ ? n_ / 2 : n_; }
  hipError_t init() {
    hipError_t err = eng_.init(c_, len());
    if (err != hipSuccess) return err;
    if (half_) err = upload_table(aux0_, root_table(n_, n_ / 2 + 1), prec_);
    if (err != hipSuccess) return err;
    // w_k = exp(-i pi k / 2n) = W_{4n}^k
    return upload_table(aux1_, root_table(4 * n_, half_ ? n_ / 2 + 1 : n_), prec_);
  }
  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    if (batch_ <= 0) return hipSuccess;
    KArgs a{};
    a.in = in;
    a.out = out;
    a.tw = eng_.tw.p;
    a.desc = eng_.d();
    a.aux0 = aux0_.p;
    a.aux1 = aux1_.p;
    a.count = batch_;
    a.istride = is_;
    a.idist = id_;
    a.ostride = os_;
    a.odist = od_;
    a.n = len();
    a.flags = flags_;
    const int64_t grid = (batch_ + c_.e->tpb - 1) / c_.e->tpb;
    return c_.e->launch(a, grid, c_.e->lds_bytes(len()), s);
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (dst_ ? "dst" : "dct") << (two_ ? "2" : "3") << (flags_ & 2 ? "_adjoint" : "")
      << " n=" << n_ << " " << chain_str(c_) << " batch=" << batch_ << " is=" << is_
      << " id=" << id_ << " os=" << os_ << " od=" << od_ << ")";
    return o.str();
  }

 private:
  bool two_, dst_, half_;
  Choice c_;
  int64_t n_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_;
  int flags_;
  EngineData eng_;
  DevMem aux0_, aux1_;
};

// DCT/DST II/III through a C2C plan (any length): pre-processing kernel,
// forward C2C in the compute precision, post-processing kernel.
class DctViaComplexNode : public Node {
 public:
  DctViaComplexNode(int flags, int64_t n, int prec, int64_t batch, int64_t is, int64_t id,
                    int64_t os, int64_t od, int64_t chunk)
      : flags_(flags), n_(n), prec_(prec), batch_(batch), is_(is), id_(id), os_(os), od_(od),
        chunk_(chunk) {}
  hipError_t init(std::string *why) {
    Problem q;
    q.kind = K_C2C;
    q.dir = -1;
    q.prec = prec_compute_bytes(prec_) == 8 ? P_F64 : P_F32;
    q.n = n_;
    q.howmany = chunk_;
    q.istride = q.ostride = 1;
    q.idist = q.odist = n_;
    q.may_alias = true;
    inner_ = build_c2c_node(q, why);
    if (!inner_) return hipErrorInvalidValue;
    return upload_table(w_, root_table(4 * n_, n_), prec_);
  }
  size_t buf_bytes() const { return (size_t)chunk_ * (size_t)n_ * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return buf_bytes() + inner_->workspace_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    char *buf = static_cast<char *>(work);
    void *iws = buf + buf_bytes();
    const size_t rb = (size_t)prec_real_bytes(prec_);
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk_) {
      const int64_t cb = std::min(chunk_, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * rb;
      a.out = buf;
      a.aux1 = w_.p;
      a.count = cb;
      a.istride = is_;
      a.idist = id_;
      a.n = n_;
      a.flags = flags_;
      hipError_t err = launch_real_stage(4, prec_, a, cb * n_, s);
      if (err == hipSuccess) err = inner_->enqueue(buf, buf, iws, s);
      if (err != hipSuccess) return err;
      a.in = buf;
      a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * rb;
      a.ostride = os_;
      a.odist = od_;
      err = launch_real_stage(5, prec_, a, cb * n_, s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (flags_ & 4 ? "dst" : "dct") << (flags_ & 8 ? "3" : "2")
      << (flags_ & 2 ? "_adjoint" : "") << "_via_c2c n=" << n_ << " chunk=" << chunk_
      << " batch=" << batch_ << " " << inner_->describe() << ")";
    return o.str();
  }

 private:
  int flags_;
  int64_t n_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  std::unique_ptr<Node> inner_;
  DevMem w_;
};

// DCT-I, DCT-IV, DST-I, DST-IV, DHT, R2HC and HC2R through one complex DFT
// of length L per member (pre and post kernels k_r2r_pre / k_r2r_post,
// launch_real_stage 6 / 7), in chunks of members.
class R2RViaComplexNode : public Node {
 public:
  R2RViaComplexNode(int kind, bool adjoint, int64_t n, int prec, int64_t batch, int64_t is,
                    int64_t id, int64_t os, int64_t od)
      : kind_(kind), adj_(adjoint), n_(n), prec_(prec), batch_(batch), is_(is), id_(id),
        os_(os), od_(od) {
    L_ = kind == 0 ? 2 * (n - 1) : kind == 4 ? 2 * (n + 1) : (kind == 3 || kind == 7) ? 2 * n : n;
    const int64_t per = L_ * 2 * prec_compute_bytes(prec);
    chunk_ = std::max<int64_t>(1, std::min<int64_t>(batch, (int64_t(64) << 20) / per));
  }
  hipError_t init(std::string *why) {
    Problem q;
    q.kind = K_C2C;
    q.dir = -1;
    q.prec = prec_compute_bytes(prec_) == 8 ? P_F64 : P_F32;
    q.n = L_;
    q.howmany = chunk_;
    q.istride = q.ostride = 1;
    q.idist = q.odist = L_;
    q.may_alias = true;
    inner_ = build_c2c_node(q, why);
    if (!inner_) return hipErrorInvalidValue;
    if (kind_ != 3 && kind_ != 7) return hipSuccess;
    // p_j = e^{-i pi j / 2n} = W_4n^j, q_k = e^{-i pi (2k+1) / 4n} = W_8n^{2k+1}
    std::vector<double> qk(2 * (size_t)n_);
    for (int64_t k = 0; k < n_; ++k) {
      double re, im;
      unit_root(2 * k + 1, 8 * n_, &re, &im);
      qk[2 * k] = re;
      qk[2 * k + 1] = -im;
    }
    hipError_t err = upload_table(pre_, root_table(4 * n_, n_), prec_);
    if (err == hipSuccess) err = upload_table(post_, qk, prec_);
    return err;
  }
  size_t buf_bytes() const { return (size_t)chunk_ * (size_t)L_ * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return buf_bytes() + inner_->workspace_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    char *buf = static_cast<char *>(work);
    void *iws = buf + buf_bytes();
    const size_t rb = (size_t)prec_real_bytes(prec_);
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk_) {
      const int64_t cb = std::min(chunk_, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * rb;
      a.out = buf;
      a.aux0 = pre_.p;
      a.aux1 = post_.p;
      a.count = cb;
      a.istride = is_;
      a.idist = id_;
      a.ostride = os_;
      a.odist = od_;
      a.n = n_;
      a.n1 = L_;
      a.flags = kind_ | (adj_ ? 256 : 0);
      hipError_t err = launch_real_stage(6, prec_, a, cb * L_, s);
      if (err == hipSuccess) err = inner_->enqueue(buf, buf, iws, s);
      if (err != hipSuccess) return err;
      a.in = buf;
      a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * rb;
      err = launch_real_stage(7, prec_, a, cb * n_, s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    static const char *names[] = {"dct1", "dct2", "dct3", "dct4", "dst1", "dst2",
                                  "dst3", "dst4", "dht",  "r2hc", "hc2r"};
    std::ostringstream o;
    o << "(" << names[kind_] << (adj_ ? "_adjoint" : "") << "_via_c2c n=" << n_ << " L=" << L_
      << " chunk=" << chunk_ << " batch=" << batch_ << " " << inner_->describe() << ")";
    return o.str();
  }

 private:
  int kind_;
  bool adj_;
  int64_t n_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_;
  int64_t L_ = 0, chunk_ = 1;
  std::unique_ptr<Node> inner_;
  DevMem pre_, post_;
};

std::unique_ptr<Node> build_r2r_node(const Problem &p, std::string *why) {
  // plfft_r2r_kind_t values: DCT_1 = 0, DCT_2 = 1, DCT_3 = 2, DCT_4 = 3,
  // DST_1 = 4, DST_2 = 5, DST_3 = 6, DST_4 = 7, DHT = 8, R2HC = 9, HC2R = 10
  const int k = p.r2r_kind;
  if (k < 0 || k > 10) {
    *why = "unknown r2r kind " + std::to_string(k);
    return nullptr;
  }
  if (k != 1 && k != 2 && k != 5 && k != 6) {
    const int64_t id = p.howmany > 1 ? p.idist : 0;
    const int64_t od = p.howmany > 1 ? p.odist : 0;
    auto node = std::make_unique<R2RViaComplexNode>(k, p.adjoint != 0, p.n, p.prec, p.howmany,
                                                    p.istride, id, p.ostride, od);
    if (node->init(why) != hipSuccess) {
      if (why->empty()) *why = "r2r setup failed";
      return nullptr;
    }
    return node;
  }
  const bool two = (k == 1 || k == 5), dst = (k == 5 || k == 6);
  const int64_t id = p.howmany > 1 ? p.idist : 0;
  const int64_t od = p.howmany > 1 ? p.odist : 0;
  Choice c;
  bool half = false;
  if (p.n % 2 == 0) {
    c = choose(two ? KK_DCT2_E : KK_DCT3_E, p.n / 2, p.prec, dst);
    half = c.e != nullptr;
  }
  if (!c.e) c = choose(two ? KK_DCT2_F : KK_DCT3_F, p.n, p.prec, dst);
  if (!c.e) {
    const int64_t per = p.n * 2 * prec_compute_bytes(p.prec);
    const int64_t chunk =
        std::max<int64_t>(1, std::min<int64_t>(p.howmany, (int64_t(64) << 20) / per));
    const int flags = (p.adjoint ? 2 : 0) | (dst ? 4 : 0) | (two ? 0 : 8);
    auto node = std::make_unique<DctViaComplexNode>(flags, p.n, p.prec, p.howmany, p.istride,
                                                    id, p.ostride, od, chunk);
    if (node->init(why) != hipSuccess) {
      if (why->empty()) *why = "DCT fallback setup failed";
      return nullptr;
    }
    return node;
  }
  auto node = std::make_unique<DctNode>(two, dst, half, c, p.n, p.prec, p.howmany, p.istride, id,
                                        p.ostride, od, p.adjoint ? 2 : 0);
  hipError_t err = node->init();
  if (err != hipSuccess) {
    *why = std::string("device upload failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  return node;
}

}  // namespace plf
