// conv_plan.cpp -- plan nodes for lengths with large prime factors.
//
// The reference tries Rader before Bluestein (src/planner/planner.cpp:178-179);
// here a prime n whose n-1 has an LDS kernel runs fused Rader, other lengths
// run fused Bluestein when the padded length M >= 2n-2 fits one LDS tile, a
// prime whose n-1 splits into compiled four-step sides runs Rader in three
// passes (65537: N = 2^16), and the rest run chirp-z in three passes or
// around the four-step transform.
// The convolution kernel B = DFT(b)/M is computed once at plan time in FP64
// on the GPU with this library's own C2C plan, then rounded to the compute
// precision (the reference likewise transforms b with its own plan,
// bluestein.hpp:100-110, rader.hpp:79-90).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <sstream>

#include "plan.hpp"

namespace plf {

static int64_t spf(int64_t n) {
  if (n % 2 == 0) return 2;
  for (int64_t d = 3; d * d <= n; d += 2)
    if (n % d == 0) return d;
  return n;
}

static int64_t group_generator(int64_t n) {
  // smallest g whose powers visit every nonzero residue (rader_generator.cpp:13-33)
  std::vector<char> seen((size_t)n);
  for (int64_t g = 2; g < n; ++g) {
    std::fill(seen.begin(), seen.end(), 0);
    int64_t x = 1;
    bool ok = true;
    for (int64_t i = 1; i < n; ++i) {
      if (seen[(size_t)x]) {
        ok = false;
        break;
      }
      seen[(size_t)x] = 1;
      x = x * g % n;
    }
    if (ok) return g;
  }
  return 0;
}

// In-place forward DFTs of `count` contiguous length-M rows of v
// (interleaved FP64) on the GPU.
hipError_t dft_f64_gpu(std::vector<double> &v, int64_t M, int64_t count) {
  Problem q;
  q.kind = K_C2C;
  q.dir = -1;
  q.prec = P_F64;
  q.n = M;
  q.howmany = count;
  q.idist = q.odist = M;
  q.may_alias = true;
  std::string why;
  std::unique_ptr<Node> node = build_c2c_node(q, &why);
  if (!node) return hipErrorInvalidValue;
  DevMem buf, ws;
  hipError_t err = buf.upload(v.data(), v.size() * sizeof(double));
  if (err == hipSuccess) err = ws.alloc(node->workspace_bytes());
  if (err == hipSuccess) err = node->enqueue(buf.p, buf.p, ws.p, nullptr);
  if (err == hipSuccess) err = hipStreamSynchronize(nullptr);
  if (err == hipSuccess)
    err = hipMemcpy(v.data(), buf.p, v.size() * sizeof(double), hipMemcpyDeviceToHost);
  return err;
}

// Chirp c_k = exp(-i pi (k^2 mod 2n)/n), phase in exact integers
// (bluestein.hpp:66-79).
std::vector<double> chirp_table(int64_t n) {
  std::vector<double> c(2 * (size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t ph = (int64_t)(((unsigned __int128)k * (unsigned __int128)k) %
                                 (unsigned __int128)(2 * n));
    double re, im;
    unit_root(ph, 2 * n, &re, &im);
    c[2 * k] = re;
    c[2 * k + 1] = -im;
  }
  return c;
}

// B = DFT_M(b)/M with b_k = b_{M-k} = conj(c_k), 0 <= k < n (bluestein.hpp:81-98).
static hipError_t bluestein_kernel(int64_t n, int64_t M, const std::vector<double> &c,
                                   std::vector<double> *B) {
  B->assign(2 * (size_t)M, 0.0);
  const double s = 1.0 / (double)M;
  for (int64_t k = 0; k < n; ++k) {
    const double re = c[2 * k] * s, im = -c[2 * k + 1] * s;
    (*B)[2 * k] = re;
    (*B)[2 * k + 1] = im;
    if (k) {
      (*B)[2 * (M - k)] = re;
      (*B)[2 * (M - k) + 1] = im;
    }
  }
  return dft_f64_gpu(*B, M, 1);
}

// ---------------------------------------------------------------------------
class FusedConvNode : public Node {
 public:
  FusedConvNode(bool rader, Choice c, int64_t n, int64_t len, int prec, int64_t batch,
                int64_t is, int64_t id, int64_t os, int64_t od)
      : rader_(rader), c_(std::move(c)), n_(n), len_(len), prec_(prec), batch_(batch), is_(is),
        id_(id), os_(os), od_(od) {}
  hipError_t init() {
    hipError_t err = eng_.init(c_, len_);
    if (err != hipSuccess) return err;
    std::vector<double> B;
    if (rader_) {
      const int64_t m = len_;
      const int64_t g = group_generator(n_);
      std::vector<int> perm((size_t)m);
      std::vector<double> b(2 * (size_t)m);
      int64_t x = 1;
      for (int64_t k = 0; k < m; ++k) {
        perm[(size_t)k] = (int)x;
        double re, im;
        unit_root(x, n_, &re, &im);  // W_n^{perm[k]}, forward sign
        b[2 * ((m - k) % m)] = re / (double)m;
        b[2 * ((m - k) % m) + 1] = -im / (double)m;
        x = x * g % n_;
      }
      err = dft_f64_gpu(b, m, 1);
      if (err == hipSuccess) err = upload_table(aux0_, b, prec_);
      if (err == hipSuccess) err = aux1_.upload(perm.data(), perm.size() * sizeof(int));
      gen_ = g;
    } else {
      std::vector<double> c = chirp_table(n_);
      err = bluestein_kernel(n_, len_, c, &B);
      if (err == hipSuccess) err = upload_table(aux0_, B, prec_);
      if (err == hipSuccess) err = upload_table(aux1_, c, prec_);
    }
    return err;
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
    a.n = len_;
    a.n1 = n_;
    const int64_t grid = (batch_ + c_.e->tpb - 1) / c_.e->tpb;
    return c_.e->launch(a, grid, c_.e->lds_bytes(len_), s);
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(" << (rader_ ? "rader" : "bluestein") << " n=" << n_ << (rader_ ? " m=" : " M=")
      << len_;
    if (rader_) o << " g=" << gen_;
    o << " " << chain_str(c_) << " batch=" << batch_ << ")";
    return o.str();
  }

 private:
  bool rader_;
  Choice c_;
  int64_t n_, len_;
  int prec_;
  int64_t batch_, is_, id_, os_, od_;
  EngineData eng_;
  DevMem aux0_, aux1_;
  int64_t gen_ = 0;
};

// ---------------------------------------------------------------------------
// Chirp-z around a length-M C2C node (four-step) for n beyond one LDS tile.
// ---------------------------------------------------------------------------
class BigBluesteinNode : public Node {
 public:
  BigBluesteinNode(int64_t n, int64_t M, int prec, bool bwd, int64_t batch, int64_t is,
                   int64_t id, int64_t os, int64_t od, int64_t chunk)
      : n_(n), M_(M), prec_(prec), bwd_(bwd), batch_(batch), is_(is), id_(id), os_(os), od_(od),
        chunk_(chunk) {}
  hipError_t init(std::string *why) {
    Problem q;
    q.kind = K_C2C;
    q.dir = -1;
    q.prec = prec_compute_bytes(prec_) == 8 ? P_F64 : P_F32;
    q.n = M_;
    q.howmany = chunk_;
    q.istride = q.ostride = 1;
    q.idist = q.odist = M_;
    q.may_alias = true;
    q.no_conv = 1;
    inner_ = build_c2c_node(q, why);
    if (!inner_) return hipErrorInvalidValue;
    std::vector<double> c = chirp_table(n_), B;
    hipError_t err = bluestein_kernel(n_, M_, c, &B);
    if (err == hipSuccess) err = upload_table(B_, B, prec_);
    if (err == hipSuccess) err = upload_table(c_, c, prec_);
    return err;
  }
  size_t buf_bytes() const { return (size_t)chunk_ * (size_t)M_ * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return buf_bytes() + inner_->workspace_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    char *buf = static_cast<char *>(work);
    void *iws = buf + buf_bytes();
    const size_t eb = 2 * (size_t)prec_real_bytes(prec_);
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk_) {
      const int64_t cb = std::min(chunk_, batch_ - b0);
      KArgs a{};
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * eb;
      a.out = buf;
      a.aux0 = B_.p;
      a.aux1 = c_.p;
      a.count = cb;
      a.istride = is_;
      a.idist = id_;
      a.ostride = os_;
      a.odist = od_;
      a.n = M_;
      a.n1 = n_;
      hipError_t err = launch_chirp_in(prec_, bwd_, a, s);
      if (err == hipSuccess) err = inner_->enqueue(buf, buf, iws, s);
      if (err == hipSuccess) err = launch_pointwise_swap(prec_, a, s);
      if (err == hipSuccess) err = inner_->enqueue(buf, buf, iws, s);
      if (err != hipSuccess) return err;
      a.in = buf;
      a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * eb;
      err = launch_chirp_out(prec_, bwd_, a, s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(chirpz n=" << n_ << " M=" << M_ << " chunk=" << chunk_ << " batch=" << batch_ << " "
      << inner_->describe() << ")";
    return o.str();
  }

 private:
  int64_t n_, M_;
  int prec_;
  bool bwd_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  std::unique_ptr<Node> inner_;
  DevMem B_, c_;
};

// Chirp-z in three passes (kernels_cz.hpp) over a power-of-two M = n1 * n2
// whose sides have compiled column and row engines: chirp-in columns, rows
// that finish DFT_M, multiply by B and start the inverse, chirp-out columns.
// Each member's workspace is M complex values of the compute precision.
class ChirpZ3Node : public Node {
 public:
  ChirpZ3Node(int64_t n, int64_t n1, int64_t n2, int prec, bool bwd, int64_t batch, int64_t is,
              int64_t id, int64_t os, int64_t od, int64_t chunk, Choice cin, Choice cmid,
              Choice cout)
      : n_(n), n1_(n1), n2_(n2), prec_(prec), bwd_(bwd), batch_(batch), is_(is), id_(id),
        os_(os), od_(od), chunk_(chunk), cin_(std::move(cin)), cmid_(std::move(cmid)),
        cout_(std::move(cout)) {}
  hipError_t init() {
    hipError_t err = ein_.init(cin_, n1_);
    if (err == hipSuccess) err = emid_.init(cmid_, n2_);
    if (err == hipSuccess) err = eout_.init(cout_, n1_);
    if (err != hipSuccess) return err;
    const int64_t M = n1_ * n2_;
    int lg = 0;
    while ((int64_t(1) << lg) < M) ++lg;
    shift_ = (lg + 1) / 2;
    const int64_t nlo = int64_t(1) << shift_, nhi = ((M - 1) >> shift_) + 1;
    // W_M split table: lo[r] = W_M^r (r < 2^shift), then hi[q] = W_M^{q 2^shift}
    std::vector<double> sp;
    for (int64_t r = 0; r < nlo; ++r) {
      double re, im;
      unit_root(r, M, &re, &im);
      sp.push_back(re);
      sp.push_back(-im);
    }
    for (int64_t q = 0; q < nhi; ++q) {
      double re, im;
      unit_root(q * nlo, M, &re, &im);
      sp.push_back(re);
      sp.push_back(-im);
    }
    err = upload_table(split_, sp, prec_);
    if (err != hipSuccess) return err;
    std::vector<double> c = chirp_table(n_), B;
    err = bluestein_kernel(n_, M, c, &B);
    if (err != hipSuccess) return err;
    // the middle pass reads row k1 of B[k1 + n1 k2]: store it row-major
    std::vector<double> Bt(B.size());
    for (int64_t k1 = 0; k1 < n1_; ++k1)
      for (int64_t k2 = 0; k2 < n2_; ++k2) {
        Bt[2 * (k1 * n2_ + k2)] = B[2 * (k1 + n1_ * k2)];
        Bt[2 * (k1 * n2_ + k2) + 1] = B[2 * (k1 + n1_ * k2) + 1];
      }
    err = upload_table(B_, Bt, prec_);
    if (err == hipSuccess) err = upload_table(c_, c, prec_);
    return err;
  }
  size_t per_bytes() const { return (size_t)(n1_ * n2_) * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return (size_t)chunk_ * per_bytes(); }
  size_t min_workspace_bytes() const override { return per_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    return run(in, out, work, chunk_, s);
  }
  // a workspace of `cap` bytes: members in chunks of cap / per_bytes()
  hipError_t enqueue_capped(const void *in, void *out, void *work, size_t cap,
                            hipStream_t s) const override {
    const int64_t c = std::min<int64_t>(chunk_, (int64_t)(cap / per_bytes()));
    if (c < 1) return hipErrorInvalidValue;
    return run(in, out, work, c, s);
  }
  hipError_t run(const void *in, void *out, void *work, int64_t chunk, hipStream_t s) const {
    const size_t eb = 2 * (size_t)prec_real_bytes(prec_);
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk) {
      const int64_t cb = std::min(chunk, batch_ - b0);
      KArgs a{};
      a.aux0 = split_.p;
      a.n1 = n1_;
      a.n2 = n2_;
      a.bigdist = n_;
      a.shift = shift_;
      // 1. chirp-in columns
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * eb;
      a.out = work;
      a.istride = is_;
      a.idist = id_;
      a.tw = ein_.tw.p;
      a.desc = ein_.d();
      a.aux1 = c_.p;
      a.n = n1_;
      a.count = cb * n2_;
      hipError_t err = cin_.e->launch(a, (a.count + cin_.e->tpb - 1) / cin_.e->tpb,
                                      cin_.e->lds_bytes(n1_), s);
      // 2. rows: DFT_n2, x B, the inverse's first DFT_n2
      if (err == hipSuccess) {
        a.tw = emid_.tw.p;
        a.desc = emid_.d();
        a.aux1 = B_.p;
        a.n = n2_;
        a.count = cb * n1_;
        err = cmid_.e->launch(a, (a.count + cmid_.e->tpb - 1) / cmid_.e->tpb,
                              cmid_.e->lds_bytes(n2_), s);
      }
      // 3. chirp-out columns
      if (err == hipSuccess) {
        a.in = work;
        a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * eb;
        a.ostride = os_;
        a.odist = od_;
        a.tw = eout_.tw.p;
        a.desc = eout_.d();
        a.aux1 = c_.p;
        a.n = n1_;
        a.count = cb * n2_;
        err = cout_.e->launch(a, (a.count + cout_.e->tpb - 1) / cout_.e->tpb,
                              cout_.e->lds_bytes(n1_), s);
      }
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(chirpz3 n=" << n_ << " M=" << n1_ * n2_ << " = " << n1_ << "x" << n2_
      << " chunk=" << chunk_ << " batch=" << batch_ << " " << chain_str(cin_) << " "
      << chain_str(cmid_) << " " << chain_str(cout_) << ")";
    return o.str();
  }

 private:
  int64_t n_, n1_, n2_;
  int prec_;
  bool bwd_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  Choice cin_, cmid_, cout_;
  EngineData ein_, emid_, eout_;
  int shift_ = 0;
  DevMem split_, B_, c_;
};

// Rader for a prime n in three passes over N = n - 1 = n1 * n2
// (rader.hpp:51-90 restated on the four-step passes): permuted-input
// columns, the chirp-z row pass (DFT_n2, x B, the inverse's first DFT_n2,
// plus A_0 kept), permuted-output columns adding x_0. Each member's
// workspace is N complex values, then two slots per member of the chunk.
class Rader3Node : public Node {
 public:
  Rader3Node(int64_t n, int64_t n1, int64_t n2, int prec, bool bwd, int64_t batch, int64_t is,
             int64_t id, int64_t os, int64_t od, int64_t chunk, Choice cin, Choice cmid,
             Choice cout)
      : n_(n), n1_(n1), n2_(n2), prec_(prec), bwd_(bwd), batch_(batch), is_(is), id_(id),
        os_(os), od_(od), chunk_(chunk), cin_(std::move(cin)), cmid_(std::move(cmid)),
        cout_(std::move(cout)) {}
  hipError_t init() {
    hipError_t err = ein_.init(cin_, n1_);
    if (err == hipSuccess) err = emid_.init(cmid_, n2_);
    if (err == hipSuccess) err = eout_.init(cout_, n1_);
    if (err != hipSuccess) return err;
    const int64_t N = n1_ * n2_;
    int lg = 0;
    while ((int64_t(1) << lg) < N) ++lg;
    shift_ = (lg + 1) / 2;
    const int64_t nlo = int64_t(1) << shift_, nhi = ((N - 1) >> shift_) + 1;
    // W_N split table: lo[r] = W_N^r (r < 2^shift), then hi[q] = W_N^{q 2^shift}
    std::vector<double> sp;
    for (int64_t r = 0; r < nlo; ++r) {
      double re, im;
      unit_root(r, N, &re, &im);
      sp.push_back(re);
      sp.push_back(-im);
    }
    for (int64_t q = 0; q < nhi; ++q) {
      double re, im;
      unit_root(q * nlo, N, &re, &im);
      sp.push_back(re);
      sp.push_back(-im);
    }
    err = upload_table(split_, sp, prec_);
    if (err != hipSuccess) return err;
    gen_ = group_generator(n_);
    if (!gen_) return hipErrorInvalidValue;
    std::vector<int> perm((size_t)N);
    std::vector<double> b(2 * (size_t)N);
    int64_t x = 1;
    for (int64_t k = 0; k < N; ++k) {
      perm[(size_t)k] = (int)x;
      double re, im;
      unit_root(x, n_, &re, &im);  // W_n^{perm[k]}, forward sign
      b[2 * ((N - k) % N)] = re / (double)N;
      b[2 * ((N - k) % N) + 1] = -im / (double)N;
      x = x * gen_ % n_;
    }
    err = dft_f64_gpu(b, N, 1);
    if (err != hipSuccess) return err;
    // the row pass reads row k1 of B[k1 + n1 k2]: store it row-major
    std::vector<double> Bt(b.size());
    for (int64_t k1 = 0; k1 < n1_; ++k1)
      for (int64_t k2 = 0; k2 < n2_; ++k2) {
        Bt[2 * (k1 * n2_ + k2)] = b[2 * (k1 + n1_ * k2)];
        Bt[2 * (k1 * n2_ + k2) + 1] = b[2 * (k1 + n1_ * k2) + 1];
      }
    err = upload_table(B_, Bt, prec_);
    if (err == hipSuccess) err = perm_.upload(perm.data(), perm.size() * sizeof(int));
    return err;
  }
  size_t per_bytes() const { return (size_t)(n1_ * n2_ + 2) * 2 * prec_compute_bytes(prec_); }
  size_t workspace_bytes() const override { return (size_t)chunk_ * per_bytes(); }
  size_t min_workspace_bytes() const override { return per_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    return run(in, out, work, chunk_, s);
  }
  hipError_t enqueue_capped(const void *in, void *out, void *work, size_t cap,
                            hipStream_t s) const override {
    const int64_t c = std::min<int64_t>(chunk_, (int64_t)(cap / per_bytes()));
    if (c < 1) return hipErrorInvalidValue;
    return run(in, out, work, c, s);
  }
  hipError_t run(const void *in, void *out, void *work, int64_t chunk, hipStream_t s) const {
    const size_t eb = 2 * (size_t)prec_real_bytes(prec_);
    const size_t cb = 2 * (size_t)prec_compute_bytes(prec_);
    const int64_t N = n1_ * n2_;
    // the chunk's members' slots follow their workspaces
    void *slots = static_cast<char *>(work) + (size_t)chunk * (size_t)N * cb;
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk) {
      const int64_t nb = std::min(chunk, batch_ - b0);
      KArgs a{};
      a.aux0 = split_.p;
      a.aux1 = perm_.p;
      a.aux3 = slots;
      a.n1 = n1_;
      a.n2 = n2_;
      a.shift = shift_;
      // 1. permuted-input columns
      a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * eb;
      a.out = work;
      a.istride = is_;
      a.idist = id_;
      a.tw = ein_.tw.p;
      a.desc = ein_.d();
      a.n = n1_;
      a.count = nb * n2_;
      hipError_t err = cin_.e->launch(a, (a.count + cin_.e->tpb - 1) / cin_.e->tpb,
                                      cin_.e->lds_bytes(n1_), s);
      // 2. rows: DFT_n2, x B, the inverse's first DFT_n2 (A_0 kept)
      if (err == hipSuccess) {
        a.in = nullptr;
        a.tw = emid_.tw.p;
        a.desc = emid_.d();
        a.aux1 = B_.p;
        a.n = n2_;
        a.count = nb * n1_;
        err = cmid_.e->launch(a, (a.count + cmid_.e->tpb - 1) / cmid_.e->tpb,
                              cmid_.e->lds_bytes(n2_), s);
      }
      // 3. permuted-output columns, one block per CU: the scattered stores
      // of fewer members in flight stay in L2 until their lines fill
      // (65537, batch 2047: 1.44 -> 1.03 ms in f32; 2.28 -> 2.08 ms for the
      // f64 plan). The launch asks for more than half of the CU's 160 KB LDS.
      if (err == hipSuccess) {
        a.in = work;
        a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * eb;
        a.ostride = os_;
        a.odist = od_;
        a.tw = eout_.tw.p;
        a.desc = eout_.d();
        a.aux1 = perm_.p;
        a.n = n1_;
        a.count = nb * n2_;
        err = cout_.e->launch(a, (a.count + cout_.e->tpb - 1) / cout_.e->tpb,
                              std::max<size_t>(81 * 1024, cout_.e->lds_bytes(n1_)), s);
      }
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(rader3 n=" << n_ << " N=" << n1_ * n2_ << " = " << n1_ << "x" << n2_ << " g=" << gen_
      << " chunk=" << chunk_ << " batch=" << batch_ << " " << chain_str(cin_) << " "
      << chain_str(cmid_) << " " << chain_str(cout_) << ")";
    return o.str();
  }

 private:
  int64_t n_, n1_, n2_;
  int prec_;
  bool bwd_;
  int64_t batch_, is_, id_, os_, od_, chunk_;
  Choice cin_, cmid_, cout_;
  EngineData ein_, emid_, eout_;
  int shift_ = 0;
  int64_t gen_ = 0;
  DevMem split_, B_, perm_;
};

// The most balanced split N = n1 * n2 (n1 a power of two) with compiled
// Rader column passes on n1 and a row pass on n2.
static bool rd3_choice(int64_t N, int prec, bool bwd, int64_t *n1_out, Choice *cin, Choice *cmid,
                       Choice *cout) {
  double best = 1e30;
  for (int64_t n1 = 2; n1 < N; n1 *= 2) {
    if (N % n1) break;
    const int64_t n2 = N / n1;
    Choice ci = choose(KK_RD_COL_IN, n1, prec, bwd);
    Choice cm = choose(KK_CZ_ROW_MID, n2, prec, bwd);
    Choice co = choose(KK_RD_COL_OUT, n1, prec, bwd);
    if (!ci.e || !cm.e || !co.e) continue;
    const double bal = std::fabs(std::log((double)n1 / (double)n2));
    if (bal < best) {
      best = bal;
      *n1_out = n1;
      *cin = ci;
      *cmid = cm;
      *cout = co;
    }
  }
  return best < 1e30;
}

bool cz3_choice(int64_t n, int prec, bool bwd, Cz3Choice *out) {
  // candidate padded lengths M >= 2n-2 in increasing order: 2^k and 3 * 2^k
  int64_t P2 = 1;
  while (P2 < 2 * n - 2) P2 *= 2;
  int64_t T3 = 3;
  while (T3 < 2 * n - 2) T3 *= 2;
  const int64_t cands[2] = {std::min(P2, T3), std::max(P2, T3)};
  for (int64_t M : cands) {
    bool found = false;
    // column sides are powers of two (the most balanced split, n1 <= n2)
    for (int64_t n1 = 2; n1 < M; n1 *= 2) {
      if (M % n1) break;
      const int64_t n2 = M / n1;
      if (n1 > n2) break;
      Choice ci = choose(KK_CZ_COL_IN, n1, prec, bwd);
      Choice cm = choose(KK_CZ_ROW_MID, n2, prec, bwd);
      Choice co = choose(KK_CZ_COL_OUT, n1, prec, bwd);
      if (!ci.e || !cm.e || !co.e) continue;
      out->M = M;
      out->n1 = n1;
      out->cin = ci;
      out->cmid = cm;
      out->cout = co;
      found = true;
    }
    if (found) return true;
  }
  return false;
}

static bool smooth7(int64_t v) {
  for (int p : {2, 3, 5, 7})
    while (v % p == 0) v /= p;
  return v == 1;
}

std::unique_ptr<Node> build_bluestein(const Problem &p, std::string *why) {
  const bool bwd = p.dir > 0;
  const int64_t id = p.howmany > 1 ? p.idist : 0;
  const int64_t od = p.howmany > 1 ? p.odist : 0;
  auto finish = [&](std::unique_ptr<FusedConvNode> node) -> std::unique_ptr<Node> {
    hipError_t err = node->init();
    if (err != hipSuccess) {
      *why = std::string("convolution table setup failed: ") + hipGetErrorString(err);
      return nullptr;
    }
    return node;
  };
  // Rader for primes whose n-1 has an LDS kernel
  if (p.n > 2 && spf(p.n) == p.n) {
    Choice c = choose(KK_RADER, p.n - 1, p.prec, bwd);
    if (c.e)
      return finish(std::make_unique<FusedConvNode>(true, c, p.n, p.n - 1, p.prec, p.howmany,
                                                    p.istride, id, p.ostride, od));
  }
  // The padded length: the cyclic convolution with b_k = b_{M-k} is exact
  // for M >= 2n-2, one less than the general 2n-1, because b is even
  // (b_{-k} = b_k): at M = 2n-2 the two lags +-(n-1) share one slot that
  // holds the same value. (n = 65537: M = 2^17 instead of 2^18.) It must
  // stay far from int64 overflow and from any length with a four-step plan.
  if (p.n > (int64_t(1) << 40)) {
    *why = "length " + std::to_string(p.n) + " is beyond the chirp-z range";
    return nullptr;
  }
  // fused Bluestein: a power of two with a compiled chain, else a 7-smooth length
  int64_t P2 = 1;
  while (P2 < 2 * p.n - 2) P2 *= 2;
  Choice c = choose(KK_BLUE, P2, p.prec, bwd);
  int64_t M = P2;
  if (!c.e || !c.e->n) {
    int64_t S = 2 * p.n - 2;
    while (!smooth7(S)) ++S;
    Choice cs = choose(KK_BLUE, S, p.prec, bwd);
    if (cs.e) {
      c = cs;
      M = S;
    }
  }
  if (c.e)
    return finish(std::make_unique<FusedConvNode>(false, c, p.n, M, p.prec, p.howmany, p.istride,
                                                  id, p.ostride, od));
  // Rader in three passes for a prime whose n - 1 splits into compiled
  // column and row sides: N = n - 1 points instead of the chirp-z's M >= 2n - 2
  if (p.n > 2 && spf(p.n) == p.n && p.n - 1 <= (int64_t(1) << 30)) {
    int64_t n1 = 0;
    Choice ci, cm, co;
    if (rd3_choice(p.n - 1, p.prec, bwd, &n1, &ci, &cm, &co)) {
      const int64_t N = p.n - 1;
      const int64_t chunk = fourstep_chunk(N, p.prec, p.howmany);
      auto node = std::make_unique<Rader3Node>(p.n, n1, N / n1, p.prec, bwd, p.howmany,
                                               p.istride, id, p.ostride, od, chunk, ci, cm, co);
      const hipError_t err = node->init();
      if (err != hipSuccess) {
        *why = std::string("Rader setup failed: ") + hipGetErrorString(err);
        return nullptr;
      }
      return node;
    }
  }
  // chirp-z in three passes when compiled column and row engines cover a
  // split of a padded length 2^k or 3 * 2^k
  {
    Cz3Choice cz;
    if (cz3_choice(p.n, p.prec, bwd, &cz)) {
      const int64_t chunk = fourstep_chunk(cz.M, p.prec, p.howmany);
      auto node = std::make_unique<ChirpZ3Node>(p.n, cz.n1, cz.M / cz.n1, p.prec, bwd,
                                                p.howmany, p.istride, id, p.ostride, od, chunk,
                                                cz.cin, cz.cmid, cz.cout);
      const hipError_t err = node->init();
      if (err != hipSuccess) {
        *why = std::string("chirp-z setup failed: ") + hipGetErrorString(err);
        return nullptr;
      }
      return node;
    }
  }
  // chirp-z around the four-step transform
  const int64_t per = P2 * 2 * prec_compute_bytes(p.prec);
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(p.howmany, (int64_t(64) << 20) / per));
  auto node = std::make_unique<BigBluesteinNode>(p.n, P2, p.prec, bwd, p.howmany, p.istride, id,
                                                 p.ostride, od, chunk);
  hipError_t err = node->init(why);
  if (err != hipSuccess) {
    if (why->empty()) *why = std::string("chirp-z setup failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  return node;
}

}  // namespace plf
