// fixed_plan.cpp -- plans for Q0.15 / Q0.7 transforms.
//
// The reference plans fixed point only for power-of-two n (no Rader or
// Bluestein, src/planner/planner.cpp:38-47, src/make_level_data.hpp:128-130)
// with radix-2/4/8/16 codelets (src/providers/aot/aot_kernel_data.cpp:
// 290-295); C2C both directions, R2C forward, C2R backward
// (scripts/gen_asm_kernels/config.py:80-97,143-164,212-237). Here C2C is one
// kernel pass through LDS (kernels_fixed.hpp) for n <= 32768; R2C and C2R
// run that transform on a complex staging buffer (widen, or Hermitian
// extension, before; bins 0..n/2, or real parts, after).
#include <algorithm>
#include <cmath>
#include <sstream>

#include "plan.hpp"

namespace plf {

namespace {

// src/plfft_convert.hpp:40-62: round to float, scale by 2^bits, truncate
// toward zero with int32 saturation, then saturate to the storage width.
int32_t q_convert(double v, int bits) {
  const float f = (float)v;
  const double scaled = (double)f * (double)(1LL << bits);
  int64_t t;
  if (scaled != scaled) t = 0;
  else if (scaled >= 2147483647.0) t = 2147483647LL;
  else if (scaled <= -2147483648.0) t = -2147483648LL;
  else t = (int64_t)scaled;
  const int64_t hi = (1LL << bits) - 1, lo = -(1LL << bits);
  return (int32_t)std::min(hi, std::max(lo, t));
}

// The stage twiddle of input h of column k at a pass of length n, as the
// premultiplied pairs (Q cos t, Q(-sin t)), (Q sin t, Q cos t) with
// t = ((dir 2 pi / n) h) k evaluated in double in that order
// (src/generate_twiddles.hpp:99-123), appended in storage pairs.
void push_tw(std::vector<int8_t> &tab, int bits, int dir, int64_t n, int64_t h, int64_t k) {
  const double base = (double)dir * 2.0 * 3.141592653589793 / (double)n;
  const double t = base * (double)h * (double)k;
  const double c = std::cos(t), si = std::sin(t);
  const int32_t q[4] = {q_convert(c, bits), q_convert(-si, bits), q_convert(si, bits),
                        q_convert(c, bits)};
  for (int i = 0; i < 4; ++i) {
    if (bits == 15) {
      const int16_t v = (int16_t)q[i];
      const int8_t *b = reinterpret_cast<const int8_t *>(&v);
      tab.insert(tab.end(), b, b + 2);
    } else {
      tab.push_back((int8_t)q[i]);
    }
  }
}

bool pow2(int64_t n) { return n >= 1 && (n & (n - 1)) == 0; }

// The descriptor of n's chain and its stage twiddles (the twiddles of input
// m of column j at (m-1) L + j of each stage s >= 1, for this direction),
// appended to `tab`; the descriptor's offsets are relative to where the
// table starts.
GenericDesc fixed_tables(int bits, int64_t n, bool bwd, std::vector<int8_t> *tab) {
  std::vector<int> chain;
  if (n > 1) chain = dynamic_chain(n);
  GenericDesc d{};
  d.n = (int)n;
  d.ns = (int)chain.size();
  const int dir = bwd ? 1 : -1;
  int64_t L = 1, off = 0;
  for (int s = 0; s < d.ns; ++s) {
    const int R = chain[s];
    d.radix[s] = R;
    d.L[s] = (int)L;
    d.twoff[s] = (int)off;
    if (s > 0) {
      for (int m = 1; m < R; ++m)
        for (int64_t j = 0; j < L; ++j) push_tw(*tab, bits, dir, L * R, m, j);
      off += L * (R - 1);
    }
    L *= R;
  }
  return d;
}

// The kernel tier of length n (the first whose longest length covers it).
int fixed_tier(int64_t n) {
  for (int t = 0; t < kNumFixedTiers; ++t)
    if (n <= kFixedTiers[t].max_n) return t;
  return -1;
}

class FixedC2CNode : public Node {
 public:
  FixedC2CNode(int bits, int64_t n, bool bwd, int64_t batch, int64_t is, int64_t id, int64_t os,
               int64_t od)
      : bits_(bits), n_(n), bwd_(bwd), batch_(batch), is_(is), id_(id), os_(os), od_(od) {}

  hipError_t init() {
    tier_ = fixed_tier(n_);
    if (tier_ < 0) return hipErrorInvalidValue;
    std::vector<int8_t> tab;
    const GenericDesc d = fixed_tables(bits_, n_, bwd_, &tab);
    // n = 4096 over 16 x 16 x 16 with unit strides: the compile-time kernel
    launch_tier_ = tier_;
    if (n_ == 4096 && is_ == 1 && os_ == 1 && d.ns == 3 && d.radix[0] == 16 &&
        d.radix[1] == 16 && d.radix[2] == 16)
      launch_tier_ = kFixedTier4096;
    hipError_t err = desc_.upload(&d, sizeof(d));
    if (err == hipSuccess && !tab.empty()) err = tab_.upload(tab.data(), tab.size());
    return err;
  }

  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    return run(in, out, batch_, s);
  }
  // the transform of `count` members (count <= the planned batch)
  hipError_t run(const void *in, void *out, int64_t count, hipStream_t s) const {
    if (count <= 0) return hipSuccess;
    const FixedTier &t = kFixedTiers[tier_];
    KArgs a{};
    a.in = in;
    a.out = out;
    a.tw = tab_.p;
    a.desc = static_cast<const GenericDesc *>(desc_.p);
    a.count = count;
    a.istride = is_;
    a.idist = id_;
    a.ostride = os_;
    a.odist = od_;
    a.n = n_;
    a.flags = bwd_ ? 1 : 0;
    const int64_t grid = (count + t.tpb - 1) / t.tpb;
    const size_t lds =
        (size_t)t.tpb * (size_t)(n_ + (n_ >> 5) + 1) * (bits_ == 15 ? 4 : 2);
    return launch_fixed(bits_, launch_tier_, a, grid, lds, s);
  }

  std::string describe() const override {
    const FixedTier &t = kFixedTiers[tier_];
    std::ostringstream o;
    o << "(fixed_q" << bits_ << "_" << (bwd_ ? "bwd" : "fwd") << " n=" << n_ << " tpt=" << t.tpt
      << " tpb=" << t.tpb << (launch_tier_ == kFixedTier4096 ? " chain=16x16x16" : "")
      << " batch=" << batch_ << ")";
    return o.str();
  }

 private:
  int bits_;
  int64_t n_;
  bool bwd_;
  int64_t batch_, is_, id_, os_, od_;
  int tier_ = -1, launch_tier_ = -1;
  DevMem desc_, tab_;
};

// R2C (forward) or C2R (backward) through a complex staging buffer.
class FixedRealNode : public Node {
 public:
  FixedRealNode(bool r2c, int bits, int64_t n, int64_t batch, int64_t is, int64_t id,
                int64_t os, int64_t od, int64_t chunk)
      : r2c_(r2c), bits_(bits), n_(n), batch_(batch), is_(is), id_(id), os_(os), od_(od),
        chunk_(chunk) {}
  hipError_t init() {
    inner_ = std::make_unique<FixedC2CNode>(bits_, n_, !r2c_, chunk_, 1, n_, 1, n_);
    return inner_->init();
  }
  size_t pair_bytes() const { return bits_ == 15 ? 4 : 2; }
  size_t workspace_bytes() const override { return (size_t)chunk_ * (size_t)n_ * pair_bytes(); }
  hipError_t enqueue(const void *in, void *out, void *work, hipStream_t s) const override {
    const size_t real_b = pair_bytes() / 2;
    for (int64_t b0 = 0; b0 < batch_; b0 += chunk_) {
      const int64_t cb = std::min(chunk_, batch_ - b0);
      KArgs a{};
      a.n = n_;
      a.count = cb;
      a.istride = is_;
      a.idist = id_;
      a.ostride = os_;
      a.odist = od_;
      hipError_t err;
      if (r2c_) {
        a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * real_b;
        a.out = work;
        err = launch_fixed_elem(bits_, 0, a, cb * n_, s);
      } else {
        a.in = static_cast<const char *>(in) + (size_t)(b0 * id_) * pair_bytes();
        a.out = work;
        err = launch_fixed_elem(bits_, 2, a, cb * n_, s);
      }
      if (err == hipSuccess) err = inner_->run(work, work, cb, s);
      if (err != hipSuccess) return err;
      a.in = work;
      if (r2c_) {
        a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * pair_bytes();
        err = launch_fixed_elem(bits_, 1, a, cb * (n_ / 2 + 1), s);
      } else {
        a.out = static_cast<char *>(out) + (size_t)(b0 * od_) * real_b;
        err = launch_fixed_elem(bits_, 3, a, cb * n_, s);
      }
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  std::string describe() const override {
    std::ostringstream o;
    o << "(fixed_" << (r2c_ ? "r2c" : "c2r") << "_q" << bits_ << " n=" << n_ << " chunk=" << chunk_
      << " " << inner_->describe() << ")";
    return o.str();
  }

 private:
  bool r2c_;
  int bits_;
  int64_t n_, batch_, is_, id_, os_, od_, chunk_;
  std::unique_ptr<FixedC2CNode> inner_;
};

}  // namespace

std::unique_ptr<Node> build_fixed_node(const Problem &p, std::string *why) {
  if (!pow2(p.n)) {
    *why = "fixed-point transforms need a power-of-two length";
    return nullptr;
  }
  if (p.n > kFixedTiers[kNumFixedTiers - 1].max_n) {
    *why = "fixed-point transforms are limited to n <= 32768";
    return nullptr;
  }
  if (p.kind == K_R2R) {
    *why = "fixed-point real-to-real transforms are not supported";
    return nullptr;
  }
  if (p.adjoint && p.kind != K_C2C) {
    *why = "the adjoint of a fixed-point real transform is not provided";
    return nullptr;
  }
  hipError_t err;
  if (p.kind == K_C2C) {
    auto node = std::make_unique<FixedC2CNode>(p.fixed_bits, p.n, p.dir > 0, p.howmany,
                                               p.istride, p.idist, p.ostride, p.odist);
    err = node->init();
    if (err == hipSuccess) return node;
  } else {
    const int64_t per = p.n * (p.fixed_bits == 15 ? 4 : 2);
    const int64_t chunk =
        std::max<int64_t>(1, std::min<int64_t>(p.howmany, (int64_t(128) << 20) / per));
    auto node = std::make_unique<FixedRealNode>(p.kind == K_R2C, p.fixed_bits, p.n, p.howmany,
                                                p.istride, p.idist, p.ostride, p.odist, chunk);
    err = node->init();
    if (err == hipSuccess) return node;
  }
  *why = std::string("fixed-point table setup failed: ") + hipGetErrorString(err);
  return nullptr;
}

// Fixed-point members of one power-of-two length at explicit offsets (the
// groups of a fixed-point mixed-length batch).
// A fixed-point mixed-length batch: one gathered launch per kernel tier.
// Members are sorted by (tier, length); each block takes up to the tier's
// members-per-block of one length (FxBlock records), so every block runs a
// single stage sequence.
class FixedMixedNode : public Node {
 public:
  FixedMixedNode(int bits, bool bwd) : bits_(bits), bwd_(bwd) {}

  hipError_t init(const std::vector<int64_t> &n, const std::vector<int64_t> &in_off,
                  const std::vector<int64_t> &out_off) {
    const size_t count = n.size();
    std::vector<size_t> order(count);
    for (size_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return n[x] < n[y]; });
    std::vector<int64_t> ioff(count), ooff(count);
    std::vector<GenericDesc> descs;
    std::vector<int8_t> tab;
    const size_t tw_bytes = bits_ == 15 ? 8 : 4;  // one QTw entry
    std::vector<FxBlock> blocks[kNumFixedTiers];
    size_t i = 0;
    while (i < count) {
      const int64_t len = n[order[i]];
      size_t j = i;
      while (j < count && n[order[j]] == len) ++j;
      const int tier = fixed_tier(len);
      if (tier < 0) return hipErrorInvalidValue;
      const int64_t tw_off = (int64_t)(tab.size() / tw_bytes);
      descs.push_back(fixed_tables(bits_, len, bwd_, &tab));
      ++lengths_;
      const int tpb = kFixedTiers[tier].tpb;
      for (size_t b = i; b < j; b += tpb) {
        FxBlock r{};
        r.desc = (int32_t)descs.size() - 1;
        r.count = (int32_t)std::min<size_t>(tpb, j - b);
        r.tw_off = tw_off;
        r.member = (int64_t)b;
        blocks[tier].push_back(r);
      }
      max_n_[tier] = std::max(max_n_[tier], len);
      for (size_t k = i; k < j; ++k) {
        ioff[k] = in_off[order[k]];
        ooff[k] = out_off[order[k]];
      }
      i = j;
    }
    hipError_t err = descs_.upload(descs.data(), descs.size() * sizeof(GenericDesc));
    if (err == hipSuccess && !tab.empty()) err = tab_.upload(tab.data(), tab.size());
    if (err == hipSuccess) err = ioff_.upload(ioff.data(), count * sizeof(int64_t));
    if (err == hipSuccess) err = ooff_.upload(ooff.data(), count * sizeof(int64_t));
    for (int t = 0; t < kNumFixedTiers && err == hipSuccess; ++t) {
      nblocks_[t] = (int64_t)blocks[t].size();
      if (nblocks_[t]) err = blocks_[t].upload(blocks[t].data(), blocks[t].size() * sizeof(FxBlock));
    }
    return err;
  }

  hipError_t enqueue(const void *in, void *out, void *, hipStream_t s) const override {
    for (int t = 0; t < kNumFixedTiers; ++t) {
      if (!nblocks_[t]) continue;
      KArgs a{};
      a.in = in;
      a.out = out;
      a.tw = tab_.p;
      a.desc = static_cast<const GenericDesc *>(descs_.p);
      a.aux0 = ioff_.p;
      a.aux1 = ooff_.p;
      a.aux2 = blocks_[t].p;
      a.istride = a.ostride = 1;
      a.n = max_n_[t];
      a.flags = (bwd_ ? 1 : 0) | 2 | 4;
      const size_t lds = (size_t)kFixedTiers[t].tpb * (size_t)(max_n_[t] + (max_n_[t] >> 5) + 1) *
                         (bits_ == 15 ? 4 : 2);
      const hipError_t err = launch_fixed(bits_, t, a, nblocks_[t], lds, s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }

  int launches() const {
    int k = 0;
    for (int t = 0; t < kNumFixedTiers; ++t) k += nblocks_[t] ? 1 : 0;
    return k;
  }
  int lengths() const { return lengths_; }
  std::string describe() const override {
    std::ostringstream o;
    o << "(fixed_q" << bits_ << "_gather lengths=" << lengths_ << " launches=" << launches()
      << ")";
    return o.str();
  }

 private:
  int bits_;
  bool bwd_;
  int lengths_ = 0;
  int64_t max_n_[kNumFixedTiers] = {}, nblocks_[kNumFixedTiers] = {};
  DevMem descs_, tab_, ioff_, ooff_, blocks_[kNumFixedTiers];
};

std::unique_ptr<Node> build_fixed_mixed_node(int bits, bool bwd, const std::vector<int64_t> &n,
                                             const std::vector<int64_t> &in_off,
                                             const std::vector<int64_t> &out_off,
                                             int *launches, std::string *why) {
  for (int64_t len : n)
    if (!pow2(len) || len > kFixedTiers[kNumFixedTiers - 1].max_n) {
      *why = "fixed-point mixed batches need power-of-two lengths <= 32768";
      return nullptr;
    }
  auto node = std::make_unique<FixedMixedNode>(bits, bwd);
  const hipError_t err = node->init(n, in_off, out_off);
  if (err != hipSuccess) {
    *why = std::string("fixed-point table setup failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  *launches = node->launches();
  return node;
}

}  // namespace plf
