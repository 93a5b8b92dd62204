// mixed_plan.cpp -- one persistent launch for a batch of mixed lengths.
//
// The reference executes a batch as one plan per length (src/plfft.cpp:
// 283-330 builds a plan for a single n; a caller with many lengths loops over
// plans). Here the whole batch is one kernel launch in which each member is
// transformed by one block of 256 threads (kernels_mixed.hpp).
// Each member gets one of four schedules:
//   direct      smooth n whose radix chain fits one block (n <= 4096)
//   four-step   smooth n above that: column tiles, then row tiles, through
//               a workspace region of its own
//   Bluestein   other n with a padded length M <= 4096: one block
//   chirp-z     other n: Bluestein around two four-step transforms of M
// (the same algorithms as the single-length planner: Stockham passes,
// src/planner/planner.cpp:38-47 for the factor choice, and Bluestein with
// the chirp phase k^2 mod 2n in integers, src/bluestein.hpp:66-98).
#include "mixed_plan.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>

namespace plf {

namespace {

constexpr int kPts = 16;           // points per thread of the mixed engines
constexpr int64_t kOneBlock = 4096;  // longest engine length of a one-block member
constexpr int64_t kMaxSide = 512;    // longest side of a four-step tile pass

int64_t padded(int64_t n) { return n + (n >> 5) + 1; }

bool radix_ok(int r) {
  switch (r) {
    case 2: case 3: case 4: case 5: case 7: case 8: case 11: case 13: case 16:
    case 17: case 19: case 23:
      return true;
    default:
      return false;
  }
}

std::vector<int> chain_of(int64_t n) { return n == 1 ? std::vector<int>{} : dynamic_chain(n); }
bool smooth(int64_t n) { return n == 1 || !dynamic_chain(n).empty(); }

bool smooth7(int64_t v) {
  for (int p : {2, 3, 5, 7})
    while (v % p == 0) v /= p;
  return v == 1;
}

// Every stage of the chain fits `tpt` threads with kPts points each.
bool engine_fits(int64_t n, int tpt) {
  if (n == 1) return true;
  const std::vector<int> ch = chain_of(n);
  if (ch.empty()) return false;
  for (int r : ch) {
    if (!radix_ok(r)) return false;
    if (n / r > (int64_t)((kPts + r - 1) / r) * tpt) return false;
  }
  return true;
}

// Tile width of a four-step pass over length `len` (16 lanes of 16 threads,
// or 8 lanes of 32 threads), 0 if neither engine fits.
int lanes_for(int64_t len) {
  if (len > kMaxSide) return 0;
  if (engine_fits(len, 16)) return 16;
  if (engine_fits(len, 32)) return 8;
  return 0;
}

struct Split {
  int64_t n1 = 0, n2 = 0;
  int l0 = 0, l1 = 0;
};

// The split N = n1 * n2 with the least padded tile work, then the most
// balanced one.
bool best_split(int64_t N, Split *out) {
  bool found = false;
  int64_t best_work = 0, best_gap = 0;
  for (int64_t n1 = 2; n1 <= kMaxSide && n1 < N; ++n1) {
    if (N % n1) continue;
    const int64_t n2 = N / n1;
    if (n2 > kMaxSide || n2 < 2) continue;
    const int l0 = lanes_for(n1), l1 = lanes_for(n2);
    if (!l0 || !l1) continue;
    const int64_t work = (n2 + l0 - 1) / l0 * l0 * n1 + (n1 + l1 - 1) / l1 * l1 * n2;
    const int64_t gap = std::llabs(n1 - n2);
    if (!found || work < best_work || (work == best_work && gap < best_gap)) {
      found = true;
      best_work = work;
      best_gap = gap;
      *out = Split{n1, n2, l0, l1};
    }
  }
  return found;
}

struct Schedule {
  int kind = -1;  // 0 direct, 1 four-step, 2 Bluestein one block, 3 chirp-z
  int64_t len = 0;  // one-block engine length (n or M); M for chirp-z
  Split sp;
};

bool schedule_for(int64_t n, Schedule *s, std::string *why) {
  if (n <= kOneBlock && engine_fits(n, 256)) {
    s->kind = 0;
    s->len = n;
    return true;
  }
  if (smooth(n) && best_split(n, &s->sp)) {
    s->kind = 1;
    s->len = n;
    return true;
  }
  for (int64_t M = 2 * n - 1; M <= kOneBlock; ++M)
    if (smooth7(M) && engine_fits(M, 256)) {
      s->kind = 2;
      s->len = M;
      return true;
    }
  for (int64_t M = std::max<int64_t>(2 * n - 1, kOneBlock + 1); M <= kMaxSide * kMaxSide; ++M)
    if (smooth7(M) && best_split(M, &s->sp)) {
      s->kind = 3;
      s->len = M;
      return true;
    }
  std::ostringstream o;
  o << "mixed batch: no schedule for length " << n;
  *why = o.str();
  return false;
}

}  // namespace

std::unique_ptr<MixedPlan> MixedPlan::build(int prec, int dir, int64_t count, const int64_t *n,
                                            const int64_t *in_off, const int64_t *out_off,
                                            std::string *why) {
  auto plan = std::unique_ptr<MixedPlan>(new MixedPlan);
  MixedPlan &P = *plan;
  P.prec_ = prec;
  P.bwd_ = dir > 0 ? 1 : 0;
  P.count_ = count;
  const int csize = 2 * prec_compute_bytes(prec);

  std::map<int64_t, Schedule> sched;
  for (int64_t i = 0; i < count; ++i) {
    if (sched.count(n[i])) continue;
    Schedule s;
    if (!schedule_for(n[i], &s, why)) return nullptr;
    sched[n[i]] = s;
  }

  // engine descriptors and twiddles, one per distinct engine length
  std::vector<GenericDesc> descs;
  std::vector<int64_t> desc_tw;
  std::vector<double> tw;
  std::map<int64_t, int> desc_of;
  auto get_desc = [&](int64_t len) -> int {
    auto it = desc_of.find(len);
    if (it != desc_of.end()) return it->second;
    const std::vector<int> ch = chain_of(len);
    GenericDesc d{};
    d.n = (int)len;
    d.ns = (int)ch.size();
    int64_t L = 1, off = 0;
    for (int s = 0; s < d.ns; ++s) {
      d.radix[s] = ch[s];
      d.L[s] = (int)L;
      d.twoff[s] = (int)off;
      if (s > 0) off += L * (ch[s] - 1);
      L *= ch[s];
    }
    desc_tw.push_back((int64_t)(tw.size() / 2));
    const std::vector<double> t = stockham_twiddles(ch);
    tw.insert(tw.end(), t.begin(), t.end());
    descs.push_back(d);
    const int idx = (int)descs.size() - 1;
    desc_of[len] = idx;
    return idx;
  };

  // aux: split twiddles per four-step length, chirps and B per Bluestein n
  std::vector<double> aux;
  std::map<int64_t, std::pair<int64_t, int>> split_of;  // N -> (offset, shift)
  std::map<int64_t, int64_t> chirp_of, B_of;
  auto get_split = [&](int64_t N) -> std::pair<int64_t, int> {
    auto it = split_of.find(N);
    if (it != split_of.end()) return it->second;
    int lg = 0;
    while ((int64_t(1) << lg) < N) ++lg;
    const int shift = (lg + 1) / 2;
    const int64_t nlo = int64_t(1) << shift, nhi = ((N - 1) >> shift) + 1;
    const int64_t off = (int64_t)(aux.size() / 2);
    for (int64_t r = 0; r < nlo; ++r) {
      double re, im;
      unit_root(r, N, &re, &im);
      aux.push_back(re);
      aux.push_back(-im);
    }
    for (int64_t q = 0; q < nhi; ++q) {
      double re, im;
      unit_root(q * nlo, N, &re, &im);
      aux.push_back(re);
      aux.push_back(-im);
    }
    return split_of[N] = {off, shift};
  };
  // Bluestein members grouped by M, for batched FP64 DFTs of b
  std::map<int64_t, std::vector<int64_t>> blue_by_M;
  for (const auto &kv : sched)
    if (kv.second.kind >= 2) {
      const int64_t nn = kv.first;
      chirp_of[nn] = (int64_t)(aux.size() / 2);
      const std::vector<double> c = chirp_table(nn);
      aux.insert(aux.end(), c.begin(), c.end());
      blue_by_M[kv.second.len].push_back(nn);
    }
  for (const auto &kv : blue_by_M) {
    const int64_t M = kv.first;
    const std::vector<int64_t> &ns = kv.second;
    std::vector<double> b(2 * (size_t)M * ns.size(), 0.0);
    const double s = 1.0 / (double)M;
    for (size_t j = 0; j < ns.size(); ++j) {
      const int64_t nn = ns[j];
      const double *c = aux.data() + 2 * chirp_of[nn];
      double *row = b.data() + 2 * (size_t)M * j;
      for (int64_t k = 0; k < nn; ++k) {
        const double re = c[2 * k] * s, im = -c[2 * k + 1] * s;  // conj(c_k) / M
        row[2 * k] = re;
        row[2 * k + 1] = im;
        if (k) {
          row[2 * (M - k)] = re;
          row[2 * (M - k) + 1] = im;
        }
      }
    }
    const hipError_t err = dft_f64_gpu(b, M, (int64_t)ns.size());
    if (err != hipSuccess) {
      *why = std::string("mixed batch: Bluestein table setup failed: ") + hipGetErrorString(err);
      return nullptr;
    }
    for (size_t j = 0; j < ns.size(); ++j) {
      B_of[ns[j]] = (int64_t)(aux.size() / 2);
      aux.insert(aux.end(), b.begin() + 2 * (size_t)M * j, b.begin() + 2 * (size_t)M * (j + 1));
    }
  }

  // members, their workspace and their predicted cost (HBM elements moved
  // plus an LDS-work term)
  std::vector<MixedXf> xfs((size_t)count);
  std::vector<double> cost((size_t)count);
  int64_t ws = 0, ipos = 0, opos = 0;
  size_t lds = 0;
  int nkind[4] = {0, 0, 0, 0};
  for (int64_t i = 0; i < count; ++i) {
    const int64_t nn = n[i];
    const Schedule &s = sched[nn];
    MixedXf &x = xfs[(size_t)i];
    x.n = (int)nn;
    x.kind = s.kind;
    x.len = (int)s.len;
    x.in_off = in_off ? in_off[i] : ipos;
    x.out_off = out_off ? out_off[i] : opos;
    ipos += nn;
    opos += nn;
    P.in_span_ = std::max(P.in_span_, x.in_off + nn);
    P.out_span_ = std::max(P.out_span_, x.out_off + nn);
    x.d0 = x.d1 = 0;
    x.ws_off = 0;
    x.chirp_off = chirp_of.count(nn) ? chirp_of[nn] : 0;
    x.B_off = B_of.count(nn) ? B_of[nn] : 0;
    ++nkind[s.kind];
    const double lwork = 0.1 * (double)s.len * std::log2((double)std::max<int64_t>(2, s.len));
    if (s.kind == 0 || s.kind == 2) {
      x.d0 = get_desc(s.len);
      lds = std::max(lds, (size_t)padded(s.len) * csize);
      cost[(size_t)i] = 2.0 * (double)nn + lwork;
      continue;
    }
    const Split &sp = s.sp;
    x.n1 = (int)sp.n1;
    x.n2 = (int)sp.n2;
    x.lanes0 = sp.l0;
    x.lanes1 = sp.l1;
    x.d0 = get_desc(sp.n1);
    x.d1 = get_desc(sp.n2);
    const auto spl = get_split(s.len);
    x.tw4_off = spl.first;
    x.shift = spl.second;
    x.ws_off = ws;
    ws += s.kind == 1 ? s.len : 2 * s.len;
    lds = std::max(lds, (size_t)sp.l0 * padded(sp.n1) * csize);
    lds = std::max(lds, (size_t)sp.l1 * padded(sp.n2) * csize);
    cost[(size_t)i] = s.kind == 1 ? 4.0 * (double)nn + lwork : 8.0 * (double)s.len + 2 * lwork;
  }
  // the most expensive first; blocks take them in snake order
  std::vector<int> order((size_t)count);
  for (int64_t i = 0; i < count; ++i) order[(size_t)i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return cost[(size_t)a] > cost[(size_t)b]; });
  P.data_bytes_ = (size_t)ws * csize;
  P.lds_ = std::max<size_t>(lds, 16);

  hipError_t err = P.xf_.upload(xfs.data(), xfs.size() * sizeof(MixedXf));
  if (err == hipSuccess) err = P.order_.upload(order.data(), order.size() * sizeof(int));
  if (err == hipSuccess) err = P.descs_.upload(descs.data(), descs.size() * sizeof(GenericDesc));
  if (err == hipSuccess) err = P.desc_tw_.upload(desc_tw.data(), desc_tw.size() * sizeof(int64_t));
  if (err == hipSuccess) err = upload_table(P.tw_, tw, prec);
  if (err == hipSuccess) err = upload_table(P.aux_, aux, prec);
  if (err != hipSuccess) {
    *why = std::string("mixed batch: table upload failed: ") + hipGetErrorString(err);
    return nullptr;
  }
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const int per_cu = mixed_blocks_per_cu(prec, P.lds_);
  if (per_cu <= 0) {
    *why = "mixed batch: the kernel does not fit a CU with this LDS size";
    return nullptr;
  }
  P.grid_ = (int)std::min<int64_t>(count, (int64_t)per_cu * cus);

  std::ostringstream o;
  o << "(mixed count=" << count << " direct=" << nkind[0] << " fourstep=" << nkind[1]
    << " bluestein=" << nkind[2] << " chirpz=" << nkind[3] << " lengths=" << sched.size()
    << " grid=" << P.grid_ << " lds=" << P.lds_
    << " workspace=" << P.workspace_bytes() << ")";
  P.summary_ = o.str();
  return plan;
}

hipError_t MixedPlan::enqueue(const void *in, void *out, void *ws, hipStream_t s) const {
  if (count_ == 0) return hipSuccess;
  MixedArgs a{};
  a.in = in;
  a.out = out;
  a.ws = ws;
  a.xf = static_cast<const MixedXf *>(xf_.p);
  a.order = static_cast<const int *>(order_.p);
  a.count = count_;
  a.descs = static_cast<const GenericDesc *>(descs_.p);
  a.desc_tw = static_cast<const int64_t *>(desc_tw_.p);
  a.tw = tw_.p;
  a.aux = aux_.p;
  a.bwd = bwd_;
  return launch_mixed(prec_, a, grid_, lds_, s);
}

}  // namespace plf
