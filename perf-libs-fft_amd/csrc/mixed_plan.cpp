// mixed_plan.cpp -- a batch of mixed lengths in a fixed set of launches.
//
// The reference executes a batch as one plan per length (src/plfft.cpp:
// 283-330 builds a plan for a single n; a caller with many lengths loops over
// plans). Here the whole batch runs as two tile launches, one per dependency
// level of the multi-pass schedules (kernels_mixed.hpp), plus grouped
// launches of the compiled Bluestein and chirp-z kernels.
// Each member gets one of five schedules:
//   direct      smooth n whose radix chain fits one block (n <= 4096)
//   four-step   smooth n above that: column tiles, then row tiles, through
//               a workspace region of its own
//   Bluestein   other n with a padded length M <= 4096: one block (on the
//               compiled fused kernel, grouped by M, for FP32 / FP64)
//   CT-Bluestein  n = n1 * n2 with n1 smooth and n2 a one-block Bluestein
//               length: column tiles, then the rows as compiled Bluestein
//               transforms grouped by M (FP32 / FP64)
//   chirp-z     other n: Bluestein around two four-step transforms of M,
//               fused into three passes
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

// The mixed kernels compile radices 2, 3, 4, 5, 7, 8 and 16 only (fewer
// engine instances keep each level's kernel small), so every engine length
// is 7-smooth; other lengths take Bluestein.
bool radix_ok(int r) {
  switch (r) {
    case 2: case 3: case 4: case 5: case 7: case 8: case 16:
      return true;
    default:
      return false;
  }
}

bool smooth7(int64_t v) {
  for (int p : {2, 3, 5, 7})
    while (v % p == 0) v /= p;
  return v == 1;
}

std::vector<int> chain_of(int64_t n) { return n == 1 ? std::vector<int>{} : dynamic_chain(n); }
bool smooth(int64_t n) { return n == 1 || smooth7(n); }

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

// Tile width of the column pass of a Bluestein-row member, whose column
// side is often short: 256 columns of one thread each, or 64 of four, keep
// every task a full block of work.
int lanes_wide(int64_t len) {
  if (engine_fits(len, 1)) return 256;
  if (engine_fits(len, 4)) return 64;
  return lanes_for(len);
}

struct Split {
  int64_t n1 = 0, n2 = 0;
  int l0 = 0, l1 = 0;
};

// Predicted cost of the two tile passes of a split, in element units: each
// tile (one block task) costs a fixed overhead plus the elements it moves.
// A column tile reads LANES consecutive elements of each row; when n2 is not
// a multiple of LANES the rows fall on different line offsets and a tile's
// row segments straddle two lines, which costs about twice the reads.
constexpr int64_t kTileOverhead = 4096;
int64_t split_cost(const Split &sp) {
  const int64_t ncol = (sp.n2 + sp.l0 - 1) / sp.l0 + 1, nrow = (sp.n1 + sp.l1 - 1) / sp.l1;
  const int64_t colmul = (sp.n2 % sp.l0) ? 2 : 1;
  return ncol * (kTileOverhead + colmul * sp.l0 * sp.n1) + nrow * (kTileOverhead + sp.l1 * sp.n2);
}

// The split N = n1 * n2 with the least predicted tile cost, then the most
// balanced one.
bool best_split(int64_t N, Split *out) {
  bool found = false;
  int64_t best_cost = 0, best_gap = 0;
  for (int64_t n1 = 2; n1 <= kMaxSide && n1 < N; ++n1) {
    if (N % n1) continue;
    const int64_t n2 = N / n1;
    if (n2 > kMaxSide || n2 < 2) continue;
    const int l0 = lanes_for(n1), l1 = lanes_for(n2);
    if (!l0 || !l1) continue;
    const Split sp{n1, n2, l0, l1};
    const int64_t cost = split_cost(sp);
    const int64_t gap = std::llabs(n1 - n2);
    if (!found || cost < best_cost || (cost == best_cost && gap < best_gap)) {
      found = true;
      best_cost = cost;
      best_gap = gap;
      *out = sp;
    }
  }
  return found;
}

struct Schedule {
  int kind = -1;  // 0 direct, 1 four-step, 2 Bluestein one block,
                  // 3 chirp-z on the three-pass compiled kernels,
                  // 4 four-step n = n1 * n2 with Bluestein rows
  int64_t len = 0;  // one-block engine length (n or M); M for chirp-z; kind 4: M of the rows
  bool compiled = false;  // kind 2: on the compiled fused kernel (a Bluestein group)
  Split sp;
};

// The padded length of a one-block Bluestein DFT of n points on the
// compiled fused kernel (FP32 / FP64 storage): the smallest power of two
// M >= max(64, 2n - 2) up to 4096 with a compiled chain, or 0.
int64_t blue_len_compiled(int64_t n, int prec, bool bwd) {
  if (prec != P_F32 && prec != P_F64) return 0;
  int64_t M = 64;
  while (M < 2 * n - 2) M *= 2;
  if (M > kOneBlock) return 0;
  const Choice c = choose(KK_BLUE, M, prec, bwd);
  return (c.e && c.e->n) ? M : 0;
}

// The padded length of a one-block Bluestein DFT of n points: the smallest
// 7-smooth M >= 2n - 2 whose chain fits one 256-thread block, or 0.
int64_t blue_len(int64_t n) {
  for (int64_t M = std::max<int64_t>(2 * n - 2, 1); M <= kOneBlock; ++M)
    if (smooth7(M) && engine_fits(M, 256)) return M;
  return 0;
}

bool schedule_for(int64_t n, int prec, bool bwd, Schedule *s, std::string *why) {
  if (n <= kOneBlock && smooth(n) && engine_fits(n, 256)) {
    s->kind = 0;
    s->len = n;
    return true;
  }
  if (smooth(n) && best_split(n, &s->sp)) {
    s->kind = 1;
    s->len = n;
    return true;
  }
  if (const int64_t M = blue_len_compiled(n, prec, bwd)) {
    s->kind = 2;
    s->len = M;
    s->compiled = true;
    return true;
  }
  if (const int64_t M = blue_len(n)) {
    s->kind = 2;
    s->len = M;
    return true;
  }
  // Cooley-Tukey n = n1 * n2 with a smooth column side n1 and Bluestein rows
  // of n2 points on the compiled fused kernel (the reference's
  // cooley_tukey_bluestein strategies, src/planner/planner.cpp:171-176,
  // bluestein.hpp:781-941): four HBM element moves per point instead of the
  // chirp-z's padded passes. The split with the least padded row work
  // n1 * M log2 M (M = the rows' padded length) wins.
  {
    double best = 0.0;
    for (int64_t n1 = 2; n1 <= kMaxSide && n1 < n; ++n1) {
      if (n % n1 || !smooth7(n1)) continue;
      const int64_t n2 = n / n1;
      const int l0 = lanes_wide(n1);
      if (!l0) continue;
      const int64_t M = blue_len_compiled(n2, prec, bwd);
      if (!M) continue;
      const double work = (double)n1 * (double)M * std::log2((double)M);
      if (s->kind == 4 && work >= best) continue;
      best = work;
      s->kind = 4;
      s->len = M;
      s->sp = Split{n1, n2, l0, 0};
    }
    if (s->kind == 4) return true;
  }
  // chirp-z on the three-pass compiled kernels (padded length 2^k or 3 2^k)
  Cz3Choice cz;
  if (cz3_choice(n, prec, bwd, &cz)) {
    s->kind = 3;
    s->len = cz.M;
    s->sp = Split{cz.n1, cz.M / cz.n1, 0, 0};
    return true;
  }
  std::ostringstream o;
  o << "mixed batch: no schedule for length " << n;
  *why = o.str();
  return false;
}

}  // namespace

std::unique_ptr<MixedPlan> MixedPlan::build(int prec, int fixed_bits, int dir, int64_t count,
                                            const int64_t *n, const int64_t *in_off,
                                            const int64_t *out_off, std::string *why) {
  auto plan = std::unique_ptr<MixedPlan>(new MixedPlan);
  MixedPlan &P = *plan;
  P.prec_ = prec;
  P.bwd_ = dir > 0 ? 1 : 0;
  P.count_ = count;
  if (fixed_bits) {
    // The reference's fixed-point plans are power-of-two only
    // (src/planner/planner.cpp:38-47): one gathered launch per kernel tier.
    std::vector<int64_t> ns((size_t)count), io((size_t)count), oo((size_t)count);
    std::map<int64_t, int> distinct;
    int64_t ipos = 0, opos = 0;
    for (int64_t i = 0; i < count; ++i) {
      ns[(size_t)i] = n[i];
      io[(size_t)i] = in_off ? in_off[i] : ipos;
      oo[(size_t)i] = out_off ? out_off[i] : opos;
      ipos += n[i];
      opos += n[i];
      P.in_span_ = std::max(P.in_span_, io[(size_t)i] + n[i]);
      P.out_span_ = std::max(P.out_span_, oo[(size_t)i] + n[i]);
      distinct[n[i]] = 1;
    }
    int launches = 0;
    std::unique_ptr<Node> node =
        build_fixed_mixed_node(fixed_bits, dir > 0, ns, io, oo, &launches, why);
    if (!node) return nullptr;
    P.fixed_groups_.push_back(std::move(node));
    std::ostringstream o;
    o << "(mixed_fixed_q" << fixed_bits << " count=" << count << " lengths=" << distinct.size()
      << " launches=" << launches << ")";
    P.summary_ = o.str();
    return plan;
  }
  const int csize = 2 * prec_compute_bytes(prec);

  std::map<int64_t, Schedule> sched;
  for (int64_t i = 0; i < count; ++i) {
    if (sched.count(n[i])) continue;
    Schedule s;
    if (!schedule_for(n[i], prec, dir > 0, &s, why)) return nullptr;
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
  std::map<int64_t, int64_t> B_of;
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
  // Bluestein lengths grouped by M, for batched FP64 DFTs of b: members of
  // kinds 2 and 3, and the rows (n2) of kind-4 members. Chirp tables are
  // placed per member below (aligned like the member's data).
  std::map<int64_t, std::vector<double>> chirp_host;
  std::map<int64_t, std::vector<int64_t>> blue_by_M;
  for (const auto &kv : sched) {
    if (kv.second.kind < 2) continue;
    const int64_t key = kv.second.kind == 4 ? kv.second.sp.n2 : kv.first;
    if (chirp_host.count(key)) continue;
    chirp_host[key] = chirp_table(key);
    blue_by_M[kv.second.len].push_back(key);
  }
  // pad aux (complex entries) up to an offset congruent to r mod 16
  auto align_aux = [&](int64_t r) {
    while ((int64_t)(aux.size() / 2) % 16 != r) {
      aux.push_back(0.0);
      aux.push_back(0.0);
    }
  };
  for (const auto &kv : blue_by_M) {
    const int64_t M = kv.first;
    const std::vector<int64_t> &ns = kv.second;
    std::vector<double> b(2 * (size_t)M * ns.size(), 0.0);
    const double s = 1.0 / (double)M;
    for (size_t j = 0; j < ns.size(); ++j) {
      const int64_t nn = ns[j];
      const double *c = chirp_host[nn].data();
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
      const auto sit = sched.find(ns[j]);
      const double *src = b.data() + 2 * (size_t)M * j;
      align_aux(0);
      B_of[ns[j]] = (int64_t)(aux.size() / 2);
      if (sit != sched.end() && sit->second.kind == 3) {
        const Schedule &sj = sit->second;
        // chirp-z: row k1 of the middle pass reads B[k1 + n1 k2], k2 < n2,
        // so B is stored row-major by k1
        const int64_t n1 = sj.sp.n1, n2 = sj.sp.n2;
        for (int64_t k1 = 0; k1 < n1; ++k1)
          for (int64_t k2 = 0; k2 < n2; ++k2) {
            aux.push_back(src[2 * (k1 + n1 * k2)]);
            aux.push_back(src[2 * (k1 + n1 * k2) + 1]);
          }
      } else {
        aux.insert(aux.end(), src, src + 2 * (size_t)M);
      }
    }
  }
  // chirp tables per (length, residue of the member's offset mod 16)
  std::map<std::pair<int64_t, int64_t>, int64_t> chirp_at;
  auto get_chirp = [&](int64_t nn, int64_t r) -> int64_t {
    auto it = chirp_at.find({nn, r});
    if (it != chirp_at.end()) return it->second;
    align_aux(r);
    const int64_t off = (int64_t)(aux.size() / 2);
    const std::vector<double> &c = chirp_host[nn];
    aux.insert(aux.end(), c.begin(), c.end());
    return chirp_at[{nn, r}] = off;
  };

  // members, their workspace, and their tasks per dependency level with a
  // predicted cost (HBM elements moved plus an LDS-work term)
  std::vector<MixedXf> xfs((size_t)count);
  struct HostTask {
    MixedTask t;
    double cost;
  };
  std::vector<HostTask> level[2];
  int64_t ws = 0, ipos = 0, opos = 0;
  size_t lds = 0;
  int nkind[5] = {0, 0, 0, 0, 0};
  std::map<int64_t, std::vector<CzMember>> cz_members;  // by M
  std::map<int64_t, std::vector<BlueMember>> blue_members, blue_rows;  // by M
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
    // the member's line offset: workspace rows, chirp tables and column
    // tiles are placed at the same residue mod 16 as its input
    const int64_t r16 = ((x.in_off % 16) + 16) % 16;
    x.chirp_off = s.kind == 4 ? get_chirp(s.sp.n2, 0) : s.kind >= 2 ? get_chirp(nn, r16) : 0;
    const int64_t bkey = s.kind == 4 ? s.sp.n2 : nn;
    x.B_off = B_of.count(bkey) ? B_of[bkey] : 0;
    ++nkind[s.kind];
    const int xi = (int)i;
    if (s.kind == 2 && s.compiled) {
      blue_members[s.len].push_back(BlueMember{nn, x.in_off, x.out_off, 1, x.chirp_off, x.B_off});
      continue;
    }
    if (s.kind == 0 || s.kind == 2) {
      x.d0 = get_desc(s.len);
      lds = std::max(lds, (size_t)padded(s.len) * csize);
      const double lwork = 0.1 * (double)s.len * std::log2((double)std::max<int64_t>(2, s.len));
      level[0].push_back({MixedTask{MT_ONE, xi, 0, 0}, 2.0 * (double)nn + lwork});
      continue;
    }
    if (s.kind == 3) {
      cz_members[s.len].push_back(CzMember{nn, x.in_off, x.out_off, 0, x.chirp_off, x.B_off});
      continue;
    }
    const Split &sp = s.sp;
    x.n1 = (int)sp.n1;
    x.n2 = (int)sp.n2;
    x.lanes0 = sp.l0;
    x.lanes1 = sp.l1;
    x.d0 = get_desc(sp.n1);
    x.d1 = s.kind == 4 ? 0 : get_desc(sp.n2);
    const auto spl = get_split(nn);
    x.tw4_off = spl.first;
    x.shift = spl.second;
    ws += ((r16 - ws % 16) + 16) % 16;
    x.ws_off = ws;
    ws += nn;
    lds = std::max(lds, (size_t)sp.l0 * padded(sp.n1) * csize);
    if (s.kind != 4) lds = std::max(lds, (size_t)sp.l1 * padded(sp.n2) * csize);
    const double ccost = sp.l0 * (double)sp.n1 * (2.0 + 0.1 * std::log2((double)sp.n1));
    const double rcost = s.kind == 4 ? 0.0
                                     : sp.l1 * (double)sp.n2 * (2.0 + 0.1 * std::log2((double)sp.n2));
    // column tiles start where in_off + column is a multiple of the tile
    // width (a short head tile first)
    const int64_t head = (sp.l0 - r16 % sp.l0) % sp.l0;
    for (int64_t c = 0; c < sp.n2;) {
      const int64_t e = std::min<int64_t>(sp.n2, c < head ? head : c + sp.l0);
      level[0].push_back({MixedTask{MT_COL, xi, (int)c, (int)e}, ccost});
      c = e;
    }
    if (s.kind == 4) {
      // rows: workspace row k1 -> bins k1 + n1 k2 (output stride n1)
      for (int64_t k1 = 0; k1 < sp.n1; ++k1)
        blue_rows[s.len].push_back(BlueMember{sp.n2, x.ws_off + k1 * sp.n2, x.out_off + k1, sp.n1,
                                              x.chirp_off, x.B_off});
      continue;
    }
    for (int64_t r = 0; r < sp.n1; r += sp.l1)
      level[1].push_back({MixedTask{MT_ROW, xi, (int)r, (int)sp.n1}, rcost});
  }
  // chirp-z groups: workspace regions (line-aligned, M apart), the W_M
  // split table in aux, the compiled engines' twiddles, member records
  for (auto &kv : cz_members) {
    const int64_t M = kv.first;
    const Schedule &s0 = sched[kv.second[0].n];
    auto g = std::make_unique<CzGroup>();
    g->n1 = s0.sp.n1;
    g->n2 = s0.sp.n2;
    g->count = (int64_t)kv.second.size();
    g->cin = choose(KK_CZ_COL_IN, g->n1, prec, dir > 0);
    g->cmid = choose(KK_CZ_ROW_MID, g->n2, prec, dir > 0);
    g->cout = choose(KK_CZ_COL_OUT, g->n1, prec, dir > 0);
    ws += (16 - ws % 16) % 16;
    for (CzMember &m : kv.second) {
      m.ws_off = ws;
      ws += M;
    }
    const auto spl = get_split(M);
    g->split_off = spl.first;
    g->shift = spl.second;
    hipError_t err = g->ein.init(g->cin, g->n1);
    if (err == hipSuccess) err = g->emid.init(g->cmid, g->n2);
    if (err == hipSuccess) err = g->eout.init(g->cout, g->n1);
    if (err == hipSuccess)
      err = g->members.upload(kv.second.data(), kv.second.size() * sizeof(CzMember));
    if (err != hipSuccess) {
      *why = std::string("mixed batch: chirp-z group setup failed: ") + hipGetErrorString(err);
      return nullptr;
    }
    P.cz_groups_.push_back(std::move(g));
  }

  // one-block Bluestein groups: members from the input, then rows from the
  // workspace, one launch per padded length each
  for (int src = 0; src < 2; ++src)
    for (auto &kv : src ? blue_rows : blue_members) {
      auto g = std::make_unique<BlueGroup>();
      g->M = kv.first;
      g->count = (int64_t)kv.second.size();
      g->from_ws = src == 1;
      g->c = choose(KK_BLUE, g->M, prec, dir > 0);
      hipError_t err = g->c.e ? g->eng.init(g->c, g->M) : hipErrorInvalidValue;
      if (err == hipSuccess)
        err = g->members.upload(kv.second.data(), kv.second.size() * sizeof(BlueMember));
      if (err != hipSuccess) {
        *why = std::string("mixed batch: Bluestein group setup failed: ") +
               hipGetErrorString(err);
        return nullptr;
      }
      P.blue_groups_.push_back(std::move(g));
    }

  std::vector<MixedTask> tasks;
  for (int L = 0; L < 2; ++L) {
    std::stable_sort(level[L].begin(), level[L].end(),
                     [](const HostTask &a, const HostTask &b) { return a.cost > b.cost; });
    P.level_first_[L] = (int64_t)tasks.size();
    P.level_count_[L] = (int64_t)level[L].size();
    for (const HostTask &h : level[L]) tasks.push_back(h.t);
  }
  P.data_bytes_ = (size_t)ws * csize;
  P.lds_ = std::max<size_t>(lds, 16);

  hipError_t err = P.xf_.upload(xfs.data(), xfs.size() * sizeof(MixedXf));
  if (err == hipSuccess) err = P.tasks_.upload(tasks.data(), tasks.size() * sizeof(MixedTask));
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
  for (int L = 0; L < 2; ++L) {
    const int per_cu = mixed_blocks_per_cu(prec, L, P.lds_);
    if (per_cu <= 0) {
      *why = "mixed batch: the kernel does not fit a CU with this LDS size";
      return nullptr;
    }
    P.level_grid_[L] = per_cu * cus;
    P.grid_ = std::max(P.grid_, P.level_grid_[L]);
  }

  std::ostringstream o;
  o << "(mixed count=" << count << " direct=" << nkind[0] << " fourstep=" << nkind[1]
    << " bluestein=" << nkind[2] << " ct_bluestein=" << nkind[4] << " ("
    << P.blue_groups_.size() << " groups) chirpz=" << nkind[3] << " (" << P.cz_groups_.size()
    << " groups) lengths=" << sched.size() << " tasks=" << P.level_count_[0] << "+"
    << P.level_count_[1]
    << " grid=" << P.grid_ << " lds=" << P.lds_
    << " workspace=" << P.workspace_bytes() << ")";
  P.summary_ = o.str();
  return plan;
}

hipError_t MixedPlan::enqueue(const void *in, void *out, void *ws, hipStream_t s) const {
  if (!fixed_groups_.empty()) {
    for (const auto &g : fixed_groups_) {
      const hipError_t err = g->enqueue(in, out, nullptr, s);
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  MixedArgs a{};
  a.in = in;
  a.out = out;
  a.ws = ws;
  a.xf = static_cast<const MixedXf *>(xf_.p);
  a.descs = static_cast<const GenericDesc *>(descs_.p);
  a.desc_tw = static_cast<const int64_t *>(desc_tw_.p);
  a.tw = tw_.p;
  a.aux = aux_.p;
  a.bwd = bwd_;
  for (int L = 0; L < 2; ++L) {
    if (level_count_[L] == 0) continue;
    a.tasks = static_cast<const MixedTask *>(tasks_.p) + level_first_[L];
    a.ntasks = level_count_[L];
    const int grid = (int)std::min<int64_t>(level_count_[L], level_grid_[L]);
    const hipError_t err = launch_mixed(prec_, L, a, grid, lds_, s);
    if (err != hipSuccess) return err;
  }
  for (const auto &g : blue_groups_) {
    KArgs k{};
    k.in = g->from_ws ? ws : in;
    k.out = out;
    k.tw = g->eng.tw.p;
    k.desc = g->eng.d();
    k.aux0 = aux_.p;
    k.aux1 = aux_.p;
    k.aux2 = g->members.p;
    k.count = g->count;
    k.n = g->M;
    const hipError_t err = g->c.e->launch(k, (k.count + g->c.e->tpb - 1) / g->c.e->tpb,
                                          g->c.e->lds_bytes(g->M), s);
    if (err != hipSuccess) return err;
  }
  const size_t cb = 2 * (size_t)prec_compute_bytes(prec_);
  for (const auto &g : cz_groups_) {
    KArgs k{};
    k.aux0 = static_cast<const char *>(aux_.p) + (size_t)g->split_off * cb;
    k.aux2 = g->members.p;
    k.n1 = g->n1;
    k.n2 = g->n2;
    k.shift = g->shift;
    // 1. chirp-in columns
    k.in = in;
    k.out = ws;
    k.aux1 = aux_.p;
    k.tw = g->ein.tw.p;
    k.n = g->n1;
    k.count = g->count * g->n2;
    hipError_t err = g->cin.e->launch(k, (k.count + g->cin.e->tpb - 1) / g->cin.e->tpb,
                                      g->cin.e->lds_bytes(g->n1), s);
    // 2. rows (B tables in aux at the members' B offsets)
    if (err == hipSuccess) {
      k.tw = g->emid.tw.p;
      k.n = g->n2;
      k.count = g->count * g->n1;
      err = g->cmid.e->launch(k, (k.count + g->cmid.e->tpb - 1) / g->cmid.e->tpb,
                              g->cmid.e->lds_bytes(g->n2), s);
    }
    // 3. chirp-out columns
    if (err == hipSuccess) {
      k.in = ws;
      k.out = out;
      k.tw = g->eout.tw.p;
      k.n = g->n1;
      k.count = g->count * g->n2;
      err = g->cout.e->launch(k, (k.count + g->cout.e->tpb - 1) / g->cout.e->tpb,
                              g->cout.e->lds_bytes(g->n1), s);
    }
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

}  // namespace plf
