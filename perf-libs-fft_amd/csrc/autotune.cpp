// autotune.cpp -- timed choice among compiled kernel configurations.
//
// The GPU counterpart of the reference's timed planner
// (src/planner/planner.cpp:286-372) and its benchmarker
// (src/planner/benchmarker.hpp:33-104, benchmarker.cpp:18-91):
//   * rigor sets a planning-time budget and a decision margin with the
//     reference's formulas (src/plfft.cpp:149-184): ESTIMATE times nothing;
//     MEASURE, PATIENT and EXHAUSTIVE get exp(a + b ln n) seconds (n < 1500:
//     exp(-13.4096 + 0.6539 ln n)) and margins 0.02 / 0.05 / 0.5;
//   * the candidates of a length are the registry entries compiled for it,
//     in registry order: the first is the incumbent (the ESTIMATE choice),
//     later ones challenge it;
//   * a candidate is measured in rounds of `niters` launches on a private
//     non-blocking stream with HIP events: niters doubles until a round
//     takes 0.25 ms, then 3 to 7 rounds give a mean and a standard
//     deviation; a challenger replaces the incumbent when a Welch t-test
//     says it is faster with probability >= 1 - margin, the incumbent stays
//     when that probability is <= margin, and an undecided comparison goes
//     to the lower mean;
//   * timing stops when the budget is spent (a candidate being measured
//     stops with its rounds so far); the incumbent always gets measured
//     first, so a spent budget leaves the registry's choice.
// Winners are cached in memory per (kernel family, n, precision, direction,
// device architecture) and, when $PLFFT_WISDOM names a file, appended to it
// and read back by later processes. ESTIMATE ignores wisdom, so it is
// deterministic.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "plan.hpp"

namespace plf {

namespace {

std::mutex g_mu;
std::map<std::string, std::string> g_wisdom;  // key -> entry name
std::string g_loaded_path;                    // wisdom file last read ("" none)
thread_local int g_rigor = 0;
thread_local std::chrono::steady_clock::time_point g_start;

// $PLFFT_WISDOM, or "" (no file: wisdom lives for the process only)
std::string wisdom_path() {
  const char *p = std::getenv("PLFFT_WISDOM");
  return p ? std::string(p) : std::string();
}

// (Re)read the wisdom file when $PLFFT_WISDOM names one not read yet; its
// entries override the in-memory ones.
void load_locked() {
  const std::string path = wisdom_path();
  if (path.empty() || path == g_loaded_path) return;
  g_loaded_path = path;
  std::ifstream f(path);
  std::string key, name;
  while (f >> key >> name) g_wisdom[key] = name;
}

void store_locked(const std::string &key, const std::string &name) {
  g_wisdom[key] = name;
  const std::string path = wisdom_path();
  if (path.empty()) return;
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    const std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)  // mkdir -p
      if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
  }
  std::ofstream f(path, std::ios::app);
  if (f) f << key << " " << name << "\n";
}

// gcnArchName and CU count of the current device, e.g. "gfx950:sramecc+:xnack-/256"
std::string device_tag() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return "unknown";
  }
  std::ostringstream o;
  o << prop.gcnArchName << "/" << prop.multiProcessorCount;
  return o.str();
}

double elapsed_secs() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - g_start).count();
}

// Mean and standard deviation of a sample built one value at a time.
struct Sample {
  double mean = 0, sd = 0, n = 0;
  void add(double x) {
    const double m = mean + (x - mean) / (n + 1);
    // sum of squared deviations about the new mean, from the old one
    const double ss = (n > 1 ? sd * sd * (n - 1) : 0.0) + (x - mean) * (x - m);
    n += 1;
    mean = m;
    sd = n > 1 ? std::sqrt(std::max(0.0, ss / (n - 1))) : 0.0;
  }
};

constexpr double kMinRoundSecs = 2.5e-4;
constexpr int kMinRounds = 3, kMaxRounds = 7;

// Probability-like score that `chal` is faster than `inc`: Welch's t of
// (inc - chal) through the CDF of Student's t with one degree of freedom
// (0.5 + atan(t) / pi), the reference's decision statistic. NaN if undefined.
double faster_score(const Sample &inc, const Sample &chal) {
  const double s = inc.sd * inc.sd / inc.n + chal.sd * chal.sd / chal.n;
  if (!(s > 0)) return NAN;
  const double t = (inc.mean - chal.mean) / std::sqrt(s);
  return 0.5 + std::atan(t) / 3.14159265358979323846;
}

// +1 challenger better, -1 incumbent better, 0 undecided
int decide(const Sample &inc, const Sample &chal, double margin) {
  if (inc.n < kMinRounds || chal.n < kMinRounds) return 0;
  const double p = faster_score(inc, chal);
  if (!std::isfinite(p)) return 0;
  if (p >= 1.0 - margin) return 1;
  if (p <= margin) return -1;
  return 0;
}

// Timing buffers and tables for the candidates of one length.
struct Bench {
  DevMem in, out, aux0, aux1;
  int64_t batch = 0;
};

// Launch arguments of candidate `c` (family `kind`, engine length `len`)
// over the bench batch.
KArgs bench_args(int kind, const Bench &b, const EngineData &eng, int64_t len) {
  KArgs a{};
  a.in = b.in.p;
  a.out = b.out.p;
  a.tw = eng.tw.p;
  a.desc = eng.d();
  a.aux0 = b.aux0.p;
  a.aux1 = b.aux1.p;
  a.count = b.batch;
  a.istride = a.ostride = 1;
  a.n = len;
  switch (kind) {
    case KK_R2C_E:  // 2 len reals in, len + 1 bins out
      a.idist = 2 * len;
      a.odist = len + 1;
      a.flags = 1;
      break;
    case KK_C2R_E:
      a.idist = len + 1;
      a.odist = 2 * len;
      a.flags = 1;
      break;
    case KK_DCT2_E:
    case KK_DCT3_E:  // 2 len reals in and out
      a.idist = a.odist = 2 * len;
      break;
    default:  // complex, contiguous
      a.idist = a.odist = len;
  }
  return a;
}

// One round of `niters` launches, seconds per launch (negative on failure).
double time_round(const Choice &c, const KArgs &a, int64_t len, int niters, hipStream_t s,
                  hipEvent_t e0, hipEvent_t e1) {
  const int64_t grid = (a.count + c.e->tpb - 1) / c.e->tpb;
  if (hipEventRecord(e0, s) != hipSuccess) return -1;
  for (int i = 0; i < niters; ++i)
    if (c.e->launch(a, grid, c.e->lds_bytes(len), s) != hipSuccess) return -1;
  if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1;
  float ms = 0;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
  return ms * 1e-3;
}

// Measure one candidate; stops early against `inc` once decided, and when
// the planning budget runs out. n == 0 in the result: nothing measured.
Sample measure(const Choice &c, int kind, int64_t len, const Bench &b, hipStream_t s,
               hipEvent_t e0, hipEvent_t e1, double budget, const Sample *inc,
               double margin) {
  Sample r;
  EngineData eng;
  if (eng.init(c, len) != hipSuccess) return r;
  const KArgs a = bench_args(kind, b, eng, len);
  if (time_round(c, a, len, 1, s, e0, e1) < 0) return r;  // warm-up, code load
  int niters = 1;
  while (r.n < kMaxRounds && elapsed_secs() < budget) {
    const double secs = time_round(c, a, len, niters, s, e0, e1);
    if (secs < 0) return Sample{};
    if (r.n == 0 && secs < kMinRoundSecs && niters < (1 << 16)) {
      niters *= 2;  // calibrate past clock-resolution noise
      continue;
    }
    r.add(secs / niters);
    if (inc && decide(*inc, r, margin) != 0) break;
  }
  return r;
}

}  // namespace

double planner_budget_secs(int rigor, int64_t n) {
  if (rigor <= 0) return 0.0;
  const double ln = std::log((double)n);
  if (ln < 7.31322038709) return std::exp(-13.4096 + 0.6539 * ln);  // n < 1500
  switch (rigor) {
    case 1: return std::exp(-15.5578 + 0.9562 * ln);
    case 2: return std::exp(-9.8701 + 0.7029 * ln);
    default: return std::exp(-5.0776 + 0.5344 * ln);
  }
}

double planner_margin(int rigor) {
  switch (rigor) {
    case 0: return 0.01;
    case 1: return 0.02;
    case 2: return 0.05;
    default: return 0.5;
  }
}

void set_planning_rigor(int rigor) {
  g_rigor = rigor;
  g_start = std::chrono::steady_clock::now();
}
int planning_rigor() { return g_rigor; }

std::string wisdom_key(int kind, int64_t n, int prec, bool bwd) {
  std::ostringstream o;
  o << kkind_name(kind) << ":" << n << ":" << prec_name(prec) << ":" << (bwd ? "b" : "f") << ":"
    << device_tag();
  return o.str();
}

// Among `cands` (all for the same family / n / precision / direction, in
// registry order), return the index to use.
int tune_pick(int kind, int64_t n, int prec, bool bwd, const std::vector<Choice> &cands) {
  // PLFFT_KERNEL_FILTER (diagnostics): use the first candidate whose name
  // contains the string; -1 (none of this kind) lets the planner fall back.
  if (const char *f = std::getenv("PLFFT_KERNEL_FILTER")) {
    for (size_t i = 0; i < cands.size(); ++i)
      if (cands[i].e->name().find(f) != std::string::npos) return (int)i;
    return -1;
  }
  const int rigor = planning_rigor();
  if (cands.size() <= 1 || rigor == 0) return 0;  // ESTIMATE: registry order
  const bool timeable = kind == KK_C2C || kind == KK_C2C_UNIT || kind == KK_R2C_E ||
                        kind == KK_C2R_E || kind == KK_DCT2_E || kind == KK_DCT3_E;
  if (!timeable) return 0;
  const std::string key = wisdom_key(kind, n, prec, bwd);
  {
    std::lock_guard<std::mutex> g(g_mu);
    load_locked();
    auto it = g_wisdom.find(key);
    if (it != g_wisdom.end())
      for (size_t i = 0; i < cands.size(); ++i)
        if (cands[i].e->name() == it->second) return (int)i;
  }
  // the reference's budget counts the whole problem's length; the kernel
  // families here time engine lengths, and the real kinds' engines are half
  // the transform length
  const int64_t tn = (kind == KK_R2C_E || kind == KK_C2R_E || kind == KK_DCT2_E ||
                      kind == KK_DCT3_E) ? 2 * n : n;
  const double budget = planner_budget_secs(rigor, tn);
  const double margin = planner_margin(rigor);
  // synthetic batch of ~2^24 complex elements (2^26 above MEASURE), zeros
  Bench b;
  const int64_t elems = rigor >= 2 ? (int64_t(1) << 26) : (int64_t(1) << 24);
  b.batch = std::max<int64_t>(1, elems / n);
  const size_t cbytes = 2 * (size_t)prec_compute_bytes(prec);
  const size_t bytes = (size_t)b.batch * (size_t)(2 * n + 2) * cbytes;
  if (b.in.alloc(bytes) != hipSuccess || b.out.alloc(bytes) != hipSuccess ||
      b.aux0.alloc((size_t)(2 * n + 2) * cbytes) != hipSuccess ||
      b.aux1.alloc((size_t)(2 * n + 2) * cbytes) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return 0;
  int best = 0;
  if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
      hipMemsetAsync(b.in.p, 0, bytes, s) == hipSuccess &&
      hipMemsetAsync(b.aux0.p, 0, b.aux0.bytes, s) == hipSuccess &&
      hipMemsetAsync(b.aux1.p, 0, b.aux1.bytes, s) == hipSuccess) {
    Sample inc = measure(cands[0], kind, n, b, s, e0, e1, budget, nullptr, margin);
    for (size_t i = 1; i < cands.size() && inc.n > 0 && elapsed_secs() < budget; ++i) {
      const Sample ch = measure(cands[i], kind, n, b, s, e0, e1, budget, &inc, margin);
      if (ch.n == 0) continue;
      const int d = decide(inc, ch, margin);
      if (d > 0 || (d == 0 && ch.mean < inc.mean)) {
        best = (int)i;
        inc = ch;
      }
    }
  }
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipGetLastError();
  std::lock_guard<std::mutex> g(g_mu);
  store_locked(key, cands[best].e->name());
  return best;
}

}  // namespace plf
