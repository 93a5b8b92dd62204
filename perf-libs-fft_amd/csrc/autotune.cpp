// autotune.cpp -- timed choice among compiled kernel configurations.
//
// Counterpart of the reference's timed planner (src/planner/planner.cpp:
// 286-372, benchmarker.hpp:33-86) and its per-creation "wisdom"
// (src/planner/wisdom.cpp:45-87), re-done for the GPU: the candidates for a
// length are the registry entries compiled for it; each is timed with HIP
// events on a synthetic batch; the winner is cached per (kind, n, precision,
// direction) in memory and in a wisdom file ($PLFFT_WISDOM, default
// ~/.cache/plfft_amd/wisdom.txt) so later plans -- ESTIMATE included -- reuse
// it without timing.
#include <sys/stat.h>

#include <algorithm>
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
bool g_loaded = false;
thread_local int g_rigor = 0;

std::string wisdom_path() {
  if (const char *p = std::getenv("PLFFT_WISDOM")) return p;
  const char *home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/plfft_amd/wisdom.txt";
}

void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(wisdom_path());
  std::string key, name;
  while (f >> key >> name) g_wisdom[key] = name;
}

void store_locked(const std::string &key, const std::string &name) {
  g_wisdom[key] = name;
  const std::string path = wisdom_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    std::string dir = path.substr(0, slash);
    // mkdir -p
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
  }
  std::ofstream f(path, std::ios::app);
  if (f) f << key << " " << name << "\n";
}

std::string key_of(int kind, int64_t n, int prec, bool bwd) {
  std::ostringstream o;
  o << kkind_name(kind) << ":" << n << ":" << prec_name(prec) << ":" << (bwd ? "b" : "f");
  return o.str();
}

}  // namespace

void set_planning_rigor(int rigor) { g_rigor = rigor; }
int planning_rigor() { return g_rigor; }

// Time one batched C2C candidate on `batch` contiguous members.
static float time_candidate(const Choice &c, int64_t n, int64_t batch, void *in, void *out) {
  EngineData eng;
  if (eng.init(c, n) != hipSuccess) return 1e30f;
  KArgs a{};
  a.in = in;
  a.out = out;
  a.tw = eng.tw.p;
  a.desc = eng.d();
  a.count = batch;
  a.istride = a.ostride = 1;
  a.idist = a.odist = n;
  a.n = n;
  const int64_t grid = (batch + c.e->tpb - 1) / c.e->tpb;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1e30f;
  float best = 1e30f;
  for (int w = 0; w < 2; ++w)
    if (c.e->launch(a, grid, c.e->lds_bytes(n), nullptr) != hipSuccess) best = -1;
  if (best < 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1e30f;
  }
  const int reps = 5;
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; ++r) (void)c.e->launch(a, grid, c.e->lds_bytes(n), nullptr);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// Among `cands` (all for the same kind/n/prec/dir), return the index to use.
int tune_pick(int kind, int64_t n, int prec, bool bwd, const std::vector<Choice> &cands) {
  // PLFFT_KERNEL_FILTER (diagnostics): use the first candidate whose name
  // contains the string; -1 (none of this kind) lets the planner fall back.
  if (const char *f = std::getenv("PLFFT_KERNEL_FILTER")) {
    for (size_t i = 0; i < cands.size(); ++i)
      if (cands[i].e->name().find(f) != std::string::npos) return (int)i;
    return -1;
  }
  if (cands.size() <= 1) return 0;
  const std::string key = key_of(kind, n, prec, bwd);
  {
    std::lock_guard<std::mutex> g(g_mu);
    load_locked();
    auto it = g_wisdom.find(key);
    if (it != g_wisdom.end())
      for (size_t i = 0; i < cands.size(); ++i)
        if (cands[i].e->name() == it->second) return (int)i;
  }
  if (planning_rigor() == 0) return 0;  // ESTIMATE: registry order
  if (kind != KK_C2C && kind != KK_C2C_UNIT) return 0;
  // synthetic batch: ~2^24 complex elements (rigor > MEASURE: 2^26)
  const int64_t elems = planning_rigor() >= 2 ? (int64_t(1) << 26) : (int64_t(1) << 24);
  const int64_t batch = std::max<int64_t>(1, elems / n);
  const size_t bytes = (size_t)(batch * n) * 2 * prec_real_bytes(prec);
  DevMem in, out;
  if (in.alloc(bytes) != hipSuccess || out.alloc(bytes) != hipSuccess) return 0;
  if (hipMemset(in.p, 0, bytes) != hipSuccess) return 0;
  int best = 0;
  float best_ms = 1e30f;
  for (size_t i = 0; i < cands.size(); ++i) {
    const float ms = time_candidate(cands[i], n, batch, in.p, out.p);
    if (ms < best_ms) {
      best_ms = ms;
      best = (int)i;
    }
  }
  (void)hipDeviceSynchronize();
  std::lock_guard<std::mutex> g(g_mu);
  store_locked(key, cands[best].e->name());
  return best;
}

}  // namespace plf
