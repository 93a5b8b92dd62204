// api.cpp -- the C ABI of include/plfft.h and include/plfft_gpu.h.
//
// Configuration objects, defaults, status strings and the validation order
// follow the reference implementation (src/plfft.cpp:22-526); plan creation
// hands the validated problem to the GPU planner, and execute launches the
// plan's kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/plfft.h"
#include "../../include/plfft_gpu.h"
#include "abi_plan.hpp"
#include "mixed_plan.hpp"
#include "plan.hpp"

namespace {

enum transform_kind_t { TRANSFORM_C2C = 0, TRANSFORM_R2C = 1, TRANSFORM_C2R = 2, TRANSFORM_R2R = 3 };

}  // namespace

// src/plfft.cpp:33-46
struct plfft_config_t {
  int64_t n;
  int64_t howmany;
  int64_t istride;
  int64_t idist;
  int64_t ostride;
  int64_t odist;
  int transform_kind;
  int payload;  // direction (C2C) or r2r kind (R2R)
  plfft_data_type_t data_type;
  plfft_io_alias_t alias;
  plfft_sme_mode_t sme_mode;
  plfft_rigor_t rigor;
};

[[noreturn]] static void out_of_memory() {
  std::fputs("Fatal error: out of memory\n", stderr);
  std::abort();
}

// ---------------------------------------------------------------------------
// validation helpers (src/plfft.cpp:66-141)
// ---------------------------------------------------------------------------
static bool valid_batch(const plfft_config_t *c) {
  if (c->howmany == 1) return c->istride > 0 && c->ostride > 0;
  return c->howmany > 0 && c->istride > 0 && c->idist > 0 && c->ostride > 0 && c->odist > 0;
}
static bool valid_direction(int d) { return d == PLFFT_FORWARD || d == PLFFT_BACKWARD; }
static bool valid_r2r_kind(int k) { return k >= PLFFT_R2R_DCT_1 && k <= PLFFT_R2R_HC2R; }
static bool valid_transform_kind(const plfft_config_t *c) {
  switch (c->transform_kind) {
    case TRANSFORM_C2C:
      return valid_direction(c->payload);
    case TRANSFORM_R2C:
    case TRANSFORM_C2R:
      return true;
    case TRANSFORM_R2R:
      return valid_r2r_kind(c->payload);
  }
  return false;
}
static bool valid_alias(const plfft_config_t *c) {
  return c->alias == PLFFT_IO_NO_ALIAS || c->alias == PLFFT_IO_MAY_ALIAS;
}
static bool valid_sme_mode(const plfft_config_t *c) {
  return c->sme_mode == PLFFT_SME_DISABLED || c->sme_mode == PLFFT_SME_ENABLED;
}
static bool want_sme(const plfft_config_t *c) { return c->sme_mode == PLFFT_SME_ENABLED; }
static bool valid_rigor(const plfft_config_t *c) {
  return c->rigor >= PLFFT_ESTIMATE && c->rigor <= PLFFT_EXHAUSTIVE;
}

extern "C" {

const char *plfft_status_to_string(plfft_status_t status) {
  switch (status) {  // src/plfft.cpp:263-296
    case PLFFT_OK: return "success";
    case PLFFT_NULL_ARGUMENT: return "null argument";
    case PLFFT_INVALID_SIZE: return "invalid transform size";
    case PLFFT_UNSUPPORTED_BATCH: return "Unsupported batch or stride configuration";
    case PLFFT_INVALID_TRANSFORM_KIND: return "invalid transform kind";
    case PLFFT_INVALID_DATA_TYPE: return "invalid data type";
    case PLFFT_INVALID_IO_ALIAS: return "invalid I/O alias mode";
    case PLFFT_PLAN_CREATION_FAILED: return "plan creation failed";
    case PLFFT_INVALID_SME_MODE: return "invalid SME mode";
    case PLFFT_SME_NOT_ENABLED: return "SME support was not enabled when building PLFFT";
    case PLFFT_SME_UNSUPPORTED_R2R: return "SME is not supported for real-to-real transforms";
    case PLFFT_FP16_NOT_AVAILABLE: return "native FP16 support is required but unavailable on this CPU";
    case PLFFT_SME_NOT_AVAILABLE: return "SME support is unavailable on this CPU";
    case PLFFT_SME_UNSUPPORTED_FIXED_POINT: return "SME is not supported for fixed-point transforms";
    case PLFFT_INVALID_PLANNER_RIGOR: return "invalid planning rigor level";
  }
  return "unknown status";
}

// src/plfft.cpp:303-319: every creator installs the same defaults, rigor
// MEASURE included.
static plfft_status_t create_config(plfft_config_t **config, int64_t n, int kind, int payload) {
  if (config == nullptr) return PLFFT_NULL_ARGUMENT;
  *config = nullptr;
  void *mem = std::malloc(sizeof(plfft_config_t));
  if (!mem) out_of_memory();
  plfft_config_t *c = new (mem) plfft_config_t;
  c->n = n;
  c->howmany = 1;
  c->istride = 1;
  c->idist = n;
  c->ostride = 1;
  c->odist = n;
  c->transform_kind = kind;
  c->payload = payload;
  c->data_type = PLFFT_DATA_TYPE_FP32;
  c->alias = PLFFT_IO_NO_ALIAS;
  c->sme_mode = PLFFT_SME_DISABLED;
  c->rigor = PLFFT_MEASURE;
  *config = c;
  return PLFFT_OK;
}

plfft_status_t plfft_config_create(plfft_config_t **config, int64_t n, plfft_direction_t direction) {
  if (config == nullptr) return PLFFT_NULL_ARGUMENT;
  *config = nullptr;
  if (!valid_direction((int)direction)) return PLFFT_INVALID_TRANSFORM_KIND;
  return create_config(config, n, TRANSFORM_C2C, (int)direction);
}

plfft_status_t plfft_config_create_r2c(plfft_config_t **config, int64_t n) {
  return create_config(config, n, TRANSFORM_R2C, 0);
}

plfft_status_t plfft_config_create_c2r(plfft_config_t **config, int64_t n) {
  return create_config(config, n, TRANSFORM_C2R, 0);
}

plfft_status_t plfft_config_create_r2r(plfft_config_t **config, int64_t n, plfft_r2r_kind_t r2r_kind) {
  if (config == nullptr) return PLFFT_NULL_ARGUMENT;
  *config = nullptr;
  if (!valid_r2r_kind((int)r2r_kind)) return PLFFT_INVALID_TRANSFORM_KIND;
  return create_config(config, n, TRANSFORM_R2R, (int)r2r_kind);
}

void plfft_config_destroy(plfft_config_t *config) {
  if (!config) return;
  config->~plfft_config_t();
  std::free(config);
}

void plfft_config_set_data_type(plfft_config_t *c, plfft_data_type_t t) { c->data_type = t; }
plfft_data_type_t plfft_config_get_data_type(const plfft_config_t *c) { return c->data_type; }
void plfft_config_set_io_alias(plfft_config_t *c, plfft_io_alias_t a) { c->alias = a; }
plfft_io_alias_t plfft_config_get_io_alias(const plfft_config_t *c) { return c->alias; }
void plfft_config_set_sme_mode(plfft_config_t *c, plfft_sme_mode_t m) { c->sme_mode = m; }
plfft_sme_mode_t plfft_config_get_sme_mode(const plfft_config_t *c) { return c->sme_mode; }
void plfft_config_set_rigor(plfft_config_t *c, plfft_rigor_t r) { c->rigor = r; }
plfft_rigor_t plfft_config_get_rigor(const plfft_config_t *c) { return c->rigor; }

void plfft_config_set_batch(plfft_config_t *c, int64_t howmany, int64_t istride, int64_t idist,
                            int64_t ostride, int64_t odist) {
  c->howmany = howmany;
  c->istride = istride;
  c->idist = idist;
  c->ostride = ostride;
  c->odist = odist;
}
int64_t plfft_config_get_howmany(const plfft_config_t *c) { return c->howmany; }
int64_t plfft_config_get_istride(const plfft_config_t *c) { return c->istride; }
int64_t plfft_config_get_idist(const plfft_config_t *c) { return c->idist; }
int64_t plfft_config_get_ostride(const plfft_config_t *c) { return c->ostride; }
int64_t plfft_config_get_odist(const plfft_config_t *c) { return c->odist; }

}  // extern "C"

// Logical element counts and byte sizes of the input and output sides.
static void side_sizes(const plf::Problem &p, size_t *in_elem, size_t *out_elem,
                       int64_t *in_count, int64_t *out_count) {
  size_t real_bytes = p.fixed_bits == 15 ? 2 : p.fixed_bits == 7 ? 1 : (size_t)plf::prec_real_bytes(p.prec);
  size_t cplx_bytes = 2 * real_bytes;
  switch (p.kind) {
    case plf::K_C2C:
      *in_elem = *out_elem = cplx_bytes;
      *in_count = *out_count = p.n;
      break;
    case plf::K_R2C:
      *in_elem = real_bytes;
      *out_elem = cplx_bytes;
      *in_count = p.n;
      *out_count = p.n / 2 + 1;
      break;
    case plf::K_C2R:
      *in_elem = cplx_bytes;
      *out_elem = real_bytes;
      *in_count = p.n / 2 + 1;
      *out_count = p.n;
      break;
    default:
      *in_elem = *out_elem = real_bytes;
      *in_count = *out_count = p.n;
      break;
  }
}

extern "C" plfft_status_t plfft_plan_create_from_config(plfft_plan_t **plan, const plfft_config_t *config) {
  // validation order of src/plfft.cpp:427-517
  if (!plan) return PLFFT_NULL_ARGUMENT;
  *plan = nullptr;
  if (!config) return PLFFT_NULL_ARGUMENT;
  if (!(config->n > 0)) return PLFFT_INVALID_SIZE;
  if (!valid_batch(config)) return PLFFT_UNSUPPORTED_BATCH;
  if (!valid_transform_kind(config)) return PLFFT_INVALID_TRANSFORM_KIND;
  if (!valid_alias(config)) return PLFFT_INVALID_IO_ALIAS;
  if (!valid_sme_mode(config)) return PLFFT_INVALID_SME_MODE;
  if (want_sme(config) && config->transform_kind == TRANSFORM_R2R) return PLFFT_SME_UNSUPPORTED_R2R;
  if (want_sme(config) &&
      (config->data_type == PLFFT_DATA_TYPE_Q7 || config->data_type == PLFFT_DATA_TYPE_Q15))
    return PLFFT_SME_UNSUPPORTED_FIXED_POINT;
  // This build has no SME support (a reference build without PLFFT_ENABLE_SME).
  if (want_sme(config)) return PLFFT_SME_NOT_ENABLED;
  // FP16 is native on the GPU: PLFFT_FP16_NOT_AVAILABLE never applies.
  if (!valid_rigor(config)) return PLFFT_INVALID_PLANNER_RIGOR;

  plf::Problem p;
  switch ((int)config->data_type) {
    case PLFFT_DATA_TYPE_FP32: p.prec = plf::P_F32; break;
    case PLFFT_DATA_TYPE_FP64: p.prec = plf::P_F64; break;
    case PLFFT_DATA_TYPE_FP16: p.prec = plf::P_F16; break;
    case 5 /* PLFFT_DATA_TYPE_BF16 */: p.prec = plf::P_BF16; break;
    case PLFFT_DATA_TYPE_Q7: p.prec = plf::P_F32; p.fixed_bits = 7; break;
    case PLFFT_DATA_TYPE_Q15: p.prec = plf::P_F32; p.fixed_bits = 15; break;
    default: return PLFFT_INVALID_DATA_TYPE;
  }
  // fixed-point R2R has no implementation (src/plfft.cpp:248-260)
  // DCT-I is undefined for n == 1 (docs/plfft.md:208-209)
  if (config->transform_kind == TRANSFORM_R2R && config->payload == PLFFT_R2R_DCT_1 &&
      config->n == 1) {
    plf::set_last_error("DCT-I needs n > 1");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  if (p.fixed_bits && config->transform_kind == TRANSFORM_R2R) {
    plf::set_last_error("fixed-point real-to-real transforms are not supported");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  p.kind = config->transform_kind;
  if (p.kind == TRANSFORM_C2C) p.dir = config->payload;
  else if (p.kind == TRANSFORM_R2C) p.dir = PLFFT_FORWARD;
  else if (p.kind == TRANSFORM_C2R) p.dir = PLFFT_BACKWARD;
  else p.r2r_kind = config->payload;
  p.n = config->n;
  p.howmany = config->howmany;
  p.istride = config->istride;
  p.ostride = config->ostride;
  // howmany == 1 ignores the distances (src/plfft.cpp:75-81)
  p.idist = config->howmany == 1 ? 0 : config->idist;
  p.odist = config->howmany == 1 ? 0 : config->odist;
  p.may_alias = config->alias == PLFFT_IO_MAY_ALIAS;
  p.rigor = config->rigor;

  int dev = 0, ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    plf::set_last_error("no HIP device available");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  if (hipGetDevice(&dev) != hipSuccess) {
    plf::set_last_error("hipGetDevice failed");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  std::string why;
  plf::set_planning_rigor(p.rigor);
  std::unique_ptr<plf::Node> node = plf::build_node(p, &why);
  plf::set_planning_rigor(0);
  if (!node) {
    plf::set_last_error(why);
    return PLFFT_PLAN_CREATION_FAILED;
  }
  void *mem = std::malloc(sizeof(plfft_plan_t));
  if (!mem) out_of_memory();
  plfft_plan_t *pl = new (mem) plfft_plan_t;
  pl->prob = p;
  pl->device = dev;
  pl->ws.bytes = node->workspace_bytes();
  pl->node = std::move(node);
  side_sizes(p, &pl->in_elem, &pl->out_elem, &pl->in_count, &pl->out_count);
  plf::set_last_error("");
  *plan = pl;
  return PLFFT_OK;
}

extern "C" void plfft_plan_destroy(plfft_plan_t *plan) {
  if (!plan) return;
  plan->~plfft_plan_t();
  std::free(plan);
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
namespace {

using plf::DeviceGuard;

bool is_device_pointer(const void *p) {
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof(a));
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void *acquire_ws(plf::WsPool &pool) {
  if (pool.bytes == 0) return nullptr;
  {
    std::lock_guard<std::mutex> g(pool.mu);
    if (!pool.free_list.empty()) {
      void *p = pool.free_list.back();
      pool.free_list.pop_back();
      return p;
    }
  }
  void *p = nullptr;
  if (hipMalloc(&p, pool.bytes) != hipSuccess) out_of_memory();
  return p;
}

void release_ws(plf::WsPool &pool, void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(pool.mu);
  pool.free_list.push_back(p);
}

[[noreturn]] void fatal_hip(const char *what, hipError_t e) {
  std::fprintf(stderr, "plfft: %s failed: %s\n", what, hipGetErrorString(e));
  std::abort();
}

// Span in bytes of a strided batched side.
size_t span_bytes(int64_t count, int64_t stride, int64_t howmany, int64_t dist, size_t elem) {
  int64_t span = (howmany - 1) * dist + (count - 1) * stride + 1;
  return (size_t)span * elem;
}

void run_sync(const plfft_plan_t *pl, const plf::Node *node, plf::WsPool &pool, const void *in,
              void *out,
              size_t in_elem, int64_t in_count, int64_t is, int64_t id, size_t out_elem,
              int64_t out_count, int64_t os, int64_t od) {
  DeviceGuard guard(pl->device);
  const int64_t hm = pl->prob.howmany;
  const bool din = is_device_pointer(in), dout = is_device_pointer(out);
  void *ws = acquire_ws(pool);
  hipStream_t s = nullptr;  // legacy default stream: ordered after other blocking streams
  hipError_t e;
  if (din && dout) {
    e = node->enqueue(in, out, ws, s);
    if (e != hipSuccess) fatal_hip("kernel launch", e);
  } else {
    // Host buffers: stage through device memory, keeping any overlap.
    const size_t ib = span_bytes(in_count, is, hm, id, in_elem);
    const size_t ob = span_bytes(out_count, os, hm, od, out_elem);
    const char *ilo = static_cast<const char *>(in), *ihi = ilo + ib;
    char *olo = static_cast<char *>(out), *ohi = olo + ob;
    const bool overlap = ilo < ohi && olo < ihi;
    if (overlap) {
      const char *lo = ilo < olo ? ilo : olo;
      const char *hi = ihi > ohi ? ihi : ohi;
      void *d = nullptr;
      if ((e = hipMalloc(&d, (size_t)(hi - lo))) != hipSuccess) out_of_memory();
      if ((e = hipMemcpy(d, lo, (size_t)(hi - lo), hipMemcpyHostToDevice)) != hipSuccess) fatal_hip("copy in", e);
      char *db = static_cast<char *>(d);
      e = node->enqueue(db + (ilo - lo), db + (olo - lo), ws, s);
      if (e != hipSuccess) fatal_hip("kernel launch", e);
      if ((e = hipMemcpy(olo, db + (olo - lo), ob, hipMemcpyDeviceToHost)) != hipSuccess) fatal_hip("copy out", e);
      (void)hipFree(d);
    } else {
      void *di = nullptr, *dout_p = nullptr;
      if (din) di = const_cast<void *>(in);
      else {
        if (hipMalloc(&di, ib) != hipSuccess) out_of_memory();
        if ((e = hipMemcpy(di, in, ib, hipMemcpyHostToDevice)) != hipSuccess) fatal_hip("copy in", e);
      }
      if (dout) dout_p = out;
      else {
        if (hipMalloc(&dout_p, ob) != hipSuccess) out_of_memory();
        // the gaps of a strided output must come back unchanged
        if ((e = hipMemcpy(dout_p, out, ob, hipMemcpyHostToDevice)) != hipSuccess) fatal_hip("copy out-init", e);
      }
      e = node->enqueue(di, dout_p, ws, s);
      if (e != hipSuccess) fatal_hip("kernel launch", e);
      if (!dout) {
        if ((e = hipMemcpy(out, dout_p, ob, hipMemcpyDeviceToHost)) != hipSuccess) fatal_hip("copy out", e);
        (void)hipFree(dout_p);
      }
      if (!din) (void)hipFree(di);
    }
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) fatal_hip("execute", e);
  release_ws(pool, ws);
}

}  // namespace

extern "C" void plfft_plan_execute(const plfft_plan_t *plan, const void *input, void *output) {
  const plf::Problem &p = plan->prob;
  run_sync(plan, plan->node.get(), plan->ws, input, output, plan->in_elem, plan->in_count,
           p.istride,
           p.idist, plan->out_elem, plan->out_count, p.ostride, p.odist);
}

extern "C" size_t plfft_plan_get_workspace_size(const plfft_plan_t *plan) {
  return plan ? plan->ws.bytes : 0;
}

static void *default_ws(plf::WsPool &pool) {
  if (pool.bytes == 0) return nullptr;
  std::lock_guard<std::mutex> g(pool.mu);
  if (!pool.dflt) {
    if (hipMalloc(&pool.dflt, pool.bytes) != hipSuccess) out_of_memory();
  }
  return pool.dflt;
}

extern "C" plfft_status_t plfft_plan_execute_async(const plfft_plan_t *plan, const void *input,
                                                   void *output, void *workspace, void *stream) {
  if (!plan || !input || !output) return PLFFT_NULL_ARGUMENT;
  DeviceGuard guard(plan->device);
  void *ws = workspace ? workspace : default_ws(plan->ws);
  hipError_t e = plan->node->enqueue(input, output, ws, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    plf::set_last_error(std::string("launch failed: ") + hipGetErrorString(e));
    return PLFFT_PLAN_CREATION_FAILED;
  }
  return PLFFT_OK;
}

// Adjoint problem: the linear map's transpose with input/output layouts
// swapped.
static const plf::Node *adjoint_node(const plfft_plan_t *pl) {
  std::lock_guard<std::mutex> g(pl->adj_mu);
  if (!pl->adj_tried) {
    pl->adj_tried = true;
    plf::Problem q = pl->prob;
    std::swap(q.istride, q.ostride);
    std::swap(q.idist, q.odist);
    bool ok = true;
    switch (q.kind) {
      case plf::K_C2C:  // adjoint of the unnormalized DFT: opposite direction
        q.dir = -q.dir;
        break;
      case plf::K_R2C:  // C2R with the interior bins halved
        q.kind = plf::K_C2R;
        q.dir = PLFFT_BACKWARD;
        q.adjoint = 1;
        break;
      case plf::K_C2R:  // R2C with the interior bins doubled
        q.kind = plf::K_R2C;
        q.dir = PLFFT_FORWARD;
        q.adjoint = 1;
        break;
      case plf::K_R2R:
        // DCT-II <-> DCT-III and DST-II <-> DST-III with end weights;
        // R2HC <-> HC2R with interior weights; DCT-I with interior weights;
        // DCT-IV, DST-I, DST-IV and DHT are symmetric (their own adjoint)
        switch (q.r2r_kind) {
          case PLFFT_R2R_DCT_2: q.r2r_kind = PLFFT_R2R_DCT_3; q.adjoint = 1; break;
          case PLFFT_R2R_DCT_3: q.r2r_kind = PLFFT_R2R_DCT_2; q.adjoint = 1; break;
          case PLFFT_R2R_DST_2: q.r2r_kind = PLFFT_R2R_DST_3; q.adjoint = 1; break;
          case PLFFT_R2R_DST_3: q.r2r_kind = PLFFT_R2R_DST_2; q.adjoint = 1; break;
          case PLFFT_R2R_R2HC: q.r2r_kind = PLFFT_R2R_HC2R; q.adjoint = 1; break;
          case PLFFT_R2R_HC2R: q.r2r_kind = PLFFT_R2R_R2HC; q.adjoint = 1; break;
          case PLFFT_R2R_DCT_1: q.adjoint = 1; break;
          case PLFFT_R2R_DCT_4:
          case PLFFT_R2R_DST_1:
          case PLFFT_R2R_DST_4:
          case PLFFT_R2R_DHT: break;
          default: ok = false;
        }
        break;
      default:
        ok = false;
    }
    if (ok) {
      std::string why;
      DeviceGuard guard(pl->device);
      pl->adj = plf::build_node(q, &why);
      if (!pl->adj) plf::set_last_error(why);
      else pl->adj_ws.bytes = pl->adj->workspace_bytes();
    } else {
      plf::set_last_error("no adjoint for this transform kind");
    }
  }
  return pl->adj.get();
}

extern "C" void plfft_plan_execute_adjoint(const plfft_plan_t *plan, const void *input, void *output) {
  const plf::Node *n = adjoint_node(plan);
  if (!n) {
    std::fprintf(stderr, "plfft: adjoint unavailable: %s\n", plf::last_error());
    std::abort();
  }
  const plf::Problem &p = plan->prob;
  run_sync(plan, n, plan->adj_ws, input, output, plan->out_elem, plan->out_count, p.ostride,
           p.odist,
           plan->in_elem, plan->in_count, p.istride, p.idist);
}

extern "C" plfft_status_t plfft_plan_execute_adjoint_async(const plfft_plan_t *plan,
                                                           const void *input, void *output,
                                                           void *workspace, void *stream) {
  if (!plan || !input || !output) return PLFFT_NULL_ARGUMENT;
  const plf::Node *n = adjoint_node(plan);
  if (!n) return PLFFT_PLAN_CREATION_FAILED;
  // a caller workspace is sized by plfft_plan_get_workspace_size(), the
  // forward node's need
  if (workspace && plan->adj_ws.bytes > plan->ws.bytes) {
    plf::set_last_error("the adjoint needs " + std::to_string(plan->adj_ws.bytes) +
                        " bytes of workspace, more than the plan's " +
                        std::to_string(plan->ws.bytes) + "; pass NULL to use the plan's own");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  DeviceGuard guard(plan->device);
  void *ws = workspace ? workspace : default_ws(plan->adj_ws);
  hipError_t e = n->enqueue(input, output, ws, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    plf::set_last_error(std::string("launch failed: ") + hipGetErrorString(e));
    return PLFFT_PLAN_CREATION_FAILED;
  }
  return PLFFT_OK;
}

extern "C" size_t plfft_plan_describe(const plfft_plan_t *plan, char *buf, size_t len) {
  std::string s = plan ? plan->node->describe() : std::string("(null)");
  if (buf && len) {
    size_t k = s.size() < len - 1 ? s.size() : len - 1;
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return s.size();
}

extern "C" const char *plfft_last_error(void) { return plf::last_error(); }

// ---------------------------------------------------------------------------
// mixed-length batches (no reference counterpart: a reference caller loops
// over one plan per length)
// ---------------------------------------------------------------------------
struct plfft_mixed_plan_t {
  std::unique_ptr<plf::MixedPlan> mp;
  int device = 0;
  size_t elem = 0;  // bytes per complex storage element
  mutable std::mutex ws_mu;
  mutable void *ws_default = nullptr;
  ~plfft_mixed_plan_t() {
    if (ws_default) (void)hipFree(ws_default);
  }
};

extern "C" plfft_status_t plfft_mixed_plan_create(plfft_mixed_plan_t **plan, int64_t count,
                                                  const int64_t *n, const int64_t *in_offset,
                                                  const int64_t *out_offset,
                                                  plfft_direction_t direction,
                                                  plfft_data_type_t data_type,
                                                  plfft_rigor_t rigor) {
  if (!plan) return PLFFT_NULL_ARGUMENT;
  *plan = nullptr;
  if (!n) return PLFFT_NULL_ARGUMENT;
  if (count < 1) return PLFFT_INVALID_SIZE;
  for (int64_t i = 0; i < count; ++i) {
    if (n[i] < 1) return PLFFT_INVALID_SIZE;
    if ((in_offset && in_offset[i] < 0) || (out_offset && out_offset[i] < 0))
      return PLFFT_INVALID_SIZE;
  }
  if (!valid_direction(direction)) return PLFFT_INVALID_TRANSFORM_KIND;
  int prec = plf::P_F32, fixed_bits = 0;
  switch ((int)data_type) {
    case PLFFT_DATA_TYPE_FP32: prec = plf::P_F32; break;
    case PLFFT_DATA_TYPE_FP64: prec = plf::P_F64; break;
    case PLFFT_DATA_TYPE_FP16: prec = plf::P_F16; break;
    case 5 /* PLFFT_DATA_TYPE_BF16 */: prec = plf::P_BF16; break;
    case PLFFT_DATA_TYPE_Q7: fixed_bits = 7; break;
    case PLFFT_DATA_TYPE_Q15: fixed_bits = 15; break;
    default: return PLFFT_INVALID_DATA_TYPE;
  }
  if ((int)rigor < PLFFT_ESTIMATE || (int)rigor > PLFFT_EXHAUSTIVE)
    return PLFFT_INVALID_PLANNER_RIGOR;
  int dev = 0, ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    plf::set_last_error("no HIP device available");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  if (hipGetDevice(&dev) != hipSuccess) {
    plf::set_last_error("hipGetDevice failed");
    return PLFFT_PLAN_CREATION_FAILED;
  }
  std::string why;
  plf::set_planning_rigor(rigor);
  std::unique_ptr<plf::MixedPlan> mp =
      plf::MixedPlan::build(prec, fixed_bits, direction, count, n, in_offset, out_offset, &why);
  plf::set_planning_rigor(0);
  if (!mp) {
    plf::set_last_error(why);
    return PLFFT_PLAN_CREATION_FAILED;
  }
  plfft_mixed_plan_t *pl = new (std::nothrow) plfft_mixed_plan_t;
  if (!pl) out_of_memory();
  pl->mp = std::move(mp);
  pl->device = dev;
  pl->elem = fixed_bits ? (fixed_bits == 15 ? 4 : 2) : 2 * (size_t)plf::prec_real_bytes(prec);
  plf::set_last_error("");
  *plan = pl;
  return PLFFT_OK;
}

extern "C" void plfft_mixed_plan_destroy(plfft_mixed_plan_t *plan) { delete plan; }

extern "C" size_t plfft_mixed_plan_get_workspace_size(const plfft_mixed_plan_t *plan) {
  return plan ? plan->mp->workspace_bytes() : 0;
}

extern "C" int64_t plfft_mixed_plan_get_span(const plfft_mixed_plan_t *plan, int which) {
  if (!plan) return 0;
  return which ? plan->mp->out_span() : plan->mp->in_span();
}

static void *mixed_default_ws(const plfft_mixed_plan_t *pl) {
  if (pl->mp->workspace_bytes() == 0) return nullptr;
  std::lock_guard<std::mutex> g(pl->ws_mu);
  if (!pl->ws_default) {
    if (hipMalloc(&pl->ws_default, pl->mp->workspace_bytes()) != hipSuccess) out_of_memory();
  }
  return pl->ws_default;
}

extern "C" plfft_status_t plfft_mixed_plan_execute_async(const plfft_mixed_plan_t *plan,
                                                         const void *input, void *output,
                                                         void *workspace, void *stream) {
  if (!plan || !input || !output) return PLFFT_NULL_ARGUMENT;
  DeviceGuard guard(plan->device);
  void *ws = workspace ? workspace : mixed_default_ws(plan);
  hipError_t e = plan->mp->enqueue(input, output, ws, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    plf::set_last_error(std::string("launch failed: ") + hipGetErrorString(e));
    return PLFFT_PLAN_CREATION_FAILED;
  }
  return PLFFT_OK;
}

extern "C" plfft_status_t plfft_mixed_plan_execute(const plfft_mixed_plan_t *plan,
                                                   const void *input, void *output) {
  if (!plan || !input || !output) return PLFFT_NULL_ARGUMENT;
  DeviceGuard guard(plan->device);
  const plf::MixedPlan &mp = *plan->mp;
  const size_t ib = (size_t)mp.in_span() * plan->elem, ob = (size_t)mp.out_span() * plan->elem;
  const bool din = is_device_pointer(input), dout = is_device_pointer(output);
  void *ws = nullptr;
  if (mp.workspace_bytes() && hipMalloc(&ws, mp.workspace_bytes()) != hipSuccess)
    out_of_memory();
  hipStream_t s = nullptr;
  hipError_t e;
  const void *di = input;
  void *dout_p = output;
  if (!din || !dout) {
    if (input == output) {
      // in place on the host: one staging buffer
      const size_t b = ib > ob ? ib : ob;
      if (hipMalloc(&dout_p, b) != hipSuccess) out_of_memory();
      if ((e = hipMemcpy(dout_p, input, ib, hipMemcpyHostToDevice)) != hipSuccess)
        fatal_hip("copy in", e);
      di = dout_p;
    } else {
      if (!din) {
        void *d = nullptr;
        if (hipMalloc(&d, ib) != hipSuccess) out_of_memory();
        if ((e = hipMemcpy(d, input, ib, hipMemcpyHostToDevice)) != hipSuccess)
          fatal_hip("copy in", e);
        di = d;
      }
      if (!dout) {
        if (hipMalloc(&dout_p, ob) != hipSuccess) out_of_memory();
        // gaps between members come back unchanged
        if ((e = hipMemcpy(dout_p, output, ob, hipMemcpyHostToDevice)) != hipSuccess)
          fatal_hip("copy out-init", e);
      }
    }
  }
  e = mp.enqueue(di, dout_p, ws, s);
  if (e != hipSuccess) fatal_hip("kernel launch", e);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) fatal_hip("execute", e);
  if (!dout) {
    if ((e = hipMemcpy(output, dout_p, ob, hipMemcpyDeviceToHost)) != hipSuccess)
      fatal_hip("copy out", e);
    (void)hipFree(dout_p);
  }
  if (!din && di != dout_p) (void)hipFree(const_cast<void *>(di));
  if (ws) (void)hipFree(ws);
  return PLFFT_OK;
}

extern "C" size_t plfft_mixed_plan_describe(const plfft_mixed_plan_t *plan, char *buf,
                                            size_t len) {
  std::string s = plan ? plan->mp->describe() : std::string("(null)");
  if (buf && len) {
    size_t k = s.size() < len - 1 ? s.size() : len - 1;
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return s.size();
}
