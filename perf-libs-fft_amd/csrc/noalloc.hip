// noalloc.hip -- plfft_plan_execute_noalloc: the stream-ordered execute of
// plfft_plan_execute_async with the plan's scratch taken from a fixed device
// array of this library instead of a caller workspace or a plan allocation.
// For callers that cannot allocate device memory (code under hipGraph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The array lives in this translation unit's code object; every device that
// runs it holds its own copy until the process ends. Host code never passes
// the array's name to a kernel (in host code it is only a stand-in for the
// symbol): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and that address is what the
// kernels receive.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/plfft_gpu.h"
#include "abi_plan.hpp"

namespace {

constexpr size_t kScratchBytes = PLFFT_NOALLOC_SCRATCH_BYTES;

// The scratch of every plfft_plan_execute_noalloc call on one device. Its
// contents carry over from call to call; nothing here relies on them being
// zero (the multi-pass nodes write each workspace element before reading it).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // device address of `scratch` per device (null: not looked up)

// Device address of the scratch array on `dev` (which must be current).
hipError_t scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace

extern "C" plfft_status_t plfft_plan_execute_noalloc(const plfft_plan_t *plan, const void *input,
                                                     void *output, void *stream) {
  if (!plan || !input || !output) return PLFFT_NULL_ARGUMENT;
  const plf::Node &node = *plan->node;
  const size_t need = node.min_workspace_bytes();
  if (need > kScratchBytes) {
    plf::set_last_error("plfft_plan_execute_noalloc: the plan needs at least " +
                        std::to_string(need) + " bytes of scratch, more than the " +
                        std::to_string(kScratchBytes) +
                        " bytes of PLFFT_NOALLOC_SCRATCH_BYTES; use plfft_plan_execute_async "
                        "with a caller workspace");
    return PLFFT_INVALID_SIZE;
  }
  plf::DeviceGuard guard(plan->device);
  void *ws = nullptr;
  if (node.workspace_bytes() > 0) {
    const hipError_t e = scratch_address(plan->device, &ws);
    if (e != hipSuccess) {
      plf::set_last_error(std::string("plfft_plan_execute_noalloc: scratch lookup failed: ") +
                          hipGetErrorString(e));
      return PLFFT_PLAN_CREATION_FAILED;
    }
  }
  const hipError_t e =
      node.enqueue_capped(input, output, ws, kScratchBytes, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    plf::set_last_error(std::string("plfft_plan_execute_noalloc: launch failed: ") +
                        hipGetErrorString(e));
    return PLFFT_PLAN_CREATION_FAILED;
  }
  return PLFFT_OK;
}
