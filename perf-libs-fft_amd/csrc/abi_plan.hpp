// abi_plan.hpp -- the object behind the opaque plfft_plan_t of include/plfft.h
// (the reference's plan handle, include/plfft.h:55), shared by the C ABI
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "../../include/plfft.h"
#include "plan.hpp"

namespace plf {

// Device workspaces of one node: a free list for concurrent synchronous
// executions and one default buffer for stream-ordered calls without a
// caller workspace. Sized by the node it serves.
struct WsPool {
  size_t bytes = 0;
  std::mutex mu;
  std::vector<void *> free_list;
  void *dflt = nullptr;
  ~WsPool() {
    for (void *p : free_list) (void)hipFree(p);
    if (dflt) (void)hipFree(dflt);
  }
};

// Makes `dev` current for the guard's lifetime.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != dev) {
      prev = cur;
      (void)hipSetDevice(dev);
    }
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace plf

struct plfft_plan_t {
  plf::Problem prob;
  int device = 0;
  std::unique_ptr<plf::Node> node;
  // adjoint node, built on first use, with a pool sized by its own
  // workspace (it need not equal the forward node's)
  mutable std::mutex adj_mu;
  mutable std::unique_ptr<plf::Node> adj;
  mutable bool adj_tried = false;
  mutable plf::WsPool ws;      // forward node
  mutable plf::WsPool adj_ws;  // adjoint node
  // element sizes (bytes) of the input and output sides
  size_t in_elem = 0, out_elem = 0;
  int64_t in_count = 0, out_count = 0;  // logical elements per transform
};
