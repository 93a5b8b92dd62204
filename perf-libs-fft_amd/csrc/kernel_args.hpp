// kernel_args.hpp -- POD argument blocks passed by value to the kernels.
#pragma once

#include <stdint.h>

namespace plf {

// Batched transform over `batch` members; element k of member b lives at
// base[b*dist + k*stride] (complex elements for complex data).
struct FusedArgs {
  const void *in;
  void *out;
  const void *tw;  // per-stage twiddle table in the compute precision
  int64_t batch;
  int64_t istride, idist;
  int64_t ostride, odist;
};

}  // namespace plf
