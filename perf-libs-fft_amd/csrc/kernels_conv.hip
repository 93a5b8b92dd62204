// kernels_conv.hip -- element-wise chirp-z helpers for long lengths.
#include "kernels_elem.hpp"
#include "launch.hpp"

namespace plf {

void register_kernels_r2r_f32(std::vector<KEntry> &r);
void register_kernels_r2r_f64(std::vector<KEntry> &r);
void register_kernels_r2r_h(std::vector<KEntry> &r);

void register_kernels_r2r(std::vector<KEntry> &r) {
  register_kernels_r2r_f32(r);
  register_kernels_r2r_f64(r);
  register_kernels_r2r_h(r);
}

void register_kernels_conv(std::vector<KEntry> &r) {
  register_kernels_conv_f32(r);
  register_kernels_conv_f64(r);
  register_kernels_conv_f16(r);
  register_kernels_conv_bf16(r);
}

template <typename S>
static hipError_t chirp_in_t(bool bwd, const KArgs &a, hipStream_t s) {
  const int64_t total = a.count * a.n;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (bwd) hipLaunchKernelGGL((k_chirp_in<S, true>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_chirp_in<S, false>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}
template <typename S>
static hipError_t chirp_out_t(bool bwd, const KArgs &a, hipStream_t s) {
  const int64_t total = a.count * a.n1;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (bwd) hipLaunchKernelGGL((k_chirp_out<S, true>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_chirp_out<S, false>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_chirp_in(int prec, bool bwd, const KArgs &a, hipStream_t s) {
  switch (prec) {
    case P_F32: return chirp_in_t<float>(bwd, a, s);
    case P_F64: return chirp_in_t<double>(bwd, a, s);
    case P_F16: return chirp_in_t<half_pair>(bwd, a, s);
    default: return chirp_in_t<bf16_pair>(bwd, a, s);
  }
}
hipError_t launch_chirp_out(int prec, bool bwd, const KArgs &a, hipStream_t s) {
  switch (prec) {
    case P_F32: return chirp_out_t<float>(bwd, a, s);
    case P_F64: return chirp_out_t<double>(bwd, a, s);
    case P_F16: return chirp_out_t<half_pair>(bwd, a, s);
    default: return chirp_out_t<bf16_pair>(bwd, a, s);
  }
}
hipError_t launch_pointwise_swap(int prec, const KArgs &a, hipStream_t s) {
  const int64_t total = a.count * a.n;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (prec == P_F64) hipLaunchKernelGGL((k_pointwise_swap<double>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_pointwise_swap<float>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename S>
static hipError_t real_stage_t(int which, const KArgs &a, int64_t total, hipStream_t s) {
  const dim3 grid((unsigned)((total + 255) / 256));
  switch (which) {
    case 0: hipLaunchKernelGGL((k_widen<S>), grid, dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_narrow<S>), grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_hermitian<S>), grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_realpart<S>), grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_dct_pre<S>), grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_dct_post<S>), grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_r2r_pre<S>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_r2r_post<S>), grid, dim3(256), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_real_stage(int which, int prec, const KArgs &a, int64_t total, hipStream_t s) {
  switch (prec) {
    case P_F32: return real_stage_t<float>(which, a, total, s);
    case P_F64: return real_stage_t<double>(which, a, total, s);
    case P_F16: return real_stage_t<half_pair>(which, a, total, s);
    default: return real_stage_t<bf16_pair>(which, a, total, s);
  }
}

}  // namespace plf
