// The generated fixed-point codelets (perf-libs-fft_amd/csrc/
// fixed_codelets.hpp) compiled for the host with a scalar ops policy of the
// device's semantics (saturating add/sub, per-lane rounding multiplies,
// one-rounding multiply-accumulates in a wrapping 2B+2-bit accumulator), so
// that a CPU test can compare the generated op order and constants with the
// oracle's interpretation of the reference's op lists.
#define __device__
#define __forceinline__ inline
#include <stdint.h>

#include "fixed_codelets.hpp"

namespace {

template <int B>
struct HostOps {
  struct V {
    int re, im;
  };
  static int sat(int v) {
    const int hi = (1 << B) - 1, lo = -(1 << B);
    return v > hi ? hi : (v < lo ? lo : v);
  }
  static int wrap(unsigned acc) {
    if (B == 7) return (int)(int16_t)(uint16_t)acc;
    return (int)acc;
  }
  static int narrow(unsigned acc) {
    const int w = wrap(acc);
    return sat((w >> B) + ((w >> (B - 1)) & 1));
  }
  static V add(V a, V b) { return {sat(a.re + b.re), sat(a.im + b.im)}; }
  static V sub(V a, V b) { return {sat(a.re - b.re), sat(a.im - b.im)}; }
  template <bool REV>
  static V mul(V a, int q0, int q1) {
    const int l0 = REV ? a.im : a.re, l1 = REV ? a.re : a.im;
    return {narrow((unsigned)(l0 * q0)), narrow((unsigned)(l1 * q1))};
  }
  template <int SG, bool RA, bool RB>
  static V mac(V a, int qa0, int qa1, V b, int qb0, int qb1) {
    const int a0 = RA ? a.im : a.re, a1 = RA ? a.re : a.im;
    const int b0 = RB ? b.im : b.re, b1 = RB ? b.re : b.im;
    const unsigned pa0 = (unsigned)(a0 * qa0), pa1 = (unsigned)(a1 * qa1);
    const unsigned pb0 = (unsigned)(b0 * qb0), pb1 = (unsigned)(b1 * qb1);
    return SG > 0 ? V{narrow(pa0 + pb0), narrow(pa1 + pb1)}
                  : V{narrow(pa0 - pb0), narrow(pa1 - pb1)};
  }
};

template <int B>
int run(int r, const int32_t *x, int32_t *y) {
  using O = HostOps<B>;
  typename O::V a[16];
  for (int m = 0; m < r; ++m) a[m] = {x[2 * m], x[2 * m + 1]};
  switch (r) {
    case 2: plf::fx_dft2<O, B>(a); break;
    case 4: plf::fx_dft4<O, B>(a); break;
    case 8: plf::fx_dft8<O, B>(a); break;
    case 16: plf::fx_dft16<O, B>(a); break;
    default: return -1;
  }
  for (int m = 0; m < r; ++m) {
    y[2 * m] = a[m].re;
    y[2 * m + 1] = a[m].im;
  }
  return 0;
}

}  // namespace

extern "C" int host_fixed_codelet(int r, int bits, const int32_t *x, int32_t *y) {
  return bits == 15 ? run<15>(r, x, y) : bits == 7 ? run<7>(r, x, y) : -1;
}
