// twiddles.cpp -- FP64 twiddle tables (host side).
//
// As in the reference (src/generate_twiddles.hpp:72-136), every factor is
// computed in double from an integer-exact angle and rounded only when it is
// stored in the kernel's compute precision.
#include <cmath>
#include <string>

#include "plan.hpp"

namespace plf {

void unit_root(int64_t num, int64_t den, double *re, double *im) {
  int64_t r = num % den;
  if (r < 0) r += den;
  // reduce to an angle in [0, pi/4] with exact integer arithmetic over 8*den
  __int128 D = (__int128)den * 8;
  __int128 t = (__int128)r * 8;
  double ss = 1.0, cs = 1.0;
  if (2 * t > D) {
    t = D - t;
    ss = -1.0;
  }
  if (4 * t > D) {
    t = D / 2 - t;
    cs = -1.0;
  }
  bool swap = false;
  if (8 * t > D) {
    t = D / 4 - t;
    swap = true;
  }
  const long double a = 6.283185307179586476925286766559005768L *
                        (long double)(int64_t)t / (long double)(int64_t)D;
  long double c = cosl(a), s = sinl(a);
  if (swap) {
    long double tmp = c;
    c = s;
    s = tmp;
  }
  *re = (double)(cs * c);
  *im = (double)(ss * s);
}

std::vector<double> stockham_twiddles(const std::vector<int> &chain) {
  std::vector<double> tw;
  int64_t L = chain.empty() ? 1 : chain[0];
  for (size_t s = 1; s < chain.size(); ++s) {
    const int64_t R = chain[s];
    for (int64_t m = 1; m < R; ++m) {
      for (int64_t j = 0; j < L; ++j) {
        double re, im;
        unit_root(j * m, L * R, &re, &im);
        tw.push_back(re);
        tw.push_back(-im);  // forward sign
      }
    }
    L *= R;
  }
  return tw;
}

static thread_local std::string g_last_error;

void set_last_error(const std::string &s) { g_last_error = s; }
const char *last_error() { return g_last_error.c_str(); }

}  // namespace plf

namespace plf {

std::vector<double> xlane_twiddles(int64_t n) {
  std::vector<double> tw;
  for (int64_t h = n / 2; h >= 2; h /= 2)
    for (int64_t j = 0; j < n; ++j) {
      double re = 1.0, im = 0.0;
      if (j & h) {
        unit_root(j % h, 2 * h, &re, &im);  // exp(+2 pi i p / 2h)
        im = -im;
      }
      tw.push_back(re);
      tw.push_back(im);
    }
  return tw;
}

}  // namespace plf

namespace plf {

std::vector<uint16_t> mfma_dft16_table() {
  // operand o = 2 kh + part (kh: K half, part 0 high / 1 low), lane l,
  // element i at ((o * 64) + l) * 8 + i
  std::vector<uint16_t> t(4 * 64 * 8);
  for (int kh = 0; kh < 2; ++kh)
    for (int l = 0; l < 64; ++l)
      for (int i = 0; i < 8; ++i) {
        const int row = l % 32, h = l / 32, k = row % 16, m = 8 * h + i;
        double cr, si;
        unit_root((int64_t)k * m, 16, &cr, &si);  // cos a, sin a, a = 2 pi k m / 16
        // rows 0..15: [Wr, -Wi] = [cos, sin]; rows 16..31: [Wi, Wr] = [-sin, cos]
        const float w = (float)(row < 16 ? (kh == 0 ? cr : si) : (kh == 0 ? -si : cr));
        const _Float16 hi = (_Float16)w;
        const _Float16 lo = (_Float16)(w - (float)hi);
        t[((2 * kh + 0) * 64 + l) * 8 + i] = __builtin_bit_cast(uint16_t, hi);
        t[((2 * kh + 1) * 64 + l) * 8 + i] = __builtin_bit_cast(uint16_t, lo);
      }
  return t;
}

}  // namespace plf
