// One-lane register FFTs of length N (power of two, 2..64).
//
// Each lane transforms a whole length-N sequence held in its own VGPRs, so a
// 2-D block transform needs no cross-lane traffic inside a 1-D pass: lanes
// exchange data only through the LDS transposes between the row and column
// passes. Twiddles are compile-time constants (evaluated in double by a
// constexpr series and rounded to float), which hipcc folds into instruction
// literals; the trivial factors 1 and -i / +i cost no multiplies.
//
// Forward: X[k] = sum_n x[n] exp(-2 pi i n k / N), unnormalised.
// Inverse: x[n] = sum_k X[k] exp(+2 pi i n k / N), unnormalised (callers fold
// the 1/N into a later multiply).
#pragma once

#include "fk_device.h"

namespace fk {
namespace tw {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series on [-pi/4, pi/4]; 14 terms are exact to double rounding there.
constexpr double sin_t(double x) {
  double t = x, s = x;
  for (int i = 1; i < 14; i++) {
    t *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double cos_t(double x) {
  double t = 1.0, s = 1.0;
  for (int i = 1; i < 14; i++) {
    t *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}
// cos(2 pi k / n) with octant reduction.
constexpr double cos2pi(int k, int n) {
  k = ((k % n) + n) % n;
  const double a = 2.0 * kPi * k / n;
  const int o8 = 8 * k;
  if (o8 <= n) return cos_t(a);
  if (o8 <= 3 * n) return -sin_t(a - kPi / 2);
  if (o8 <= 5 * n) return -cos_t(a - kPi);
  if (o8 <= 7 * n) return sin_t(a - 3 * kPi / 2);
  return cos_t(a - 2 * kPi);
}
// sin(2 pi k / n) = cos(2 pi (k - n/4) / n), evaluated on the 4n grid.
constexpr double sin2pi(int k, int n) { return cos2pi(4 * k - n, 4 * n); }

template <int N>
struct Table {
  float c[N];
  float s[N];
  constexpr Table() : c(), s() {
    for (int k = 0; k < N; k++) {
      c[k] = (float)cos2pi(k, N);
      s[k] = (float)sin2pi(k, N);
    }
  }
};

constexpr int bitrev(int i, int n) {
  int r = 0;
  for (int b = 1; b < n; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

} // namespace tw

// In-place radix-2 decimation-in-time FFT of (re[i], im[i]), i < N.
template <int N, bool INV>
__device__ __forceinline__ void fft_reg(float (&re)[N], float (&im)[N]) {
  constexpr tw::Table<N> T{};
  float ar[N], ai[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    ar[i] = re[tw::bitrev(i, N)];
    ai[i] = im[tw::bitrev(i, N)];
  }
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int i = 0; i < N; i += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        const int k = j * (N / len);
        const int p = i + j, q = i + j + half;
        float tr, ti;
        if (k == 0) {
          tr = ar[q];
          ti = ai[q];
        } else if (4 * k == N) {
          // forward w = -i, inverse w = +i
          tr = INV ? -ai[q] : ai[q];
          ti = INV ? ar[q] : -ar[q];
        } else {
          const float wr = T.c[k];
          const float wi = INV ? T.s[k] : -T.s[k];
          tr = wr * ar[q] - wi * ai[q];
          ti = wr * ai[q] + wi * ar[q];
        }
        const float ur = ar[p], ui = ai[p];
        ar[p] = ur + tr;
        ai[p] = ui + ti;
        ar[q] = ur - tr;
        ai[q] = ui - ti;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    re[i] = ar[i];
    im[i] = ai[i];
  }
}

} // namespace fk
