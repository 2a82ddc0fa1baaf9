// Blocks with a side above 64 (bw, bh powers of two in [4, 512], bw*bh <=
// 4096: 128x32, 256x16, 512x8 and their transposes, and the smaller
// rectangles with a long side): one 256-thread workgroup per block, the 2-D
// transforms as radix-2 decimation-in-time FFTs over the block held in LDS
// (at most 36 KB), twiddles from a host table built in double. The reference
// GPU takes these shapes with one work-group per block and LDS FFT stages
// (src/cl/gpu.cpp:70-87, src/cl/kernels.h:84-176, :368-404); its limit is
// bw*bh*8 <= 32 KB.
//
//   k_analysis_lds        window + 2-D R2C into the packed ring layout
//   k_filter_inverse_lds  temporal filter (filter_bin, every temporal mode)
//                         + 2-D C2R + synthesis window into the blocks buffer,
//                         which k_synthesize overlap-adds
#pragma once

#include "fk_kernels.h"

namespace fk {
namespace lg {

constexpr int NT = 256;  // threads per workgroup

__device__ __forceinline__ int bit_reverse(int i, int logn) {
  int r = 0;
  for (int b = 0; b < logn; b++) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// count transforms of length N (a power of two), element i of transform t at
// buf[t * ts + i * es]; forward exp(-2 pi i nk/N) or inverse exp(+...),
// unnormalised. tw[k] = (cos, sin)(2 pi k / Nmax), k < Nmax/2, stride
// Nmax/N. Every thread of the workgroup calls it; it ends with a barrier.
template <bool INV>
__device__ __forceinline__ void lds_fft(c32* buf, int N, int count, int ts, int es,
                                        const float2* tw, int twStride) {
  const int tid = threadIdx.x;
  int logn = 0;
  while ((1 << logn) < N) logn++;
  for (int q = tid; q < count * N; q += NT) {
    const int t = q / N, i = q - t * N, j = bit_reverse(i, logn);
    if (j > i) {
      c32* a = buf + t * ts + i * es;
      c32* b = buf + t * ts + j * es;
      const c32 x = *a;
      *a = *b;
      *b = x;
    }
  }
  __syncthreads();
  const int hn = N / 2;
  for (int L = 2; L <= N; L <<= 1) {
    const int half = L >> 1;
    for (int q = tid; q < count * hn; q += NT) {
      const int t = q / hn, u = q - t * hn;
      const int grp = u / half, j = u - grp * half;
      c32* p0 = buf + t * ts + (grp * L + j) * es;
      c32* p1 = p0 + half * es;
      const float2 w = tw[j * (N / L) * twStride];
      const float wr = w.x, wi = INV ? w.y : -w.y;
      const c32 a = *p0, b = *p1;
      const float tr = wr * b.x - wi * b.y, ti = wr * b.y + wi * b.x;
      *p0 = cmake(a.x + tr, a.y + ti);
      *p1 = cmake(a.x - tr, a.y - ti);
    }
    __syncthreads();
  }
}

template <int BW, int BH>
struct LShape {
  static constexpr int NROW = BH / 2, NCOL = BW / 2, FBW = BW / 2 + 1;
  static constexpr int BSTR = BH * NCOL;
  static constexpr int NMAX = BW > BH ? BW : BH;
  // LDS: Z (row pairs, NROW x BW) then Y (BH x FBW), complex
  static constexpr int ZN = NROW * BW, YN = BH * FBW;
};

}  // namespace lg

template <int BW, int BH>
__global__ __launch_bounds__(lg::NT) void k_analysis_lds(PlaneArgs P, const void* plane,
                                                         c32* spec) {
  using L = lg::LShape<BW, BH>;
  __shared__ c32 Z[L::ZN];
  __shared__ c32 Y[L::YN];
  const int tid = threadIdx.x, b = blockIdx.x;
  int x0, y0, par;
  block_origin(P, b, x0, y0, par);
  // windowed samples, two real rows packed per complex row
  for (int q = tid; q < BW * BH; q += lg::NT) {
    const int ly = q / BW, lx = q - ly * BW;
    const size_t i = (size_t)mirror(y0 + ly, P.H) * P.W + mirror(x0 + lx, P.W);
    float v;
    if (P.pix == 8) v = (float)static_cast<const uint8_t*>(plane)[i] * (1.0f / 255.0f);
    else if (P.pix == 16) v = (float)static_cast<const uint16_t*>(plane)[i] * P.inNorm;
    else v = static_cast<const float*>(plane)[i];
    v = (v - P.chromaNorm) * P.winAna[ly * BW + lx];
    float* z = reinterpret_cast<float*>(Z + (ly >> 1) * BW + lx);
    z[ly & 1] = v;
  }
  __syncthreads();
  lg::lds_fft<false>(Z, BW, L::NROW, BW, 1, P.tw, L::NMAX / BW);
  // unpack the row pairs: X_2p[k] = (Z[k] + conj Z[-k]) / 2,
  // X_2p+1[k] = (Z[k] - conj Z[-k]) / 2i, k <= BW/2
  for (int q = tid; q < L::NROW * L::FBW; q += lg::NT) {
    const int p = q / L::FBW, k = q - p * L::FBW;
    const c32 zc = Z[p * BW + k], zm = Z[p * BW + ((BW - k) & (BW - 1))];
    Y[(2 * p) * L::FBW + k] = cmake(0.5f * (zc.x + zm.x), 0.5f * (zc.y - zm.y));
    Y[(2 * p + 1) * L::FBW + k] = cmake(0.5f * (zc.y + zm.y), -0.5f * (zc.x - zm.x));
  }
  __syncthreads();
  lg::lds_fft<false>(Y, BH, L::FBW, 1, L::FBW, P.tw, L::NMAX / BH);
  // packed store: column c >= 1 as is, column 0 W[r] = X[r][0] + i X[r][BW/2]
  const size_t base = (size_t)b * L::BSTR;
  for (int q = tid; q < L::BSTR; q += lg::NT) {
    const int r = q / L::NCOL, c = q - r * L::NCOL;
    c32 v;
    if (c == 0) {
      const c32 a = Y[r * L::FBW], h = Y[r * L::FBW + L::NCOL];
      v = cmake(a.x - h.y, a.y + h.x);
    } else {
      v = Y[r * L::FBW + c];
    }
    spec[base + q] = v;
  }
}

// Temporal filter of every bin of block b into Y (BH x FBW, unpacked), the
// DC and Nyquist columns replaced by their Hermitian parts as the real
// inverse keeps them (src/cpu/fft.cpp:88-119).
template <int BW, int BH, int BM>
__device__ __forceinline__ void filter_lds(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                           c32* Y, float4* kstate, bool kfirst) {
  using L = lg::LShape<BW, BH>;
  const int tid = threadIdx.x;
  const size_t base = (size_t)b * L::BSTR;
  const float dc = P.dg ? Wn.s[Wn.cur][base].x : 0.0f;  // block DC X[0][0] = Re W[0]
  float4* kb = kstate ? kstate + (size_t)b * BH * L::FBW : nullptr;
  for (int q = tid; q < BH * L::FBW; q += lg::NT) {
    const int r = q / L::FBW, c = q - r * L::FBW;
    c32 v[4];
    for (int t = 0; t < 4; t++) {
      const c32* s = Wn.s[t] + base;
      if (c != 0 && c != L::NCOL) {
        v[t] = s[r * L::NCOL + c];
      } else {
        const c32 wa = s[r * L::NCOL], wb = s[((BH - r) & (BH - 1)) * L::NCOL];
        v[t] = c == 0 ? cmake(0.5f * (wa.x + wb.x), 0.5f * (wa.y - wb.y))
                      : cmake(0.5f * (wa.y + wb.y), -0.5f * (wa.x - wb.x));
      }
    }
    const BinConst k = bin_const<BW, 2, 2, 2>(P, r, c, dc);
    Y[q] = filter_bin<BM, 2, 2>(P, v, k.d, k.sX, k.sY, k.lut, k.kcn, kb ? kb + q : nullptr,
                                kfirst);
  }
  __syncthreads();
  // Hermitian parts of columns 0 and BW/2: H[r] = (F[r] + conj F[-r]) / 2
  constexpr int HN = (2 * BH + lg::NT - 1) / lg::NT;  // entries per thread
  c32 h[HN];
  int hr[HN];
#pragma unroll
  for (int i = 0; i < HN; i++) {
    const int q = tid + i * lg::NT;
    hr[i] = -1;
    if (q < 2 * BH) {
      const int c = q < BH ? 0 : L::NCOL, r = q < BH ? q : q - BH;
      const c32 a = Y[r * L::FBW + c], m = Y[((BH - r) & (BH - 1)) * L::FBW + c];
      h[i] = cmake(0.5f * (a.x + m.x), 0.5f * (a.y - m.y));
      hr[i] = r * L::FBW + c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < HN; i++)
    if (hr[i] >= 0) Y[hr[i]] = h[i];
  __syncthreads();
}

template <int BW, int BH>
__global__ __launch_bounds__(lg::NT) void k_filter_inverse_lds(PlaneArgs P, WindowArgs Wn,
                                                               float4* kstate, int kfirst,
                                                               float* blocks) {
  using L = lg::LShape<BW, BH>;
  __shared__ c32 Z[L::ZN];
  __shared__ c32 Y[L::YN];
  const int tid = threadIdx.x, b = blockIdx.x;
  switch (P.bm) {
  case 4: filter_lds<BW, BH, 4>(P, Wn, b, Y, kstate, kfirst != 0); break;
  case 3: filter_lds<BW, BH, 3>(P, Wn, b, Y, kstate, kfirst != 0); break;
  case 2: filter_lds<BW, BH, 2>(P, Wn, b, Y, kstate, kfirst != 0); break;
  case 1: filter_lds<BW, BH, 1>(P, Wn, b, Y, kstate, kfirst != 0); break;
  case 0: filter_lds<BW, BH, 0>(P, Wn, b, Y, kstate, kfirst != 0); break;
  default: filter_lds<BW, BH, -1>(P, Wn, b, Y, kstate, kfirst != 0);
  }
  // vertical inverse of every column, then two real rows per complex row:
  // Z[p][k] = Y[2p][k] + i Y[2p+1][k], Y[r][BW-k] = conj Y[r][k]
  lg::lds_fft<true>(Y, BH, L::FBW, 1, L::FBW, P.tw, L::NMAX / BH);
  for (int q = tid; q < L::ZN; q += lg::NT) {
    const int p = q / BW, k = q - p * BW;
    const bool up = k > L::NCOL;
    const int kk = up ? BW - k : k;
    c32 a = Y[(2 * p) * L::FBW + kk], e = Y[(2 * p + 1) * L::FBW + kk];
    if (up) {
      a.y = -a.y;
      e.y = -e.y;
    }
    Z[q] = cmake(a.x - e.y, a.y + e.x);
  }
  __syncthreads();
  lg::lds_fft<true>(Z, BW, L::NROW, BW, 1, P.tw, L::NMAX / BW);
  int x0, y0, par;
  block_origin(P, b, x0, y0, par);
  const float* win = par ? P.winSynB : P.winSynA;
  const float sc = 1.0f / (float)(BW * BH);
  float* dst = blocks + (size_t)b * BW * BH;
  for (int q = tid; q < BW * BH; q += lg::NT) {
    const int ly = q / BW, lx = q - ly * BW;
    const c32 z = Z[(ly >> 1) * BW + lx];
    dst[q] = ((ly & 1 ? z.y : z.x) * sc) * win[q];
  }
}

}  // namespace fk
