// The block DFTs as dense matrix products on the matrix cores (gfx950 MFMA
// v_mfma_f32_32x32x16_{f16,bf16}: 16-bit operands, FP32 accumulation), the
// "DFT-as-GEMM" form of the analysis and of the sweep's inverse transform.
// It replaces the register FFTs (fk_fft.h) for 32x32 blocks; the reference's
// radix-2 work-group FFTs are src/cl/kernels.h:84-129 (fft_batch), used by
// fft2d_fwd :141-176 and fft2d_inv :368-404.
//
// Precision. Every operand is split into a high and a low 16-bit part
// (hi = round(x), lo = round(x - hi)) and each product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation; the dropped lo*lo term is
// about 2^-22 (FP16) or 2^-16 (BF16) of the product. The DFT matrices are
// evaluated in double on the host and split the same way.
//
// MFMA 32x32x16 lane layout (wave64; lane l, h = l / 32, i = l % 32),
// checked on gfx950 with small-integer operands:
//   A (32x16): lane l holds A[i][8h + j], j < 8
//   B (16x32): lane l holds B[8h + j][i], j < 8
//   D (32x32): lane l, register v holds D[8(v/4) + 4h + v%4][i]
// D keeps the row index in registers and the column index in lanes. A
// product that sums over D's rows can take D directly as its A operand (row
// index of A = D's column): chunk c of the sum covers D's registers
// 8c..8c+7, i.e. rows rho(c, h, j) = 16c + 8(j/4) + 4h + j%4; the constant
// operand is permuted the same way. So the stages of a 2-D DFT meet in
// registers, with no LDS transpose.
//
// Forward (one 32x32 real block x[r][k] per wave):
//   Z = x M          M[k][2p] = cos(2 pi kp/32), M[k][2p+1] = -sin(2 pi kp/32),
//                    M[k][1] = (-1)^k: the real row DFT, column 2p / 2p+1 the
//                    real / imaginary part of bin p, column 1 bin 16
//   P^T = Z^T C^T, Q^T = Z^T S^T     C[v][r] = cos(2 pi vr/32), S = sin
//   X[v][p] = P[v][2p] + Q[v][2p+1] + i (P[v][2p+1] - Q[v][2p])
// which for p = 0 is W[v] = X[v][0] + i X[v][16], the packed DC/Nyquist
// column of the stored spectrum (fk_kernels.h). Lane (v = l%32, h) ends with
// the packed columns p = 4t + 2h and 4t + 2h + 1 (t < 4) of row v.
// Inverse (spectrum F[v][p] of a filtered block, packed, in LDS):
//   Pi^T = F^T C^T, Qi^T = F^T S^T   (vertical inverse, e^{+i})
//   Y[r][p] = Pi[r][2p] - Qi[r][2p+1] + i (Pi[r][2p+1] + Qi[r][2p])
//   x^T = N^T Y^T    N[0][k] = 1, N[1][k] = (-1)^k, N[2p][k] = 2 cos(2 pi pk/32),
//                    N[2p+1][k] = -2 sin(2 pi pk/32): the real row inverse
// Lane (r = l%32, h) ends with samples k = 8t + 4h .. 8t + 4h + 3 (t < 4) of
// row r, unnormalised (x bw*bh), as the register inverse.
//
// Constant operands: 24 slots of 64 lanes x 8 16-bit values (uint4 per lane),
// built by fft3d::dft_operands (fft3d_params.h):
//   0-3   M, chunk c = s/2, hi (s even) / lo                    (forward B)
//   4-7   C permuted by rho, 8-11 S permuted by rho              (forward B)
//   12-15 C, 16-19 S, natural row order                          (inverse B)
//   20-23 N permuted by rho                                      (inverse A)
#pragma once

#include "fk_kernels.h"

namespace fk {
namespace mm {

enum { F16 = 1, BF16 = 2 };  // operand precision of the MFMA form
constexpr int kSlots = 24;    // constant operand slots per (block size, precision)

#if defined(FFT3D_CPU_EMU)
struct V8 {
  uint16_t h[8];
};
struct Acc {
  float v[16];
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
};
inline Acc acc_zero() { return Acc{}; }
#else
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float Acc __attribute__((ext_vector_type(16)));
struct V8 {
  uint4 u;
};
__device__ __forceinline__ Acc acc_zero() {
  Acc a;
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = 0.0f;
  return a;
}
#endif

// Round to the operand type (nearest even) as 16 bits, and back.
template <int PR>
__device__ __forceinline__ uint16_t rnd16(float x) {
#if defined(FFT3D_CPU_EMU)
  if constexpr (PR == F16) {
    return emu_f2h(x);
  } else {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
#else
  if constexpr (PR == F16) return __builtin_bit_cast(uint16_t, (_Float16)x);
  else return __builtin_bit_cast(uint16_t, (__bf16)x);
#endif
}
template <int PR>
__device__ __forceinline__ float val16(uint16_t b) {
#if defined(FFT3D_CPU_EMU)
  if constexpr (PR == F16) {
    return emu_h2f(b);
  } else {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  }
#else
  if constexpr (PR == F16) return (float)__builtin_bit_cast(_Float16, b);
  else return (float)__builtin_bit_cast(__bf16, b);
#endif
}

// x[0..7] -> high and low operand parts.
template <int PR>
__device__ __forceinline__ void split8(const float* x, V8& hi, V8& lo) {
  uint16_t a[8], b[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    a[j] = rnd16<PR>(x[j]);
    b[j] = rnd16<PR>(x[j] - val16<PR>(a[j]));
  }
#if defined(FFT3D_CPU_EMU)
#pragma unroll
  for (int j = 0; j < 8; j++) {
    hi.h[j] = a[j];
    lo.h[j] = b[j];
  }
#else
  hi.u = make_uint4(a[0] | ((uint32_t)a[1] << 16), a[2] | ((uint32_t)a[3] << 16),
                    a[4] | ((uint32_t)a[5] << 16), a[6] | ((uint32_t)a[7] << 16));
  lo.u = make_uint4(b[0] | ((uint32_t)b[1] << 16), b[2] | ((uint32_t)b[3] << 16),
                    b[4] | ((uint32_t)b[5] << 16), b[6] | ((uint32_t)b[7] << 16));
#endif
}

// A constant operand (slot s) of this lane.
__device__ __forceinline__ V8 cop(const uint4* tab, int s, int lane) {
  const uint4 u = tab[s * 64 + lane];
#if defined(FFT3D_CPU_EMU)
  V8 v;
  std::memcpy(v.h, &u, 16);
  return v;
#else
  return V8{u};
#endif
}

#if defined(FFT3D_CPU_EMU)
// One MFMA of a wave: every lane posts its operands, then forms its own
// accumulator registers from all of them (FP32 sums, k in order).
inline float emu_mm_a_[16][64][8], emu_mm_b_[16][64][8];
template <int PR>
inline void emu_mfma(Acc& c, const V8& a, const V8& b) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  for (int j = 0; j < 8; j++) {
    emu_mm_a_[w][l][j] = val16<PR>(a.h[j]);
    emu_mm_b_[w][l][j] = val16<PR>(b.h[j]);
  }
  emu_wave_sync();
  const int h = l / 32, n = l % 32;
  for (int v = 0; v < 16; v++) {
    const int i = 8 * (v / 4) + 4 * h + v % 4;
    float s = c[v];
    for (int k = 0; k < 16; k++)
      s += emu_mm_a_[w][i + 32 * (k / 8)][k % 8] * emu_mm_b_[w][n + 32 * (k / 8)][k % 8];
    c[v] = s;
  }
  emu_wave_sync();
}
#endif

template <int PR>
__device__ __forceinline__ void mma(Acc& c, const V8& a, const V8& b) {
#if defined(FFT3D_CPU_EMU)
  emu_mfma<PR>(c, a, b);
#else
  if constexpr (PR == F16)
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a.u),
                                                __builtin_bit_cast(f16x8, b.u), c, 0, 0, 0);
  else
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a.u),
                                                 __builtin_bit_cast(bf16x8, b.u), c, 0, 0, 0);
#endif
}
// c += (ah + al)(bh + bl) without the al*bl term, small products first.
template <int PR>
__device__ __forceinline__ void mma3(Acc& c, const V8& ah, const V8& al, const V8& bh,
                                     const V8& bl) {
  mma<PR>(c, al, bh);
  mma<PR>(c, ah, bl);
  mma<PR>(c, ah, bh);
}

// Samples x0..x0+7 of plane row y (already mirrored) as floats minus
// chromaNorm: one vector load when the span lies inside the plane at an
// aligned address, else sample by sample through the mirror.
__device__ __forceinline__ void load8(const PlaneArgs& P, const void* plane, int x0, int y,
                                      float* z) {
  const int bps = P.pix == 8 ? 1 : (P.pix == 16 ? 2 : 4);
  const char* row = static_cast<const char*>(plane) + ((size_t)y * P.W + x0) * bps;
  const bool vec = x0 >= 0 && x0 + 8 <= P.W && (reinterpret_cast<uintptr_t>(row) & 15) == 0;
  if (vec && P.pix == 8) {
    const uint2 v = *reinterpret_cast<const uint2*>(row);
#pragma unroll
    for (int k = 0; k < 8; k++)
      z[k] = (float)(((k < 4 ? v.x : v.y) >> (8 * (k & 3))) & 0xffu) * (1.0f / 255.0f);
  } else if (vec && P.pix == 16) {
    const uint4 v = *reinterpret_cast<const uint4*>(row);
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 8; k++) z[k] = (float)((w[k >> 1] >> (16 * (k & 1))) & 0xffffu) * P.inNorm;
  } else if (vec) {
    const float4 a = reinterpret_cast<const float4*>(row)[0];
    const float4 b = reinterpret_cast<const float4*>(row)[1];
    z[0] = a.x; z[1] = a.y; z[2] = a.z; z[3] = a.w;
    z[4] = b.x; z[5] = b.y; z[6] = b.z; z[7] = b.w;
  } else {
    const size_t base = (size_t)y * P.W;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const size_t i = base + mirror(x0 + k, P.W);
      float v;
      if (P.pix == 8) v = (float)static_cast<const uint8_t*>(plane)[i] * (1.0f / 255.0f);
      else if (P.pix == 16) v = (float)static_cast<const uint16_t*>(plane)[i] * P.inNorm;
      else v = static_cast<const float*>(plane)[i];
      z[k] = v;
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) z[k] -= P.chromaNorm;
}

// The lane's 16 samples of 32x32 block b (row l%32, columns 8h..8h+7 and
// 16+8h..16+8h+7) times the lane's analysis window factors wf.
__device__ __forceinline__ void load_block32(const PlaneArgs& P, const void* plane, int b,
                                             int lane, const float* wf, float* xs) {
  const int h = lane >> 5, li = lane & 31;
  int x0, y0, par;
  block_origin(P, b, x0, y0, par);
  const int y = mirror(y0 + li, P.H);
  load8(P, plane, x0 + 8 * h, y, xs);
  load8(P, plane, x0 + 16 + 8 * h, y, xs + 8);
#pragma unroll
  for (int k = 0; k < 16; k++) xs[k] *= wf[k];
}

// The lane's window factors: the same for every block (mode 1: the separable
// factors' product anaX[k] * anaY[r], as the register form applies them).
__device__ __forceinline__ void window32(const PlaneArgs& P, int lane, float* wf) {
  const int h = lane >> 5, li = lane & 31;
  if (P.mode == 1) {
    const float fy = P.anaY[li];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int j = 0; j < 8; j++) wf[8 * c + j] = P.anaX[16 * c + 8 * h + j] * fy;
  } else {
    const float* wr = P.winAna + li * 32;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int j = 0; j < 8; j++) wf[8 * c + j] = wr[16 * c + 8 * h + j];
  }
}

// 2-D R2C DFT of the windowed 32x32 block b whose samples the wave holds
// (xs, as load_block32 leaves them) by all 64 lanes, stored packed into spec
// as k_analysis does. T: the forward constant operands (slots 0-11).
template <int PR, bool H>
__device__ __forceinline__ void forward_block_mfma(const float* xs, const uint4* T, int b,
                                                   c32* spec, int lane) {
  constexpr int NCOL = 16, BSTR = 32 * 16;
  const int h = lane >> 5, li = lane & 31;
  V8 ah[2], al[2];
  split8<PR>(xs, ah[0], al[0]);
  split8<PR>(xs + 8, ah[1], al[1]);
  Acc z = acc_zero();
#pragma unroll
  for (int c = 0; c < 2; c++)
    mma3<PR>(z, ah[c], al[c], cop(T, 2 * c, lane), cop(T, 2 * c + 1, lane));
  float zf[16];
#pragma unroll
  for (int v = 0; v < 16; v++) zf[v] = z[v];
  split8<PR>(zf, ah[0], al[0]);
  split8<PR>(zf + 8, ah[1], al[1]);
  Acc p = acc_zero(), q = acc_zero();
#pragma unroll
  for (int c = 0; c < 2; c++) {
    mma3<PR>(p, ah[c], al[c], cop(T, 4 + 2 * c, lane), cop(T, 5 + 2 * c, lane));
    mma3<PR>(q, ah[c], al[c], cop(T, 8 + 2 * c, lane), cop(T, 9 + 2 * c, lane));
  }
  // lane (row v = li, h): registers 4t, 4t+1 are (re, im) of column 4t + 2h,
  // 4t+2, 4t+3 of column 4t + 2h + 1
  c32* row = spec + (size_t)b * BSTR + li * NCOL;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int u = 4 * t;
    const c32 a = cmake(p[u] + q[u + 1], p[u + 1] - q[u]);
    const c32 e = cmake(p[u + 2] + q[u + 3], p[u + 3] - q[u + 2]);
    const int col = 4 * t + 2 * h;
    if constexpr (H) {
      *reinterpret_cast<uint2*>(reinterpret_cast<uint32_t*>(spec) + (size_t)b * BSTR +
                                li * NCOL + col) = make_uint2(pack_h2(a.x, a.y), pack_h2(e.x, e.y));
    } else {
      *reinterpret_cast<float4*>(row + col) = make_float4(a.x, a.y, e.x, e.y);
    }
  }
}

// The inverse constant operands (slots 12-23) of this lane, loaded once for
// all the blocks a wave transforms in a row.
struct InvOps {
  V8 c[4], s[4], n[4];
};
__device__ __forceinline__ void load_inv_ops(const uint4* T, int lane, InvOps& o) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    o.c[i] = cop(T, 12 + i, lane);
    o.s[i] = cop(T, 16 + i, lane);
    o.n[i] = cop(T, 20 + i, lane);
  }
}

// Inverse 2-D C2R of the filtered 32x32 block in its LDS area Yb (packed
// columns, row pitch `pitch` complex), by the calling wave; the windowed rows
// (x synX[k] synY[r], factors that include 1/(bw*bh)) are written back over
// Yb as rows of `op` floats. Contains the wave barrier between the reads of
// Yb and the writes.
template <int PR>
__device__ __forceinline__ void inverse_block_mfma(const InvOps& O, c32* Yb, int pitch, int op,
                                                   const float* synX, const float* synY,
                                                   int lane) {
  const int h = lane >> 5, li = lane & 31;
  // A = F^T: lane li is real column li (complex column li/2, component li&1)
  const float* Yf = reinterpret_cast<const float*>(Yb);
  float fa[16];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int j = 0; j < 8; j++) fa[8 * c + j] = Yf[2 * pitch * (16 * c + 8 * h + j) + li];
  V8 ah[2], al[2];
  split8<PR>(fa, ah[0], al[0]);
  split8<PR>(fa + 8, ah[1], al[1]);
  Acc p = acc_zero(), q = acc_zero();
#pragma unroll
  for (int c = 0; c < 2; c++) {
    mma3<PR>(p, ah[c], al[c], O.c[2 * c], O.c[2 * c + 1]);
    mma3<PR>(q, ah[c], al[c], O.s[2 * c], O.s[2 * c + 1]);
  }
  // lane (row r = li, h): registers (2s, 2s+1) are (re, im) of a column
  float yp[16];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    yp[2 * s] = p[2 * s] - q[2 * s + 1];
    yp[2 * s + 1] = p[2 * s + 1] + q[2 * s];
  }
  V8 bh[2], bl[2];
  split8<PR>(yp, bh[0], bl[0]);
  split8<PR>(yp + 8, bh[1], bl[1]);
  Acc x = acc_zero();
#pragma unroll
  for (int c = 0; c < 2; c++) mma3<PR>(x, O.n[2 * c], O.n[2 * c + 1], bh[c], bl[c]);
  wave_sync();  // every lane has read Yb
  const float fy = synY[li];
  float* o = reinterpret_cast<float*>(Yb) + li * op;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int k = 8 * t + 4 * h;
    *reinterpret_cast<float4*>(o + k) =
        make_float4(x[4 * t] * (synX[k] * fy), x[4 * t + 1] * (synX[k + 1] * fy),
                    x[4 * t + 2] * (synX[k + 2] * fy), x[4 * t + 3] * (synX[k + 3] * fy));
  }
}

}  // namespace mm

// Analysis by MFMA DFT: 4 waves per workgroup. The plane's blocks form
// chunks of 4 consecutive blocks, one block per wave; workgroup wg of nwg
// takes a contiguous run of nch / nwg chunks (balanced, XCD-contiguous, see
// xcd_contiguous), so neighbouring blocks, which share pixel rows, are read
// together. A grid of one round of resident workgroups (each with a run of
// many chunks) leaves no ragged last round.
// The forward constant operands go to LDS once per workgroup; each wave
// loads the samples of its next block while it transforms the current one.
template <int PR, bool H>
__device__ __forceinline__ void analysis_mfma_body(const PlaneArgs& P, const void* plane,
                                                   c32* spec, int wg, int nwg) {
  __shared__ uint4 ctab[12 * 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < 12 * 64; i += 256) ctab[i] = P.dft[i];
  const int nch = (P.nb + 3) / 4;
  const long long j = xcd_contiguous(wg, nwg);
  const int c0 = (int)(j * nch / nwg), c1 = (int)((j + 1) * nch / nwg);
  float wf[16], xs[16], xn[16];
  mm::window32(P, lane, wf);
  int b = c0 * 4 + wv;
  if (c0 < c1 && b < P.nb) mm::load_block32(P, plane, b, lane, wf, xs);
  __syncthreads();  // the constant operands are in LDS
  for (int c = c0; c < c1 && b < P.nb; c++) {  // b is wave-uniform
    const int bn = b + 4;
    if (c + 1 < c1 && bn < P.nb) mm::load_block32(P, plane, bn, lane, wf, xn);
    mm::forward_block_mfma<PR, H>(xs, ctab, b, spec, lane);
#pragma unroll
    for (int k = 0; k < 16; k++) xs[k] = xn[k];
    b = bn;
  }
}
template <int PR, bool H>
__global__ __launch_bounds__(256) void k_analysis_mfma(PlaneArgs P, const void* plane, c32* spec) {
  analysis_mfma_body<PR, H>(P, plane, spec, blockIdx.x, gridDim.x);
}

}  // namespace fk
