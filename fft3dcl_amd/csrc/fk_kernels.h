// HIP kernels of the FFT3D denoise path (gfx950, wave64).
//
// Data layout in HBM
//   plane    : W x H samples, packed rows (u8, u16 or f32), read in place.
//   spectrum : per frame, per block b, bh rows x bw/2 packed columns of
//              complex float ([b][r][c], c fastest). Column c >= 1 holds the
//              2-D DFT bin (r, c); column 0 holds W[r] = X[r][0] + i X[r][bw/2],
//              the two real-signal columns packed into one complex column
//              (both are Hermitian in r, so the packing is exact). This is
//              bw*bh*4 bytes per block instead of the reference's
//              (bw/2+1)*bh*8 (src/cpu/reference.cpp:161-167).
//   kalman   : per bin (b, r, col<bw/2+1) float4 {last.re, last.im, covar, proc}.
//
// Work mapping. A 64-lane workgroup owns G = 64 / max(bh/2, bw/2) blocks. In
// a row pass lane j of block g transforms the packed row pair (2j, 2j+1) as
// one complex sequence; in a column pass lane c transforms column c (c = 0 is
// the packed DC/Nyquist column). Each 1-D transform runs inside one lane's
// registers (fk_fft.h); rows and columns meet through one padded LDS
// transpose per direction.
//
// Reference kernels this replaces (src/cl/kernels.h): fft2d_fwd :141-176,
// filter :185-345, kalman_init :348-363 (folded into the filter), fft2d_inv
// :368-404, synthesize :409-448.
#pragma once

#include "fk_fft.h"

namespace fk {

struct PlaneArgs {
  int W, H;            // plane size in samples
  int bw, bh, ow, oh;  // block size and overlap
  int mode;            // 0: two 1:1 grids, 1: overlapped grid
  int nx, ny;          // grid cells (src/params.h:131-139)
  int ncols, nrows;    // mode 1 block columns / rows
  int nb;              // blocks per frame
  int pix;             // 0 float samples, 8 u8, 16 u16
  float inNorm;        // u16 -> float factor 2^(16-bits)/65535
  float invNorm;       // float -> u16 factor 1/2^(16-bits)
  float maxPix;        // 2^bits - 1
  float chromaNorm;    // chroma centring (0 when off)
  int bm;              // temporal mode -1..4
  int dg, pat, shp;    // degrid, sigma pattern, sharpen on/off
  float sigmaX, sigmaY, betaX, beta;
  float thrX;  // wiener_threshold(sigmaX): the Wiener test is psd > thrX
  float ssmin, ssmax, kalmanCN, kr2;
  int khalf;  // Kalman state as FP16 pairs (the FP16 ring on integer samples)
  const float* winAna;  // bw*bh analysis window
  const float* winSynA; // bw*bh synthesis window, even block columns (mode 0: winAna)
  const float* winSynB; // bw*bh synthesis window, odd block columns
  const float* pattern; // (bw/2+1)*bh sigma pattern (pat)
  const float* sharpen; // (bw/2+1)*bh sharpen gains (shp)
  const c32* gs;        // (bw/2+1)*bh degrid sample (dg)
  // Separable factors of the mode-1 windows (window = fx[lx] * fy[ly]); the
  // synthesis factors already include the inverse transform's 1/(bw*bh).
  const float* anaX;   // bw
  const float* anaY;   // bh
  const float* synX;   // bw (half overlap: both block parities share it)
  const float* synY;   // bh
  // separable factors of the sharpen and degrid tables (mode 1): row r
  // {shY, gyRe, gyIm, strength}, column c {shX, gxRe, gxIm, 0}
  const float4* sepRow;  // bh
  const float4* sepCol;  // bw/2+1
  // constant operands of the MFMA DFT form (fk_mfma.h), null when not used
  const uint4* dft;
  // twiddles (cos, sin)(2 pi k / max(bw, bh)), k < max(bw, bh) / 2, for the
  // LDS FFTs of blocks with a side above 64 (fk_large.h)
  const float2* tw;
};

// The temporal window handed to the filter: frames in temporal order.
struct WindowArgs {
  const c32* s[4];
  int nwin;  // bt for bm >= 2, else 1
  int cur;   // index of the current frame in s[]
};

// Element i of a stored spectrum: complex FP32 (8 bytes), or with H the two
// halves rounded to FP16 (4 bytes, round to nearest even) in the same layout.
template <bool H>
__device__ __forceinline__ c32 spec_load(const c32* s, size_t i) {
  if constexpr (H) {
    const float2 v = unpack_h2(reinterpret_cast<const uint32_t*>(s)[i]);
    return cmake(v.x, v.y);
  } else {
    return s[i];
  }
}
template <bool H>
__device__ __forceinline__ c32 spec_round(c32 v) {
  if constexpr (H) {
    const float2 u = unpack_h2(pack_h2(v.x, v.y));
    return cmake(u.x, u.y);
  } else {
    return v;
  }
}
template <bool H>
__device__ __forceinline__ void spec_store(c32* s, size_t i, c32 v) {
  if constexpr (H) reinterpret_cast<uint32_t*>(s)[i] = pack_h2(v.x, v.y);
  else s[i] = v;
}

template <int BW, int BH>
struct Shape {
  static constexpr int NROW = BH / 2;  // packed row pairs per block
  static constexpr int NCOL = BW / 2;  // stored columns per block
  static constexpr int LPB = NROW > NCOL ? NROW : NCOL;
  static constexpr int G = 64 / LPB;   // blocks per workgroup
  static constexpr int FBW = BW / 2 + 1;
  static constexpr int BSTR = BH * NCOL;  // complex values stored per block
  static constexpr int ZP = BW + 1;       // LDS pitch, row-pass buffer
  // Per-block LDS area of the filter/inverse stage: BH rows of PITCH complex,
  // row-major. Slots 0..NCOL-1 hold the filtered columns; while filtering, the
  // packed DC/Nyquist column W[r] of window frame t is staged in slot st(t)
  // (frame 0 in slot 0, which pass 2 overwrites row by row with the result,
  // frames 1-3 in the pad slots NCOL..NCOL+2, which also break the LDS bank
  // aliasing of the row pass).
  static constexpr int PITCH = NCOL + 3;
  static constexpr int YS = PITCH * BH;
  static constexpr int LDS_C = (G * NROW * ZP > G * YS) ? G * NROW * ZP : G * YS;
  // k_analysis: the transpose in two rounds
  static constexpr int SPLIT_A = NROW >= 2 ? 2 : 1;
  static constexpr int LDS_A = G * (NROW / SPLIT_A) * ZP;
  __device__ __forceinline__ static int yidx(int c, int r) { return r * PITCH + c; }
  __device__ __forceinline__ static int st(int t, int r) {
    return r * PITCH + (t == 0 ? 0 : NCOL + t - 1);
  }
};

// Direct3D MIRROR addressing (src/cpu/reference.cpp:10-17).
__device__ __forceinline__ int mirror(int x, int w) {
  if (x >= 0 && x < w) return x;
  // one reflection covers every block of a plane at least a block wide: no
  // integer division on the edge blocks
  if (x >= -w && x < 2 * w) return x < 0 ? -x - 1 : 2 * w - 1 - x;
  int p = x % (2 * w);
  if (p < 0) p += 2 * w;
  if (p >= w) p = 2 * w - 1 - p;
  return p;
}

// Block origin and synthesis parity (src/cpu/reference.cpp:23-46).
__device__ __forceinline__ void block_origin(const PlaneArgs& P, int b, int& x0, int& y0,
                                             int& par) {
  if (P.mode == 1) {
    const int m = b / P.ncols, n = b - m * P.ncols;
    x0 = -P.ow + n * (P.bw - P.ow);
    y0 = -P.oh + m * (P.bh - P.oh);
    par = n & 1;
  } else {
    const int per = P.nx * P.ny;
    const int grid = b >= per ? 1 : 0;
    const int r = b - grid * per, j = r / P.nx, i = r - j * P.nx;
    x0 = i * P.bw - (grid ? P.bw / 2 : 0);
    y0 = j * P.bh - (grid ? P.bh / 2 : 0);
    par = 0;
  }
}

// Sample -> float (src/cpu/reference.cpp:82-105; src/cl/kernels.h:22-34):
// Samples x0..x0+BW-1 of plane row y (already mirrored) as floats, minus
// chromaNorm. Rows that lie inside the plane at a 16-byte aligned address are
// read with 16-byte loads; other rows go sample by sample through the mirror.
template <int BW>
__device__ __forceinline__ void load_row(const PlaneArgs& P, const void* plane, int x0, int y,
                                         float (&z)[BW]) {
  const int bps = P.pix == 8 ? 1 : (P.pix == 16 ? 2 : 4);
  const char* row = static_cast<const char*>(plane) + ((size_t)y * P.W + x0) * bps;
  const bool vec = (BW * bps) % 16 == 0 && x0 >= 0 && x0 + BW <= P.W &&
                   (reinterpret_cast<uintptr_t>(row) & 15) == 0;
  if (vec) {
    const uint4* q = reinterpret_cast<const uint4*>(row);
    if (P.pix == 8) {
#pragma unroll
      for (int i = 0; i < BW / 16; i++) {
        const uint4 v = q[i];
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; k++)
          z[16 * i + k] = (float)((w[k >> 2] >> (8 * (k & 3))) & 0xffu) * (1.0f / 255.0f);
      }
    } else if (P.pix == 16) {
#pragma unroll
      for (int i = 0; i < BW / 8; i++) {
        const uint4 v = q[i];
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; k++)
          z[8 * i + k] = (float)((w[k >> 1] >> (16 * (k & 1))) & 0xffffu) * P.inNorm;
      }
    } else {
#pragma unroll
      for (int i = 0; i < BW / 4; i++) {
        const float4 v = reinterpret_cast<const float4*>(row)[i];
        z[4 * i] = v.x;
        z[4 * i + 1] = v.y;
        z[4 * i + 2] = v.z;
        z[4 * i + 3] = v.w;
      }
    }
  } else {
    const size_t base = (size_t)y * P.W;
#pragma unroll
    for (int k = 0; k < BW; k++) {
      const size_t i = base + mirror(x0 + k, P.W);
      float v;
      if (P.pix == 8) v = (float)static_cast<const uint8_t*>(plane)[i] * (1.0f / 255.0f);
      else if (P.pix == 16) v = (float)static_cast<const uint16_t*>(plane)[i] * P.inNorm;
      else v = static_cast<const float*>(plane)[i];
      z[k] = v;
    }
  }
#pragma unroll
  for (int k = 0; k < BW; k++) z[k] -= P.chromaNorm;
}

// --------------------------------------------------------------- analysis

// Windowed 2-D R2C FFT of block b. Lanes j < NCOL end with packed column j
// of the spectrum and call emit(yr, yi) with it (r < BH). All lanes of the wave
// call forward_block; its LDS area (NROW/SPLIT rows of ZP complex per block)
// belongs to the wave. With SPLIT > 1 the row-to-column transpose goes through
// it in SPLIT rounds, which divides the LDS a workgroup needs.
template <int BW, int BH, int SPLIT = 1, class Emit>
__device__ __forceinline__ void forward_block(const PlaneArgs& P, const void* plane, int b,
                                              bool live, int g, int j, c32* lds, Emit&& emit) {
  using S = Shape<BW, BH>;
  constexpr int HR = S::NROW / SPLIT;  // row pairs per round through LDS
  float zr[BW], zi[BW];
  if (live && j < S::NROW) {
    int x0, y0, par;
    block_origin(P, b, x0, y0, par);
    const int r0 = 2 * j;
    load_row<BW>(P, plane, x0, mirror(y0 + r0, P.H), zr);
    load_row<BW>(P, plane, x0, mirror(y0 + r0 + 1, P.H), zi);
    if (P.mode == 1) {
      // separable window: a uniform column factor times a per-lane row factor
      const float fy0 = P.anaY[r0], fy1 = P.anaY[r0 + 1];
#pragma unroll
      for (int k = 0; k < BW; k++) {
        const float fx = P.anaX[k];
        zr[k] *= fx * fy0;
        zi[k] *= fx * fy1;
      }
    } else {
      const float* w0 = P.winAna + r0 * BW;
#pragma unroll
      for (int k = 0; k < BW; k++) {
        zr[k] *= w0[k];
        zi[k] *= w0[BW + k];
      }
    }
    fft_reg<BW, false>(zr, zi);
  }
  // row pairs [h*HR, (h+1)*HR) go through the block's LDS rows in round h
  float yr[BH], yi[BH];
  const c32* Zg = lds + g * HR * S::ZP;
#pragma unroll
  for (int h = 0; h < SPLIT; h++) {
    if (live && j < S::NROW && j / HR == h) {
      c32* Z = lds + (g * HR + j - h * HR) * S::ZP;
#pragma unroll
      for (int k = 0; k < BW; k++) Z[k] = cmake(zr[k], zi[k]);
    }
    wave_sync();
    if (live && j < S::NCOL) {
      const int c = j;
#pragma unroll
      for (int pp = 0; pp < HR; pp++) {
        const int p = h * HR + pp;
        const c32* Z = Zg + pp * S::ZP;
        if (c == 0) {
          // rows 2p, 2p+1 are real at DC and Nyquist: X = (Re Z, Im Z)
          const c32 a = Z[0], hh = Z[BW / 2];
          yr[2 * p] = a.x;
          yi[2 * p] = hh.x;
          yr[2 * p + 1] = a.y;
          yi[2 * p + 1] = hh.y;
        } else {
          // X_2p[c] = (Z[c] + conj Z[-c]) / 2, X_2p+1[c] = (Z[c] - conj Z[-c]) / 2i
          const c32 zc = Z[c], zm = Z[BW - c];
          yr[2 * p] = 0.5f * (zc.x + zm.x);
          yi[2 * p] = 0.5f * (zc.y - zm.y);
          yr[2 * p + 1] = 0.5f * (zc.y + zm.y);
          yi[2 * p + 1] = -0.5f * (zc.x - zm.x);
        }
      }
    }
    if (h + 1 < SPLIT) wave_sync();  // the next round overwrites the rows
  }
  if (live && j < S::NCOL) {
    fft_reg<BH, false>(yr, yi);
    emit(yr, yi);
  }
}

// Workgroup i of a launch runs on XCD i % 8, and each XCD has its own L2.
// Mapping XCD x to a contiguous run of the n work items (a bijection on
// [0, n)) keeps neighbouring blocks, which read overlapping pixel rows, under
// one L2.
__device__ __forceinline__ int xcd_contiguous(int i, int n) {
  const int q = n / 8, r = n % 8, x = i % 8;
  return x * q + min(x, r) + i / 8;
}

// Analysis of the G blocks of work item wg of nwg (one 64-lane workgroup).
template <int BW, int BH, bool H>
__device__ __forceinline__ void analysis_body(const PlaneArgs& P, const void* plane, c32* spec,
                                              int wg, int nwg) {
  using S = Shape<BW, BH>;
  __shared__ c32 lds[S::LDS_A];
  const int lane = threadIdx.x, g = lane / S::LPB, j = lane - g * S::LPB;
  const int b = xcd_contiguous(wg, nwg) * S::G + g;
  const bool live = b < P.nb;
  const size_t d = (size_t)b * S::BSTR + j;
  forward_block<BW, BH, S::SPLIT_A>(P, plane, b, live, g, j, lds,
                                    [&](const float(&yr)[BH], const float(&yi)[BH]) {
#pragma unroll
                                      for (int r = 0; r < BH; r++)
                                        spec_store<H>(spec, d + r * S::NCOL, cmake(yr[r], yi[r]));
                                    });
}
template <int BW, int BH, bool H>
__global__ __launch_bounds__(64) void k_analysis(PlaneArgs P, const void* plane, c32* spec) {
  analysis_body<BW, BH, H>(P, plane, spec, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- filter

// Wiener gain (src/cpu/reference.cpp:201-209). The reference's test
// 1/sigmaX > 1/psd is evaluated as psd > thr with thr = wiener_threshold(sX)
// (computed once on the host, PlaneArgs::thrX, unless a pattern varies sX)
// (no reciprocal per call); the gain is continuous at the threshold.
__host__ __device__ inline float wiener_threshold(float sX) {
  if (sX > 0.0f) return sX;
  // 1/(+0) = +inf passes every psd; 1/(-0) and 1/negative never pass.
  return (sX == 0.0f && copysignf(1.0f, sX) > 0.0f) ? -1.0f : __builtin_inff();
}
// With `one` a power of two (the mean's 1/2 or 1/4 folded in, sY and betaX
// scaled by it) the result is exactly `one` times the unscaled gain's. The
// 1e-37 goes in with y*y: it only decides psd > 0 for a zero bin, and a bin
// whose psd it could change takes the betaX branch either way.
__device__ __forceinline__ c32 wiener(c32 v, float thr, float sY, float betaX,
                                      float one = 1.0f) {
  const float psd = fmaf(v.x, v.x, fmaf(v.y, v.y, 1e-37f));
  const float mul = psd > thr ? one - sY * frcp(psd) : betaX;
  return cscale(v, mul);
}

// A feature flag: fixed at compile time (F = 0 or 1) or read from the plane
// (F = 2).
template <int F>
__device__ __forceinline__ bool flag(int v) {
  return F == 2 ? v != 0 : F != 0;
}

// The Kalman state of one bin, {last.re, last.im, covar, proc}: a float4, or
// with KH (P.khalf: the FP16 ring on integer samples) the same four values as
// two FP16 pairs rounded to nearest even, 8 bytes instead of 16.
template <bool KH>
__device__ __forceinline__ float4 kstate_load(const void* p) {
  if constexpr (KH) {
    const uint2 u = *static_cast<const uint2*>(p);
    const float2 a = unpack_h2(u.x), b = unpack_h2(u.y);
    return make_float4(a.x, a.y, b.x, b.y);
  } else {
    return *static_cast<const float4*>(p);
  }
}
template <bool KH>
__device__ __forceinline__ void kstate_store(void* p, float4 v) {
  if constexpr (KH) *static_cast<uint2*>(p) = make_uint2(pack_h2(v.x, v.y), pack_h2(v.z, v.w));
  else *static_cast<float4*>(p) = v;
}

// Temporal filter + sharpen of one bin (src/cpu/reference.cpp:246-409,
// src/cl/kernels.h:206-342). v[t] are the bin's values in temporal order,
// d the degrid correction, ks the Kalman state of the bin (bm == 0).
template <int BM, int DG, int SHP, bool KH = false>
__device__ __forceinline__ c32 filter_bin(const PlaneArgs& P, const c32 (&v)[4], c32 d,
                                          float sX, float sY, float lut, float kcn, void* ks,
                                          bool kfirst) {
  const bool dg = flag<DG>(P.dg), shp = flag<SHP>(P.shp);
  c32 f;
  if constexpr (BM == 2) {
    const c32 pv = v[0], cu = v[1];
    c32 a = csub(cu, pv), bb = cadd(cu, pv);
    if (dg) bb = csub(bb, cscale(d, 2.0f));
    // the mean (1/2) folded into the gains
    const float sY2 = 0.5f * sY, bX2 = 0.5f * P.betaX;
    a = wiener(a, sX, sY2, bX2, 0.5f);
    bb = wiener(bb, sX, sY2, bX2, 0.5f);
    f = cadd(a, bb);
    if (dg) f = cadd(f, d);
  } else if constexpr (BM == 3) {
    const c32 pv = v[0], cu = v[1], nx = v[2];
    const float K = 0.8660254f;
    const c32 t1 = cadd(pv, nx);
    const c32 y1 = csub(cu, cscale(t1, 0.5f));
    const c32 dn = csub(nx, pv);
    const c32 y2 = cmake(K * dn.y, -K * dn.x);
    c32 d0 = cadd(y1, y2), d1 = cadd(cu, t1), d2 = csub(y1, y2);
    if (dg) d1 = csub(d1, cscale(d, 3.0f));
    d0 = wiener(d0, sX, sY, P.betaX);
    d1 = wiener(d1, sX, sY, P.betaX);
    d2 = wiener(d2, sX, sY, P.betaX);
    f = cscale(cadd(cadd(d0, d1), d2), 1.0f / 3.0f);
    if (dg) f = cadd(f, d);
  } else if constexpr (BM == 4) {
    const c32 p2 = v[0], p1 = v[1], cu = v[2], nx = v[3];
    // 4-point temporal DFT as two radix-2 stages
    const c32 s02 = cadd(p2, cu), d02 = csub(cu, p2);
    const c32 s13 = cadd(p1, nx), d13 = csub(nx, p1);
    c32 fc = cadd(s02, s13), fp2 = csub(s02, s13);
    c32 fp1 = cmake(d02.x - d13.y, d02.y + d13.x);  // (c - p2) + i (n - p1)
    c32 fn = cmake(d02.x + d13.y, d02.y - d13.x);   // (c - p2) - i (n - p1)
    if (dg) fp1 = csub(fp1, cscale(d, 4.0f));
    // the mean (1/4) folded into the gains
    const float sY4 = 0.25f * sY, bX4 = 0.25f * P.betaX;
    fp2 = wiener(fp2, sX, sY4, bX4, 0.25f);
    fp1 = wiener(fp1, sX, sY4, bX4, 0.25f);
    fc = wiener(fc, sX, sY4, bX4, 0.25f);
    fn = wiener(fn, sX, sY4, bX4, 0.25f);
    f = cadd(cadd(cadd(fp2, fp1), fc), fn);
    if (dg) f = cadd(f, d);
  } else if constexpr (BM == 1) {
    c32 x = v[0];
    if (dg) x = csub(x, d);
    x = wiener(x, sX, sY, P.betaX);
    f = dg ? cadd(x, d) : x;
  } else if constexpr (BM == 0) {
    // Kalman (src/cpu/reference.cpp:334-383); first frame initialises state.
    const c32 cu = v[0];
    float4 st = kfirst ? make_float4(cu.x, cu.y, kcn, kcn) : kstate_load<KH>(ks);
    const float dr = cu.x - st.x, di = cu.y - st.y;
    const float mask = fmaxf(dr * dr, di * di);
    const float s = st.z + st.w;
    const float gain = s / (s + kcn);
    float np = gain * gain * kcn, nc = (1.0f - gain) * s;
    float fr = gain * cu.x + (1.0f - gain) * st.x;
    float fi = gain * cu.y + (1.0f - gain) * st.y;
    if (mask > P.kr2 * kcn) {
      fr = cu.x;
      fi = cu.y;
      nc = kcn;
      np = kcn;
    }
    kstate_store<KH>(ks, make_float4(fr, fi, nc, np));
    f = cmake(fr, fi);
  } else {
    f = v[0];  // bm == -1: the current frame passes through
  }
  if (shp) {
    c32 x = dg ? csub(f, d) : f;
    const float psd = fmaf(x.x, x.x, fmaf(x.y, x.y, 1e-37f));
    const float q = psd * P.ssmax * frcp((psd + P.ssmin) * (psd + P.ssmax));
    const float gsh = 1.0f + lut * fsqrt(q);
    x = cscale(x, gsh);
    f = dg ? cadd(x, d) : x;
  }
  return f;
}

// Per-bin constants of (row r, column col < bw/2+1): the Wiener and Kalman
// sigma terms, the sharpen gain and the degrid term d = gs * dc of the block.
struct BinConst {
  float sX, sY, lut, kcn;
  c32 d;
};
template <int PAT>
__device__ __forceinline__ void bin_sigma(const PlaneArgs& P, int i, BinConst& k) {
  if (flag<PAT>(P.pat)) {
    const float pv = P.pattern[i];
    k.sX = wiener_threshold(pv * P.beta);
    k.sY = pv;
    k.kcn = pv;
  } else {
    k.sX = P.thrX;
    k.sY = P.sigmaY;
    k.kcn = P.kalmanCN;
  }
}
// From the full sharpen and degrid tables in device memory.
template <int BW, int DG, int SHP, int PAT>
__device__ __forceinline__ BinConst bin_const(const PlaneArgs& P, int r, int col, float dc) {
  constexpr int FBW = BW / 2 + 1;
  BinConst k;
  const int i = r * FBW + col;
  bin_sigma<PAT>(P, i, k);
  k.lut = flag<SHP>(P.shp) ? P.sharpen[i] : 0.0f;
  k.d = flag<DG>(P.dg) ? cscale(P.gs[i], dc) : cmake(0.0f, 0.0f);
  return k;
}
// From the separable factors: rf = row factor {shY, gyRe, gyIm, strength},
// cf = column factor with its degrid part already scaled by the block DC.
template <int BW, int DG, int SHP, int PAT>
__device__ __forceinline__ BinConst bin_const_sep(const PlaneArgs& P, int r, int col, float4 rf,
                                                  float4 cf) {
  constexpr int FBW = BW / 2 + 1;
  BinConst k;
  bin_sigma<PAT>(P, r * FBW + col, k);
  k.lut = flag<SHP>(P.shp) ? rf.w - rf.x * cf.x : 0.0f;
  k.d = flag<DG>(P.dg) ? cmake(cf.y * rf.y - cf.z * rf.z, cf.y * rf.z + cf.z * rf.y)
                       : cmake(0.0f, 0.0f);
  return k;
}
// Column factor of column col scaled for a block of DC dc.
__device__ __forceinline__ float4 col_factor(const float4* tab, int bh, int col, float dc) {
  float4 cf = tab[bh + col];
  cf.y *= dc;
  cf.z *= dc;
  return cf;
}

// Temporal filter of block b into its LDS area Yb (Shape::YS complex).
//
// Pass 1 streams the stored spectrum row by row: lanes with a column c >= 1
// filter bin (r, c) into the swizzled column buffer; the lane holding column
// 0 (the packed DC/Nyquist pair W[r]) stages W[r] of every window frame in
// LDS instead. All lanes of the block read row r of every frame together, so
// each 128-byte row line is fetched once. Loads go out U rows at a time, one
// batch ahead of the arithmetic.
// Pass 2 (after a barrier) filters the DC and Nyquist bins from the staged
// column: unit q pairs rows q and bh-q (unit 0 bundles the self-paired rows 0
// and bh/2), i.e. four bins per unit and bh/2 units per block, and writes the
// repacked W'[r] = H0[r] + i Hh[r], H = (F[r] + conj F[-r]) / 2 the Hermitian
// part that the real inverse transform keeps.
// Both passes are called by every lane of the block (lb = lane in block).
// A stored bin as loaded: complex FP32, or with H the packed FP16 pair
// (unpacked when the bin is used, so a load in flight holds one register).
template <bool H>
struct SpecWord {
  using T = c32;
};
template <>
struct SpecWord<true> {
  using T = uint32_t;
};
template <bool H>
__device__ __forceinline__ typename SpecWord<H>::T spec_word(const c32* s, size_t i) {
  if constexpr (H) return reinterpret_cast<const uint32_t*>(s)[i];
  else return s[i];
}
template <bool H>
__device__ __forceinline__ c32 spec_value(typename SpecWord<H>::T v) {
  if constexpr (H) {
    const float2 u = unpack_h2(v);
    return cmake(u.x, u.y);
  } else {
    return v;
  }
}

// Rows per load batch and batch buffers of pass 1: one batch loads while the
// one before is filtered. Measured on gfx950 at 4K bw=32 (sweep us, FP32 /
// FP16 ring): 4 rows x 2 buffers 158 / 137; 2 x 4: 162 / 137; 1 x 8: 173 /
// 137; 4 x 4: 207 (spills) / 138; FP16 2 x 8: 141, 8 x 2: 137. Deeper
// prefetch does not pay: the sweep is bound by its compute phases.
constexpr int kPass1Rows = 4, kPass1Buffers = 2;

template <int BW, int BH, bool SEP, bool H, int BM, int DG, int SHP, int PAT, bool KH = false>
__device__ __forceinline__ void filter_pass1(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                             int lb, c32* Yb, const float4* tab, float4* kstate,
                                             bool kfirst, float dc) {
  using S = Shape<BW, BH>;
  using Word = typename SpecWord<H>::T;
  constexpr int FBW = S::FBW, NCOL = S::NCOL;
  constexpr int RS = S::LPB / NCOL;  // lanes sharing a column split its rows
  constexpr int NIT = BH / RS;       // rows per lane
  constexpr int U = NIT < kPass1Rows ? NIT : kPass1Rows;  // rows per batch; divides NIT
  constexpr int NB = NIT / U;                               // batches
  constexpr int D = NB < kPass1Buffers ? NB : kPass1Buffers;  // batch buffers
  static_assert(NB % D == 0, "whole rounds of batch buffers");
  const size_t base = (size_t)b * S::BSTR;
  const int c = lb % NCOL, r0 = lb / NCOL;
  constexpr int kst = KH ? 8 : 16;  // bytes of a bin's Kalman state
  char* kb = kstate ? reinterpret_cast<char*>(kstate) + (size_t)b * BH * FBW * kst : nullptr;
  float4 cf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if constexpr (SEP) cf = col_factor(tab, BH, c, dc);
  // All four slots are loaded and staged: slots past nwin repeat a window
  // spectrum (the host fills them), which keeps every load unconditional, so
  // the next batches stay in flight while this one is filtered.
  auto load = [&](int i0, Word(&va)[U][4]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = r0 + (i0 + u) * RS;
#pragma unroll
      for (int t = 0; t < 4; t++) va[u][t] = spec_word<H>(Wn.s[t], base + r * NCOL + c);
    }
  };
  auto work = [&](int i0, const Word(&va)[U][4]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = r0 + (i0 + u) * RS;
      c32 v[4];
#pragma unroll
      for (int t = 0; t < 4; t++) v[t] = spec_value<H>(va[u][t]);
      if (c == 0) {
#pragma unroll
        for (int t = 0; t < 4; t++) Yb[S::st(t, r)] = v[t];
      } else {
        BinConst kc;
        if constexpr (SEP) kc = bin_const_sep<BW, DG, SHP, PAT>(P, r, c, tab[r], cf);
        else kc = bin_const<BW, DG, SHP, PAT>(P, r, c, dc);
        Yb[S::yidx(c, r)] = filter_bin<BM, DG, SHP, KH>(P, v, kc.d, kc.sX, kc.sY, kc.lut, kc.kcn,
                                       kb ? kb + (size_t)(r * FBW + c) * kst : nullptr, kfirst);
      }
    }
  };
  Word buf[D][U][4];
#pragma unroll
  for (int d = 0; d + 1 < D; d++) load(d * U, buf[d]);
  if constexpr (NB == D) {
    load((D - 1) * U, buf[D - 1]);
#pragma unroll
    for (int d = 0; d < D; d++) work(d * U, buf[d]);
  } else {
#pragma unroll 1
    for (int k = 0; k < NB; k += D) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        // batch k + d + D - 1 goes into the buffer batch k + d - 1 used
        const int kn = k + d + D - 1;
        if (kn < NB) load(kn * U, buf[(d + D - 1) % D]);
        work((k + d) * U, buf[d]);
      }
    }
  }
}

template <int BW, int BH, bool SEP, bool H, int BM, int DG, int SHP, int PAT, bool KH = false>
__device__ __forceinline__ void filter_pass2(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                             int lb, c32* Yb, const float4* tab, float4* kstate,
                                             bool kfirst, float dc) {
  using S = Shape<BW, BH>;
  constexpr int FBW = S::FBW, NCOL = S::NCOL, H2 = BH / 2;
  const int nwin = Wn.nwin;
  constexpr int kst = KH ? 8 : 16;  // bytes of a bin's Kalman state
  char* kb = kstate ? reinterpret_cast<char*>(kstate) + (size_t)b * BH * FBW * kst : nullptr;
  float4 cf0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cfh = cf0;
  if constexpr (SEP) {
    cf0 = col_factor(tab, BH, 0, dc);
    cfh = col_factor(tab, BH, NCOL, dc);
  }
  for (int q = lb; q < H2; q += S::LPB) {
    // unit 0: self-paired rows 0 and bh/2; unit q: rows q and bh-q
    const int rA = q, rB = q == 0 ? H2 : BH - q;
    c32 f[4];  // filtered (rA,0), (rA,h), (rB,0), (rB,h)
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
      const int r = k < 2 ? rA : rB;
      const int col = (k & 1) ? NCOL : 0;
      c32 v[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const c32 wa = t < nwin ? Yb[S::st(t, r)] : cmake(0.0f, 0.0f);
        const c32 wb = t < nwin ? Yb[S::st(t, q == 0 ? r : BH - r)] : cmake(0.0f, 0.0f);
        // X[r][0] = (W[r] + conj W[-r]) / 2, X[r][h] = (W[r] - conj W[-r]) / 2i;
        // for a self-paired row (W[-r] = W[r]) that is (Re W[r], Im W[r])
        v[t] = col == 0 ? cmake(0.5f * (wa.x + wb.x), 0.5f * (wa.y - wb.y))
                        : cmake(0.5f * (wa.y + wb.y), -0.5f * (wa.x - wb.x));
      }
      BinConst kc;
      if constexpr (SEP) kc = bin_const_sep<BW, DG, SHP, PAT>(P, r, col, tab[r], col == 0 ? cf0 : cfh);
      else kc = bin_const<BW, DG, SHP, PAT>(P, r, col, dc);
      const c32 y = filter_bin<BM, DG, SHP, KH>(P, v, kc.d, kc.sX, kc.sY, kc.lut, kc.kcn,
                               kb ? kb + (size_t)(r * FBW + col) * kst : nullptr, kfirst);
      // f[k] with a runtime k would live in scratch: select instead
      f[0] = k == 0 ? y : f[0];
      f[1] = k == 1 ? y : f[1];
      f[2] = k == 2 ? y : f[2];
      f[3] = k == 3 ? y : f[3];
    }
    if (q == 0) {
      // the Hermitian part of a self-paired bin is its real part
      Yb[S::yidx(0, rA)] = cmake(f[0].x, f[1].x);
      Yb[S::yidx(0, rB)] = cmake(f[2].x, f[3].x);
    } else {
      const c32 h0 = cscale(cadd(f[0], cconj(f[2])), 0.5f);
      const c32 hh = cscale(cadd(f[1], cconj(f[3])), 0.5f);
      Yb[S::yidx(0, rA)] = cmake(h0.x - hh.y, h0.y + hh.x);
      Yb[S::yidx(0, rB)] = cmake(h0.x + hh.y, -h0.y + hh.x);
    }
  }
}

template <int BW, int BH, bool SEP, bool H, int BM, int DG, int SHP, int PAT, bool KH = false>
__device__ __forceinline__ void filter_passes(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                              bool live, int lb, c32* Yb, const float4* tab,
                                              float4* kstate, bool kfirst, float dc) {
  if (live)
    filter_pass1<BW, BH, SEP, H, BM, DG, SHP, PAT, KH>(P, Wn, b, lb, Yb, tab, kstate, kfirst, dc);
  wave_sync();
  if (live)
    filter_pass2<BW, BH, SEP, H, BM, DG, SHP, PAT, KH>(P, Wn, b, lb, Yb, tab, kstate, kfirst, dc);
}

// The Wiener modes get one instance per (degrid, sharpen, sigma pattern)
// combination; Kalman and the one-frame modes read the flags at run time.
template <int BW, int BH, bool SEP, bool H, int BM>
__device__ __forceinline__ void filter_mode(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                            bool live, int lb, c32* Yb, const float4* tab,
                                            float4* kstate, bool kfirst, float dc) {
#define FK_PASSES(DG, SHP, PAT) \
  filter_passes<BW, BH, SEP, H, BM, DG, SHP, PAT>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc)
  if constexpr (BM >= 2) {
    const int sel = (P.dg ? 4 : 0) | (P.shp ? 2 : 0) | (P.pat ? 1 : 0);
    switch (sel) {
    case 7: FK_PASSES(1, 1, 1); break;
    case 6: FK_PASSES(1, 1, 0); break;
    case 5: FK_PASSES(1, 0, 1); break;
    case 4: FK_PASSES(1, 0, 0); break;
    case 3: FK_PASSES(0, 1, 1); break;
    case 2: FK_PASSES(0, 1, 0); break;
    case 1: FK_PASSES(0, 0, 1); break;
    default: FK_PASSES(0, 0, 0);
    }
  } else if constexpr (BM == 0 && H) {
    // Kalman over an FP16 ring: the state is FP16 pairs on integer samples
    if (P.khalf)
      filter_passes<BW, BH, SEP, H, 0, 2, 2, 2, true>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst,
                                                       dc);
    else
      FK_PASSES(2, 2, 2);
  } else {
    FK_PASSES(2, 2, 2);
  }
#undef FK_PASSES
}

// Filter of block b into Yb; contains the wave barrier between its two passes,
// so every lane of the wave calls it (live = this lane's block exists). Yb must
// belong to the calling wave. With SEP the sharpen and degrid constants come
// from the separable factors in tab (LDS: bh row factors, then bw/2+1 column
// factors); otherwise from the full tables. When bit `cur` of fwdMask is set,
// the current frame was transformed by this wave just before and its block DC
// is read from the last slot of Yb.
template <int BW, int BH, bool SEP, bool H>
__device__ __forceinline__ void filter_block(const PlaneArgs& P, const WindowArgs& Wn, int b,
                                             bool live, int lb, c32* Yb, const float4* tab,
                                             float4* kstate, bool kfirst, int fwdMask = 0) {
  using S = Shape<BW, BH>;
  // block DC X[0][0] = Re W[0] of the current frame, for the degrid term
  float dc = 0.0f;
  if (live && P.dg)
    dc = ((fwdMask >> Wn.cur) & 1) ? Yb[S::YS - 1].x
                                   : spec_load<H>(Wn.s[Wn.cur], (size_t)b * S::BSTR).x;
  if (fwdMask) wave_sync();  // pass 1 stages over that slot
  // one instance per temporal mode: only the plane's mode is in the hot loop
  switch (P.bm) {
  case 4: filter_mode<BW, BH, SEP, H, 4>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc); break;
  case 3: filter_mode<BW, BH, SEP, H, 3>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc); break;
  case 2: filter_mode<BW, BH, SEP, H, 2>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc); break;
  case 1: filter_mode<BW, BH, SEP, H, 1>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc); break;
  case 0: filter_mode<BW, BH, SEP, H, 0>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc); break;
  default: filter_mode<BW, BH, SEP, H, -1>(P, Wn, b, live, lb, Yb, tab, kstate, kfirst, dc);
  }
  wave_sync();
}

// Inverse 2-D C2R of the filtered block held in its LDS area (column 0
// packed). On return lanes j < NROW hold real rows 2j (xr) and 2j+1 (xi),
// unnormalised (x bw*bh). All lanes of the wave call it; Yb belongs to the
// wave.
template <int BW, int BH>
__device__ __forceinline__ void inverse_block(bool live, int j, c32* Yb, float (&xr)[BW],
                                              float (&xi)[BW]) {
  using S = Shape<BW, BH>;
  if (live && j < S::NCOL) {
    float yr[BH], yi[BH];
#pragma unroll
    for (int r = 0; r < BH; r++) {
      const c32 v = Yb[S::yidx(j, r)];
      yr[r] = v.x;
      yi[r] = v.y;
    }
    fft_reg<BH, true>(yr, yi);
#pragma unroll
    for (int r = 0; r < BH; r++) Yb[S::yidx(j, r)] = cmake(yr[r], yi[r]);
  }
  wave_sync();
  if (live && j < S::NROW) {
    const int p = j;
    {
      const c32 v = Yb[S::yidx(0, 2 * p)], u = Yb[S::yidx(0, 2 * p + 1)];
      xr[0] = v.x;
      xi[0] = u.x;
      xr[BW / 2] = v.y;
      xi[BW / 2] = u.y;
    }
#pragma unroll
    for (int c = 1; c < S::NCOL; c++) {
      const c32 a = Yb[S::yidx(c, 2 * p)], q = Yb[S::yidx(c, 2 * p + 1)];
      xr[c] = a.x - q.y;
      xi[c] = a.y + q.x;
      xr[BW - c] = a.x + q.y;
      xi[BW - c] = -a.y + q.x;
    }
    fft_reg<BW, true>(xr, xi);
  }
}

// Filter + inverse + synthesis window; writes windowed blocks [b][ly][lx].
template <int BW, int BH, bool H>
__global__ __launch_bounds__(64) void k_filter_inverse(PlaneArgs P, WindowArgs Wn,
                                                       float4* kstate, int kfirst,
                                                       float* blocks) {
  using S = Shape<BW, BH>;
  __shared__ c32 lds[S::LDS_C];
  const int lane = threadIdx.x, g = lane / S::LPB, j = lane - g * S::LPB;
  const int b = blockIdx.x * S::G + g;
  const bool live = b < P.nb;
  c32* Yb = lds + g * S::YS;
  filter_block<BW, BH, false, H>(P, Wn, b, live, j, Yb, nullptr, kstate, kfirst != 0);
  float xr[BW], xi[BW];
  inverse_block<BW, BH>(live, j, Yb, xr, xi);
  if (live && j < S::NROW) {
    int x0, y0, par;
    block_origin(P, b, x0, y0, par);
    const float* win = par ? P.winSynB : P.winSynA;
    float* dst = blocks + (size_t)b * BW * BH + (2 * j) * BW;
    const float sc = 1.0f / (float)(BW * BH);
#pragma unroll
    for (int k = 0; k < BW; k++) {
      dst[k] = (xr[k] * sc) * win[2 * j * BW + k];
      dst[BW + k] = (xi[k] * sc) * win[(2 * j + 1) * BW + k];
    }
  }
}

// ---------------------------------------------------------------- synthesis

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// Float -> sample (src/cpu/reference.cpp:107-138; src/cl/kernels.h:436-447).
__device__ __forceinline__ void store_sample(const PlaneArgs& P, void* out, size_t i, float v) {
#pragma clang fp contract(off)
  if (P.pix == 8) {
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    static_cast<uint8_t*>(out)[i] = (uint8_t)floorf(v * 255.0f + 0.5f);
  } else if (P.pix == 16) {
    v = v * P.invNorm;
    v = v < P.maxPix ? v : P.maxPix;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    static_cast<uint16_t*>(out)[i] = (uint16_t)floorf(v * 65535.0f + 0.5f);
  } else {
    static_cast<float*>(out)[i] = v;
  }
}

#ifdef FK_COMMON_KERNELS
// Overlap-add of the windowed blocks covering each pixel, in block-index
// order (src/cpu/reference.cpp:415-436), plus chroma centring, then pack.
__global__ __launch_bounds__(256) void k_synthesize(PlaneArgs P, const float* blocks,
                                                    void* out) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= P.W || y >= P.H) return;
  const int bw = P.bw, bh = P.bh;
  float sum = 0.0f;
  if (P.mode == 1) {
    const int xs = bw - P.ow, ys = bh - P.oh;
    int n0 = floordiv(x + P.ow - bw + 1, xs), n1 = floordiv(x + P.ow, xs);
    int m0 = floordiv(y + P.oh - bh + 1, ys), m1 = floordiv(y + P.oh, ys);
    if (n0 < 0) n0 = 0;
    if (m0 < 0) m0 = 0;
    if (n1 >= P.ncols) n1 = P.ncols - 1;
    if (m1 >= P.nrows) m1 = P.nrows - 1;
    for (int m = m0; m <= m1; m++) {
      const int ly = y + P.oh - m * ys;
      for (int n = n0; n <= n1; n++) {
        const int lx = x + P.ow - n * xs;
        if (lx < 0 || lx >= bw || ly < 0 || ly >= bh) continue;
        sum += blocks[((size_t)m * P.ncols + n) * bw * bh + ly * bw + lx];
      }
    }
  } else {
    for (int grid = 0; grid < 2; grid++) {
      const int sx = x + (grid ? bw / 2 : 0), sy = y + (grid ? bh / 2 : 0);
      const int i = sx / bw, j = sy / bh;
      if (i < P.nx && j < P.ny) {
        const size_t b = (size_t)grid * P.nx * P.ny + (size_t)j * P.nx + i;
        sum += blocks[b * bw * bh + (sy - j * bh) * bw + (sx - i * bw)];
      }
    }
  }
  if (P.chromaNorm != 0.0f) sum += P.chromaNorm;
  store_sample(P, out, (size_t)y * P.W + x, sum);
}
#endif  // FK_COMMON_KERNELS

} // namespace fk
