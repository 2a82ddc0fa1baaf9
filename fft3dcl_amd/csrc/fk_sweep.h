// Fused filter -> inverse FFT -> windowed overlap-add -> pack (mode 1, half
// overlap: ow = bw/2, oh = bh/2, every output pixel covered by 2x2 blocks).
//
// Replaces the reference's fft2d_inv + synthesize pair (src/cl/kernels.h:
// 368-448) and its `blocks` round trip through device memory (nb*bw*bh
// floats written and read again per frame).
//
// Tiling. Output tiles are (bw/2) x (bh/2); block (m, n) covers tiles
// (m-1..m) x (n-1..n). A workgroup of WPW waves owns a strip of NBW = WPW*G
// adjacent block columns and a band of R block rows, and sweeps down the
// band one block row at a time:
//   1. each wave filters + inverse-transforms its G blocks (fk_kernels.h),
//      after transforming them for a frame being analysed in the same pass;
//   2. each block's windowed output goes to its own LDS area;
//   3. one pass over the strip forms each pixel of a tile row: the two
//      horizontally overlapping block halves of this row plus the lower half
//      carried in registers from the previous block row (fixed order: the
//      result is bit-identical run to run, no atomics); interior tiles are
//      packed and stored to the output plane; tiles on the band's or strip's
//      border are stored as float partial sums into one of four partial
//      planes chosen by the (band, strip) parity, so neighbours never write
//      the same word;
//   4. k_combine_seams then adds the partials of those border tiles in a
//      fixed (band, strip) order and packs them.
#pragma once

#include "fk_mfma.h"

namespace fk {

// A border tile and the partial planes that hold its contributions, in the
// order they are added (bands, then strips): plane k is (planes >> 2k) & 3.
struct SeamTile {
  int x, y;    // tile column, tile row
  int planes;  // packed plane indices
  int np;      // 1..4
};

struct SweepArgs {
  int R;             // block rows per band
  int nstrips, nbands;
  int ntx, nty;      // output tiles covering the plane
  float* part;       // 4 partial planes of H rows of ppitch floats
  int ppitch;        // row pitch of the partial planes: W rounded up to 4
  void* out;         // packed output plane (u8/u16/f32)
  const SeamTile* seams;  // border tiles with their contributing planes
  int nseams;
  // Fused analysis of one new frame (the FWD instance of the sweep): its
  // plane, the ring spectrum to write, and the window slots that hold it
  // (bit t = slot t).
  const void* fwdPlane;
  c32* fwdSpec;
  int fwdMask;
};

template <int BW, int BH, int WPW>
struct SweepShape {
  using S = Shape<BW, BH>;
  static constexpr int SX = BW / 2, SY = BH / 2;  // tile size
  static constexpr int NBW = S::G * WPW;          // block columns per strip
  static constexpr int CW = (NBW + 1) * SX;       // strip width in pixels
  static constexpr int NT = 64 * WPW;             // threads per workgroup
  static constexpr int NCARRY = (SY * CW + NT - 1) / NT;
  static constexpr int YFLOATS = 2 * S::G * S::YS * WPW;
  // row pitch of a windowed block in its LDS area (16-byte aligned rows)
  static constexpr int OP = BW + 4;
  // tile-row pass: 4-sample quads when the tile width allows, else samples
  static constexpr bool QUADS = SX % 4 == 0;
  static constexpr int QW = CW / 4;                      // quads per strip row
  static constexpr int NQ = (SY * QW + NT - 1) / NT;     // quads per thread
  static constexpr int NCF = QUADS && 4 * NQ > NCARRY ? 4 * NQ : NCARRY;  // carry floats
  static_assert(BH * OP <= 2 * S::YS, "windowed block must fit its block area");
  static constexpr int AREA = (YFLOATS + 3) & ~3;
  // constant tables kept in LDS after the block/strip area: the separable
  // sharpen/degrid factors (bh rows, then bw/2+1 columns, float4 each), then
  // the synthesis window factors synX (bw) and synY (bh)
  static constexpr int NFAC = 4 * (BH + BW / 2 + 1);
  static constexpr int TAB = NFAC + BW + BH;
  static constexpr int LDS_FLOATS = AREA + TAB;
};

// Stores the first n <= 4 of 4 consecutive samples at idx with store_sample's
// arithmetic: one packed store when all 4 are wanted and idx is a multiple of
// 4, else sample by sample.
__device__ __forceinline__ void store_sample4(const PlaneArgs& P, void* out, size_t idx,
                                              float4 v, int n) {
#pragma clang fp contract(off)
  const float s[4] = {v.x, v.y, v.z, v.w};
  if (n < 4 || (idx & 3) != 0) {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (k < n) store_sample(P, out, idx + k, s[k]);
    return;
  }
  if (P.pix == 8) {
    unsigned w = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float e = s[k] < 0.0f ? 0.0f : (s[k] > 1.0f ? 1.0f : s[k]);
      w |= (unsigned)(uint8_t)floorf(e * 255.0f + 0.5f) << (8 * k);
    }
    *reinterpret_cast<unsigned*>(static_cast<uint8_t*>(out) + idx) = w;
  } else if (P.pix == 16) {
    unsigned w[2] = {0u, 0u};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float e = s[k] * P.invNorm;
      e = e < P.maxPix ? e : P.maxPix;
      e = e < 0.0f ? 0.0f : (e > 1.0f ? 1.0f : e);
      w[k >> 1] |= (unsigned)(uint16_t)floorf(e * 65535.0f + 0.5f) << (16 * (k & 1));
    }
    *reinterpret_cast<uint2*>(static_cast<uint16_t*>(out) + idx) = make_uint2(w[0], w[1]);
  } else {
    *reinterpret_cast<float4*>(static_cast<float*>(out) + idx) = v;
  }
}

// The sweep of one workgroup: strip `strip`, band `band` of the plane.
// DFT: 0 register FFTs for the inverse; mm::F16 / mm::BF16 the MFMA form
// (32x32 blocks, no fused analysis).
template <int BW, int BH, int WPW, bool FWD, bool H, int DFT>
__device__ __forceinline__ void sweep_body(const PlaneArgs& P, const WindowArgs& Wn,
                                           float4* kstate, int kfirst, const SweepArgs& A,
                                           int strip, int band) {
  using S = Shape<BW, BH>;
  using T = SweepShape<BW, BH, WPW>;
  constexpr int SX = T::SX, SY = T::SY, NBW = T::NBW, CW = T::CW;
  __shared__ __attribute__((aligned(16))) float lds[T::LDS_FLOATS];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / S::LPB, j = lane - g * S::LPB;
  const int n0 = strip * NBW, m0 = band * A.R;
  const int mEnd = min(m0 + A.R, P.nrows);
  const int nlive = min(NBW, P.ncols - n0);
  const int n = n0 + w * S::G + g;  // this lane's block column
  const bool colLive = n < P.ncols;
  c32* Yb = reinterpret_cast<c32*>(lds) + (w * S::G + g) * S::YS;
  float* part = A.part + (size_t)((band & 1) * 2 + (strip & 1)) * A.ppitch * P.H;
  const int validCols = (nlive + 1) * SX;
  const int x0 = (n0 - 1) * SX;  // image column of strip column 0
  float carry[T::NCF];
  // constant tables (visible to every wave after the row loop's first barrier)
  float* tabf = lds + T::AREA;
  for (int i = tid; i < T::TAB; i += T::NT) {
    float v;
    if (i < 4 * BH) v = reinterpret_cast<const float*>(P.sepRow)[i];
    else if (i < T::NFAC) v = reinterpret_cast<const float*>(P.sepCol)[i - 4 * BH];
    else if (i < T::NFAC + BW) v = P.synX[i - T::NFAC];
    else v = P.synY[i - T::NFAC - BW];
    tabf[i] = v;
  }
  const float4* tab = reinterpret_cast<const float4*>(tabf);
  const float* synX = tabf + T::NFAC;
  const float* synY = synX + BW;

  for (int m = m0; m < mEnd; m++) {
    const int b = m * P.ncols + n;
    __syncthreads();  // the previous row's strip reads are done: LDS is the block areas again
    if (FWD) {
      // analysis of the new frame into the ring; the block DC also goes to
      // the last slot of Yb, which the forward transpose never touches
      const size_t d = (size_t)b * S::BSTR + j;
      forward_block<BW, BH, S::SPLIT_A>(
          P, A.fwdPlane, b, colLive, 0, j, Yb, [&](const float(&yr)[BH], const float(&yi)[BH]) {
#pragma unroll
            for (int r = 0; r < BH; r++)
              spec_store<H>(A.fwdSpec, d + r * S::NCOL, cmake(yr[r], yi[r]));
            // the DC as the filter would read it back from the ring
            if (j == 0) Yb[S::YS - 1] = spec_round<H>(cmake(yr[0], yi[0]));
          });
      wave_sync_global();  // the ring stores are read back by the filter
    }
    filter_block<BW, BH, true, H>(P, Wn, b, colLive, j, Yb, tab, kstate, kfirst != 0,
                               FWD ? A.fwdMask : 0);
    if constexpr (DFT != 0) {
      static_assert(BW == 32 && BH == 32 && !FWD, "the MFMA inverse is for 32x32 blocks");
      // the wave's blocks one after the other, each by all 64 lanes, with
      // the constant operands loaded once for them
      mm::InvOps ops;
      mm::load_inv_ops(P.dft, lane, ops);
      for (int gg = 0; gg < S::G; gg++) {
        if (n0 + w * S::G + gg >= P.ncols) break;  // wave-uniform
        c32* Yg = reinterpret_cast<c32*>(lds) + (w * S::G + gg) * S::YS;
        mm::inverse_block_mfma<DFT>(ops, Yg, S::PITCH, T::OP, synX, synY, lane);
      }
    } else {
      float xr[BW], xi[BW];
      inverse_block<BW, BH>(colLive, j, Yb, xr, xi);
      const bool rowLane = colLive && j < S::NROW;
      if (rowLane) {
        // half overlap: both parities use the same separable synthesis window
        const float fy0 = synY[2 * j], fy1 = synY[2 * j + 1];
#pragma unroll
        for (int k = 0; k < BW; k++) {
          const float fx = synX[k];
          xr[k] *= fx * fy0;
          xi[k] *= fx * fy1;
        }
      }
      // windowed block -> its own LDS area as rows of OP floats (the area
      // belongs to this wave, whose inverse reads of it are done)
      wave_sync();
      float* const blk = reinterpret_cast<float*>(Yb);
      if (rowLane) {
        float* o = blk + (2 * j) * T::OP;
#pragma unroll
        for (int k = 0; k < BW; k++) {
          o[k] = xr[k];
          o[T::OP + k] = xi[k];
        }
      }
    }
    __syncthreads();  // every block of the strip is in LDS
    // tile row m-1: carried lower half of row m-1 + upper half of row m.
    // Strip column c lies in tile column tc = c / SX, covered by the right
    // half of local block tc-1 and the left half of local block tc.
    if constexpr (T::QUADS) {
      // 4 samples per step: 16-byte LDS reads and stores (x0 is a multiple
      // of 4; a quad crossing the plane's right edge stores its inside part)
      const int y0 = (m - 1) * SY;
      const bool firstRow = m == m0;
      int r = tid / T::QW, q = tid - (tid / T::QW) * T::QW;
#pragma unroll
      for (int i = 0; i < T::NQ; i++) {
        if (r < SY) {
          const int c = 4 * q;
          const bool ok = c < validCols;
          const int tc = c / SX, x = c - tc * SX;
          const int tl = tc > 0 ? tc - 1 : 0, tr = tc < nlive ? tc : 0;
          const float* L = lds + (size_t)tl * 2 * S::YS + SX + x;
          const float* Rb = lds + (size_t)tr * 2 * S::YS + x;
          float4 top = make_float4(0.0f, 0.0f, 0.0f, 0.0f), bot = top;
          if (ok) {
            if (tc == 0) {
              top = *reinterpret_cast<const float4*>(Rb + r * T::OP);
              bot = *reinterpret_cast<const float4*>(Rb + (r + SY) * T::OP);
            } else if (tc == nlive) {
              top = *reinterpret_cast<const float4*>(L + r * T::OP);
              bot = *reinterpret_cast<const float4*>(L + (r + SY) * T::OP);
            } else {
              const float4 a0 = *reinterpret_cast<const float4*>(L + r * T::OP);
              const float4 b0 = *reinterpret_cast<const float4*>(Rb + r * T::OP);
              const float4 a1 = *reinterpret_cast<const float4*>(L + (r + SY) * T::OP);
              const float4 b1 = *reinterpret_cast<const float4*>(Rb + (r + SY) * T::OP);
              top = make_float4(a0.x + b0.x, a0.y + b0.y, a0.z + b0.z, a0.w + b0.w);
              bot = make_float4(a1.x + b1.x, a1.y + b1.y, a1.z + b1.z, a1.w + b1.w);
            }
          }
          float4 v = top;
          if (!firstRow)
            v = make_float4(carry[4 * i] + top.x, carry[4 * i + 1] + top.y,
                            carry[4 * i + 2] + top.z, carry[4 * i + 3] + top.w);
          const int xx = x0 + c, y = y0 + r;
          if (ok && xx >= 0 && xx < P.W && y >= 0 && y < P.H) {
            if (firstRow || c < SX || c >= nlive * SX) {
              // border tile: partial sums (the pitch keeps the quad aligned)
              *reinterpret_cast<float4*>(part + (size_t)y * A.ppitch + xx) = v;
            } else {
              if (P.chromaNorm != 0.0f) {
                v.x += P.chromaNorm;
                v.y += P.chromaNorm;
                v.z += P.chromaNorm;
                v.w += P.chromaNorm;
              }
              store_sample4(P, A.out, (size_t)y * P.W + xx, v, min(4, P.W - xx));
            }
          }
          carry[4 * i] = bot.x;
          carry[4 * i + 1] = bot.y;
          carry[4 * i + 2] = bot.z;
          carry[4 * i + 3] = bot.w;
        }
        q += T::NT;
        while (q >= T::QW) {
          q -= T::QW;
          r++;
        }
      }
    } else {
      const int y0 = (m - 1) * SY;
      const bool firstRow = m == m0;
      int r = tid / CW, c = tid - (tid / CW) * CW;  // element tid + NT*i, advanced below
#pragma unroll
      for (int i = 0; i < T::NCARRY; i++) {
        if (r < SY) {
          const bool ok = c < validCols;
          const int tc = c / SX, x = c - tc * SX;
          const int tl = tc > 0 ? tc - 1 : 0, tr = tc < nlive ? tc : 0;  // used when in range
          const float* L = lds + (size_t)tl * 2 * S::YS + SX + x;  // block tc-1, right half
          const float* Rb = lds + (size_t)tr * 2 * S::YS + x;      // block tc, left half
          float top = 0.0f, bot = 0.0f;
          if (ok) {
            if (tc == 0) {
              top = Rb[r * T::OP];
              bot = Rb[(r + SY) * T::OP];
            } else if (tc == nlive) {
              top = L[r * T::OP];
              bot = L[(r + SY) * T::OP];
            } else {
              top = L[r * T::OP] + Rb[r * T::OP];
              bot = L[(r + SY) * T::OP] + Rb[(r + SY) * T::OP];
            }
          }
          const float v = firstRow ? top : carry[i] + top;
          const int xx = x0 + c, y = y0 + r;
          if (ok && xx >= 0 && xx < P.W && y >= 0 && y < P.H) {
            if (firstRow || c < SX || c >= nlive * SX)
              part[(size_t)y * A.ppitch + xx] = v;  // border tile: partial sum
            else
              store_sample(P, A.out, (size_t)y * P.W + xx,
                           P.chromaNorm != 0.0f ? v + P.chromaNorm : v);
          }
          carry[i] = bot;
        }
        c += T::NT;
        while (c >= CW) {
          c -= CW;
          r++;
        }
      }
    }
  }
  // tile row mEnd-1: lower half of the band's last block row (a border row)
  if constexpr (T::QUADS) {
    const int y0 = (mEnd - 1) * SY;
    int r = tid / T::QW, q = tid - (tid / T::QW) * T::QW;
#pragma unroll
    for (int i = 0; i < T::NQ; i++) {
      const int c = 4 * q, x = x0 + c, y = y0 + r;
      if (r < SY && c < validCols && x >= 0 && x < P.W && y >= 0 && y < P.H)
        *reinterpret_cast<float4*>(part + (size_t)y * A.ppitch + x) =
            make_float4(carry[4 * i], carry[4 * i + 1], carry[4 * i + 2], carry[4 * i + 3]);
      q += T::NT;
      while (q >= T::QW) {
        q -= T::QW;
        r++;
      }
    }
  } else {
    const int y0 = (mEnd - 1) * SY;
    int r = tid / CW, c = tid - (tid / CW) * CW;
#pragma unroll
    for (int i = 0; i < T::NCARRY; i++) {
      const int x = x0 + c, y = y0 + r;
      if (r < SY && c < validCols && x >= 0 && x < P.W && y >= 0 && y < P.H)
        part[(size_t)y * A.ppitch + x] = carry[i];
      c += T::NT;
      while (c >= CW) {
        c -= CW;
        r++;
      }
    }
  }
}

// MINB: workgroups per CU the register allocation is sized for (its VGPR
// budget is 512 / (MINB * WPW / 4) per lane; more resident waves hide more
// latency when the LDS holds them).
template <int BW, int BH, int WPW, bool FWD, bool H, int DFT = 0, int MINB = 2>
__global__ __launch_bounds__(64 * WPW, MINB) void k_filter_synth(PlaneArgs P, WindowArgs Wn,
                                                              float4* kstate, int kfirst,
                                                              SweepArgs A) {
  sweep_body<BW, BH, WPW, FWD, H, DFT>(P, Wn, kstate, kfirst, A, blockIdx.x, blockIdx.y);
}

// Adds the partial sums of border tile si from every workgroup that touched
// it, in (band, strip) order (the planes listed with the tile), adds chroma
// centring and packs: one wave per tile, four samples per lane when the tile
// width is a multiple of 4.
__device__ __forceinline__ void combine_body(const PlaneArgs& P, const SweepArgs& A, int si) {
  const int lane = threadIdx.x & 63;
  if (si >= A.nseams) return;  // no barriers below
  const SeamTile tile = A.seams[si];
  const int sx = P.bw / 2, sy = P.bh / 2;
  const size_t plane = (size_t)A.ppitch * P.H;
  const int np = tile.np;
  int planes[4];
#pragma unroll
  for (int k = 0; k < 4; k++) planes[k] = (tile.planes >> (2 * k)) & 3;
  if (sx % 4 == 0) {
    const int qpr = sx / 4;  // 4-sample groups per tile row
    for (int q = lane; q < qpr * sy; q += 64) {
      const int x = tile.x * sx + (q % qpr) * 4, y = tile.y * sy + q / qpr;
      if (x >= P.W || y >= P.H) continue;
      const size_t pidx = (size_t)y * A.ppitch + x;
      float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      for (int k = 0; k < np; k++) {
        const float4 u = *reinterpret_cast<const float4*>(A.part + planes[k] * plane + pidx);
        v.x += u.x;
        v.y += u.y;
        v.z += u.z;
        v.w += u.w;
      }
      if (P.chromaNorm != 0.0f) {
        v.x += P.chromaNorm;
        v.y += P.chromaNorm;
        v.z += P.chromaNorm;
        v.w += P.chromaNorm;
      }
      store_sample4(P, A.out, (size_t)y * P.W + x, v, min(4, P.W - x));
    }
  } else {
    for (int q = lane; q < sx * sy; q += 64) {
      const int x = tile.x * sx + q % sx, y = tile.y * sy + q / sx;
      if (x >= P.W || y >= P.H) continue;
      const size_t pidx = (size_t)y * A.ppitch + x;
      float v = 0.0f;
      for (int k = 0; k < np; k++) v += A.part[planes[k] * plane + pidx];
      store_sample(P, A.out, (size_t)y * P.W + x, P.chromaNorm != 0.0f ? v + P.chromaNorm : v);
    }
  }
}

#ifdef FK_COMMON_KERNELS
__global__ __launch_bounds__(256) void k_combine_seams(PlaneArgs P, SweepArgs A) {
  combine_body(P, A, blockIdx.x * 4 + (threadIdx.x >> 6));
}
#endif  // FK_COMMON_KERNELS

}  // namespace fk
