// Parameter contract and constant tables of the FFT3D denoise path (host side).
//
// Restates the reference's contract so the GPU path receives exactly the
// same validated parameters and tables as its CPU backend:
//   validate()          src/params.h:76-148 (roundPow2Clamp :62-72)
//   window tables       src/cpu/tables.cpp:8-105, reference.cpp:47-58
//   sigma pattern       src/cpu/tables.cpp:107-137
//   sharpen LUT         src/cpu/tables.cpp:139-156
//   degrid sample       src/cpu/reference.cpp:172-198 (computed here with an
//                       FP64 DFT of the analysis window instead of FP32 FFT)
// All tables are evaluated in double and stored as float, as the reference
// does; float expressions keep the reference's operand order.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace fft3d {

// Script-level parameters (src/params.h:8-27), user units, reference defaults.
struct Script {
  float sigma = 2.0f, beta = 1.0f;
  int bw = 32, bh = 32, bt = 3;
  float sharpen = 0.0f;
  int plane = 0, mode = 1, bordersize = 1, precision = 0;
  float degrid = 1.0f, scutoff = 0.3f, svr = 1.0f, smin = 4.0f, smax = 20.0f, kratio = 2.0f;
  int ow = 16, oh = 16, wintype = 0, interlaced = 0;
  float sigma2 = 2.0f, sigma3 = 2.0f, sigma4 = 2.0f;
  int oldfft = -1;
};

// Validated, derived parameters for one plane (src/params.h:30-59).
struct Params {
  float sigma, beta;
  int bw, bh, bm, bt;
  float sharpen;
  bool sharpenOn;
  int mode, bordersize, precision;
  float degrid;
  bool degridOn;
  float scutoff, svr, smin, smax, kratio;
  int ow, oh, wintype, interlaced;
  float sigma2, sigma3, sigma4;
  bool usePattern;
  int current;
  int nx, ny, totw, toth, oldfft;
  float wienerS, betaX, kalmanCN, ssmin, ssmax;
};

inline int roundPow2Clamp(int v) {
  unsigned logn = 0;
  for (unsigned i = static_cast<unsigned>(v - 1); i > 0; i /= 2) logn++;
  int r = 1 << logn;
  return r < 4 ? 4 : (r > 512 ? 512 : r);
}

// Returns an empty string on success, else the reference's error text.
inline std::string validate(const Script& s, int width, int height, int bits, Params& o) {
  o = Params{};
  o.sigma = s.sigma / 255.0f;
  o.beta = s.beta;
  o.bm = s.bt;
  if (o.bm < -1 || o.bm > 4) return "FFT3DCL: valid modes for bt are -1,0,1,2,3,4";
  o.bt = s.bt > 0 ? s.bt : 1;
  o.sharpen = s.sharpen;
  o.sharpenOn = s.sharpen != 0.0f;
  o.mode = (s.mode < 0 || s.mode > 2) ? 0 : s.mode;
  o.bordersize = s.bordersize;
  o.degrid = s.degrid;
  o.degridOn = s.degrid != 0.0f;
  o.scutoff = s.scutoff;
  o.svr = s.svr;
  o.smin = s.smin / 255.0f;
  o.smax = s.smax / 255.0f;
  o.kratio = s.kratio;
  o.wintype = (s.wintype < 0 || s.wintype > 2) ? 0 : s.wintype;
  o.interlaced = s.interlaced;
  o.sigma2 = s.sigma2;
  o.sigma3 = s.sigma3;
  o.sigma4 = s.sigma4;
  o.usePattern = s.sigma != s.sigma2 || s.sigma3 != s.sigma4 || s.sigma != s.sigma3;
  o.oldfft = s.oldfft;
  o.current = o.bm == 1 ? 0 : (o.bm == 4 ? 2 : 1);
  o.bw = roundPow2Clamp(s.bw);
  o.bh = roundPow2Clamp(s.bh);
  // ow/oh follow the block-size rounding; ow is forced even, both clamped to half a block.
  o.ow = static_cast<int>(s.ow * (float)o.bw / (float)s.bw) & ~1;
  o.oh = static_cast<int>(s.oh * (float)o.bh / (float)s.bh);
  if (o.oh * 2 > o.bh) o.oh = o.bh / 2;
  if (o.ow * 2 > o.bw) o.ow = o.bw / 2;
  int precision = bits > 8 ? 2 : s.precision;
  o.precision = (precision < 0 || precision > 2) ? 1 : precision;
  if (o.mode == 2) return "FFT3DCL: mode 2 is not implemented yet in this port (use mode 0 or 1)";
  if (o.interlaced) return "FFT3DCL: interlaced is not implemented yet in this port";
  if (o.mode == 0 || (o.bw == o.ow * 2 && o.bh == o.oh * 2)) {
    o.ny = static_cast<int>((height + o.bh * 1.5 - 1) / o.bh);
    o.nx = static_cast<int>((width + o.bw * 1.5 - 1) / o.bw);
  } else {
    o.nx = ((width + o.ow - 1) / (o.bw - o.ow) + 2) / 2;
    o.ny = ((height + o.oh - 1) / (o.bh - o.oh) + 2) / 2;
  }
  o.totw = o.nx * o.bw;
  o.toth = o.ny * o.bh;
  o.wienerS = o.sigma * o.sigma * o.bt * o.bw * o.bh;
  o.betaX = (o.beta - 1.0f) / o.beta;
  o.kalmanCN = o.sigma * o.sigma * o.bw * o.bh;
  o.ssmin = o.smin * o.smin * o.bw * o.bh * o.bt;
  o.ssmax = o.smax * o.smax * o.bw * o.bh * o.bt;
  return {};
}

// Block grid of one plane (src/cpu/reference.cpp:23-46). Mode 1: ncols x nrows
// blocks at (-ow + n*(bw-ow), -oh + m*(bh-oh)), index m*ncols+n, parity n&1.
// Mode 0: grid U (nx*ny at (i*bw, j*bh)) then grid S shifted by -bw/2,-bh/2.
struct Grid {
  int ncols = 0, nrows = 0, nb = 0;
  static Grid of(const Params& p) {
    Grid g;
    if (p.mode == 0) {
      g.ncols = p.nx;
      g.nrows = 2 * p.ny;
      g.nb = 2 * p.nx * p.ny;
    } else {
      g.ncols = 2 * p.nx;
      g.nrows = 2 * p.ny;
      g.nb = g.ncols * g.nrows;
    }
    return g;
  }
};

struct Tables {
  std::vector<float> winAna, winSynA, winSynB; // bw*bh, row-major
  std::vector<float> pattern;                   // fbw*bh, empty when off
  std::vector<float> sharpen;                   // fbw*bh, empty when off
  std::vector<float> degrid;                    // 2*fbw*bh (re,im), empty when off
  // Mode 1 windows are separable: winAna = anaX[lx]*anaY[ly] and, at half
  // overlap, winSynA = winSynB = synX[lx]*synY[ly]/(bw*bh) (the inverse
  // transform's normalisation folded into synY). Empty in mode 0.
  std::vector<float> anaX, anaY, synX, synY;
  // Mode 1 separable factors of the sharpen and degrid tables, 4 floats per
  // entry. Row r: {shY, gyRe, gyIm, strength}; column c < fbw: {shX, gxRe,
  // gxIm, 0}. sharpen[r][c] = strength - shY[r] * shX[c] (the exponential of a
  // sum of a row and a column term) and gs[r][c] = gx[c] * gy[r] (the DFT of a
  // separable window is the product of the 1-D DFTs), both to float rounding.
  std::vector<float> sepRow, sepCol;
};

namespace detail {
constexpr double kPi = 3.14159265358979323846;

// Half-cosine ramps of the mode-1 windows: analysis an[i], synthesis sy[i], i < o.
inline void ramps(int o, int wintype, std::vector<double>& an, std::vector<double>& sy) {
  an.assign(o, 1.0);
  sy.assign(o, 1.0);
  for (int i = 0; i < o; i++) {
    double c = std::cos(kPi * (i - o + 0.5) / (o * 2.0));
    if (wintype == 1) {
      an[i] = std::sqrt(c);
      sy[i] = an[i] * an[i] * an[i];
    } else if (wintype == 2) {
      an[i] = 1.0;
      sy[i] = c * c;
    } else {
      an[i] = c;
      sy[i] = c;
    }
  }
}

// Separable 1-D profile: rising ramp, flat middle, falling ramp.
inline std::vector<double> profile(int n, int o, const std::vector<double>& r, bool reversed) {
  std::vector<double> w(n, 1.0);
  for (int l = 0; l < o; l++) w[l] = reversed ? r[o - 1 - l] : r[l];
  for (int l = n - o; l < n; l++) {
    int x = l - (n - o);
    w[l] = reversed ? r[x] : r[o - 1 - x];
  }
  return w;
}
} // namespace detail

inline Tables buildTables(const Params& p) {
  using namespace detail;
  const int bw = p.bw, bh = p.bh, fbw = bw / 2 + 1;
  Tables t;
  t.winAna.resize(bw * bh);
  t.winSynA.resize(bw * bh);
  t.winSynB.resize(bw * bh);
  if (p.mode == 0) {
    // sqrt(0.5 (cos^2 x + cos^2 y)) over pixel centres (tables.cpp:8-19).
    for (int ly = 0; ly < bh; ly++) {
      double cy = std::cos((ly - bh / 2.0 + 0.5) * kPi / bh);
      for (int lx = 0; lx < bw; lx++) {
        double cx = std::cos((lx - bw / 2.0 + 0.5) * kPi / bw);
        t.winAna[ly * bw + lx] = (float)std::sqrt(0.5 * (cx * cx + cy * cy));
      }
    }
    t.winSynA = t.winAna;
    t.winSynB = t.winAna;
  } else {
    std::vector<double> anx, syx, any, syy;
    ramps(p.ow, p.wintype, anx, syx);
    ramps(p.oh, p.wintype, any, syy);
    std::vector<double> wxA = profile(bw, p.ow, anx, false);
    std::vector<double> wxSA = profile(bw, p.ow, syx, false);
    std::vector<double> wxSB = profile(bw, p.ow, syx, true);
    std::vector<double> wyA = profile(bh, p.oh, any, false);
    std::vector<double> wySA = profile(bh, p.oh, syy, false);
    t.anaX.assign(wxA.begin(), wxA.end());
    t.anaY.assign(wyA.begin(), wyA.end());
    t.synX.assign(wxSA.begin(), wxSA.end());
    t.synY.resize(bh);
    for (int ly = 0; ly < bh; ly++) t.synY[ly] = (float)(wySA[ly] / ((double)bw * bh));
    for (int ly = 0; ly < bh; ly++)
      for (int lx = 0; lx < bw; lx++) {
        t.winAna[ly * bw + lx] = (float)(wxA[lx] * wyA[ly]);
        t.winSynA[ly * bw + lx] = (float)(wxSA[lx] * wySA[ly]);
        // odd-parity synthesis: reversed-ramp profile sampled (bw-ow) to the right
        int sx = (lx + bw - p.ow) % bw;
        t.winSynB[ly * bw + lx] = (float)(wxSB[sx] * wySA[ly]);
      }
  }
  if (p.usePattern) {
    auto scale = [&](float s) {
      double v = s / 255.0;
      return (float)(v * v * p.bt * bw * bh);
    };
    const float s1 = scale(p.sigma * 255.0f), s2 = scale(p.sigma2), s3 = scale(p.sigma3),
                s4 = scale(p.sigma4);
    const float ft2 = (float)(std::sqrt(0.5) / 2.0), ft3 = (float)(std::sqrt(0.5) / 4.0);
    t.pattern.resize(fbw * bh);
    for (int h = 0; h < bh; h++) {
      float fy = (bh - 2.0f * std::fabs(h - bh / 2.0f)) / bh;
      for (int w = 0; w < fbw; w++) {
        float fx = w * 1.0f / fbw;
        float f = std::sqrt((fx * fx + fy * fy) * 0.5f);
        float v;
        if (f < ft3)
          v = s4 + (s3 - s4) * f / ft3;
        else if (f < ft2)
          v = s3 + (s2 - s3) * (f - ft3) / (ft2 - ft3);
        else
          v = s1 + (s2 - s1) * (1.0f - f) / (1.0f - ft2);
        t.pattern[h * fbw + w] = v;
      }
    }
  }
  if (p.sharpenOn) {
    t.sharpen.resize(fbw * bh);
    for (int j = 0; j < bh; j++) {
      int dj = j < bh / 2 ? j : bh - j;
      float d2v = (float)(dj * dj) * (p.svr * p.svr) / ((bh / 2.0f) * (bh / 2.0f));
      for (int i = 0; i < fbw; i++) {
        float d2 = d2v + (float)(i * i) / ((bw / 2.0f) * (bw / 2.0f));
        t.sharpen[j * fbw + i] = p.sharpen * (1.0f - std::exp(-d2 / (2.0f * p.scutoff * p.scutoff)));
      }
    }
  }
  if (p.mode == 1) {
    t.sepRow.assign(4 * bh, 0.0f);
    t.sepCol.assign(4 * fbw, 0.0f);
    std::vector<double> anx, syx, any, syy;
    ramps(p.ow, p.wintype, anx, syx);
    ramps(p.oh, p.wintype, any, syy);
    const std::vector<double> wx = profile(bw, p.ow, anx, false);
    const std::vector<double> wy = profile(bh, p.oh, any, false);
    const double sc2 = 2.0 * (double)p.scutoff * p.scutoff;
    double gx0 = 0.0, gy0 = 0.0;
    for (double v : wx) gx0 += v;
    for (double v : wy) gy0 += v;
    for (int r = 0; r < bh; r++) {
      const int dj = r < bh / 2 ? r : bh - r;
      const double d2v = (double)dj * dj * ((double)p.svr * p.svr) / ((bh / 2.0) * (bh / 2.0));
      double gr = 0.0, gi = 0.0;
      for (int y = 0; y < bh; y++) {
        const double ang = -2.0 * kPi * ((long)r * y % bh) / bh;
        gr += wy[y] * std::cos(ang);
        gi += wy[y] * std::sin(ang);
      }
      const double g = (double)p.degrid / (gx0 * gy0);
      t.sepRow[4 * r] = p.sharpenOn ? (float)(p.sharpen * std::exp(-d2v / sc2)) : 0.0f;
      t.sepRow[4 * r + 1] = (float)(gr * g);
      t.sepRow[4 * r + 2] = (float)(gi * g);
      t.sepRow[4 * r + 3] = p.sharpen;
    }
    for (int c = 0; c < fbw; c++) {
      const double d2 = (double)c * c / ((bw / 2.0) * (bw / 2.0));
      double gr = 0.0, gi = 0.0;
      for (int x = 0; x < bw; x++) {
        const double ang = -2.0 * kPi * ((long)c * x % bw) / bw;
        gr += wx[x] * std::cos(ang);
        gi += wx[x] * std::sin(ang);
      }
      t.sepCol[4 * c] = p.sharpenOn ? (float)std::exp(-d2 / sc2) : 0.0f;
      t.sepCol[4 * c + 1] = (float)gr;
      t.sepCol[4 * c + 2] = (float)gi;
    }
  }
  if (p.degridOn) {
    // gs = DFT2(winAna) / DFT2(winAna)[0,0].re * degrid, FP64 direct DFT.
    std::vector<double> cr(bw * bw), ci(bw * bw), rr(bh * bh), ri(bh * bh);
    for (int a = 0; a < bw; a++)
      for (int x = 0; x < bw; x++) {
        double ang = -2.0 * kPi * ((long)a * x % bw) / bw;
        cr[a * bw + x] = std::cos(ang);
        ci[a * bw + x] = std::sin(ang);
      }
    for (int a = 0; a < bh; a++)
      for (int y = 0; y < bh; y++) {
        double ang = -2.0 * kPi * ((long)a * y % bh) / bh;
        rr[a * bh + y] = std::cos(ang);
        ri[a * bh + y] = std::sin(ang);
      }
    // rows: R[y][c] = sum_x w[y][x] e^{-2pi i c x/bw}
    std::vector<double> Rr(bh * fbw), Ri(bh * fbw);
    for (int y = 0; y < bh; y++)
      for (int c = 0; c < fbw; c++) {
        double sr = 0, si = 0;
        for (int x = 0; x < bw; x++) {
          double w = t.winAna[y * bw + x];
          sr += w * cr[c * bw + x];
          si += w * ci[c * bw + x];
        }
        Rr[y * fbw + c] = sr;
        Ri[y * fbw + c] = si;
      }
    std::vector<double> Sr(bh * fbw), Si(bh * fbw);
    for (int r = 0; r < bh; r++)
      for (int c = 0; c < fbw; c++) {
        double sr = 0, si = 0;
        for (int y = 0; y < bh; y++) {
          double a = Rr[y * fbw + c], b = Ri[y * fbw + c];
          double e = rr[r * bh + y], f = ri[r * bh + y];
          sr += a * e - b * f;
          si += a * f + b * e;
        }
        Sr[r * fbw + c] = sr;
        Si[r * fbw + c] = si;
      }
    const double dc = Sr[0];
    t.degrid.resize(2 * fbw * bh);
    for (int i = 0; i < fbw * bh; i++) {
      t.degrid[2 * i] = (float)(Sr[i] / dc * p.degrid);
      t.degrid[2 * i + 1] = (float)(Si[i] / dc * p.degrid);
    }
  }
  return t;
}

// ----- constant operands of the MFMA DFT form (fk_mfma.h), host side -----

// float -> IEEE binary16 / bfloat16 bits, round to nearest even
inline uint16_t to_f16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ex = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  const int e = (int)ex - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}
inline float from_f16_bits(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-8f;  // m * 2^-24
    std::memcpy(&x, &v, 4);
    x |= sign;
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e - 15u + 127u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}
inline uint16_t to_bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float from_bf16_bits(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// The 24 operand slots x 64 lanes x 8 values of the 32-point DFT matrices for
// the MFMA form (slot layout and lane layout: fk_mfma.h), each entry split
// into hi = round(c) and lo = round(c - hi) of operand type prec (1 binary16,
// 2 bfloat16), from matrices evaluated in double.
inline std::vector<uint16_t> dft_operands32(int prec) {
  const double kPi = 3.14159265358979323846;
  auto cs = [&](long long a) { return std::cos(2.0 * kPi * (double)(a % 32) / 32.0); };
  auto sn = [&](long long a) { return std::sin(2.0 * kPi * (double)(a % 32) / 32.0); };
  // M[k][m]: real row DFT (column 2p: cos, 2p+1: -sin, column 1: bin 16)
  auto M = [&](int k, int m) {
    if (m == 1) return (k & 1) ? -1.0 : 1.0;
    const int p = m / 2;
    return (m & 1) ? -sn((long long)k * p) : cs((long long)k * p);
  };
  // N[m][k]: real row inverse (row 0: 1, row 1: (-1)^k, 2p: 2cos, 2p+1: -2sin)
  auto N = [&](int m, int k) {
    if (m == 0) return 1.0;
    if (m == 1) return (k & 1) ? -1.0 : 1.0;
    const int p = m / 2;
    return (m & 1) ? -2.0 * sn((long long)p * k) : 2.0 * cs((long long)p * k);
  };
  auto rho = [](int c, int h, int j) { return 16 * c + 8 * (j >> 2) + 4 * h + (j & 3); };
  std::vector<uint16_t> out(24 * 64 * 8);
  auto put = [&](int slot0, int c, int lane, int j, double v) {
    const float hv = (float)v;
    uint16_t hi, lo;
    if (prec == 1) {
      hi = to_f16_bits(hv);
      lo = to_f16_bits((float)(v - (double)from_f16_bits(hi)));
    } else {
      hi = to_bf16_bits(hv);
      lo = to_bf16_bits((float)(v - (double)from_bf16_bits(hi)));
    }
    out[((size_t)(slot0 + 2 * c) * 64 + lane) * 8 + j] = hi;
    out[((size_t)(slot0 + 2 * c + 1) * 64 + lane) * 8 + j] = lo;
  };
  for (int c = 0; c < 2; c++)
    for (int lane = 0; lane < 64; lane++) {
      const int h = lane / 32, i = lane % 32;
      for (int j = 0; j < 8; j++) {
        const int kn = 16 * c + 8 * h + j, kr = rho(c, h, j);
        put(0, c, lane, j, M(kn, i));                          // B: M[k][m = i]
        put(4, c, lane, j, cs((long long)i * kr));             // B: C[v = i][rho]
        put(8, c, lane, j, sn((long long)i * kr));             // B: S[v = i][rho]
        put(12, c, lane, j, cs((long long)kn * i));            // B: C[v = kn][r = i]
        put(16, c, lane, j, sn((long long)kn * i));            // B: S[v = kn][r = i]
        put(20, c, lane, j, N(kr, i));                         // A: N[rho][k = i]
      }
    }
  return out;
}

} // namespace fft3d
