/*
 * FFT3D denoise oracle (test infrastructure only): plain-C restatement of the
 * reference's CPU backend. See fft3d_oracle.h for scope and the rule that the
 * product never uses this file.
 *
 * Every function cites the reference file:line it follows. Floating-point
 * expressions keep the reference's operand order so that the FP32 build is
 * bit-identical to the reference compiled with g++ (checked in
 * tests/test_oracle_ref.py against oracle/_ref when it is built).
 */
#include "fft3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef orc_real real;

/* sqrt in the arithmetic type (sqrtf for the FP32 build, as std::sqrt(float)). */
#define RSQRT_(x) (sizeof(real) == 4 ? (real)sqrtf((float)(x)) : (real)sqrt((double)(x)))

int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------ params */

/* src/fft3dcl.cpp:171-195: defaults when the script passes no keys. */
void orc_script_defaults(orc_script* sp) {
  memset(sp, 0, sizeof(*sp));
  sp->sigma = 2.0f;
  sp->beta = 1.0f;
  sp->bw = 32;
  sp->bh = 32;
  sp->bt = 3;
  sp->sharpen = 0.0f;
  sp->plane = 0;
  sp->mode = 1;
  sp->bordersize = 1;
  sp->precision = 0;
  sp->degrid = 1.0f; /* mode==1 ? 1 : 0 */
  sp->scutoff = 0.3f;
  sp->svr = 1.0f;
  sp->smin = 4.0f;
  sp->smax = 20.0f;
  sp->kratio = 2.0f;
  sp->ow = 16;
  sp->oh = 16;
  sp->wintype = 0;
  sp->interlaced = 0;
  sp->sigma2 = sp->sigma3 = sp->sigma4 = 2.0f;
  sp->oldfft = -1;
}

/* src/params.h:62-72: round down to a power of two, clamp to [4,512]. */
int orc_round_pow2_clamp(int v) {
  unsigned int logn = 0;
  for (unsigned int i = (unsigned int)(v - 1); i > 0; i /= 2, logn++)
    ;
  int r = 1;
  for (unsigned int i = 1; i <= logn; i++) r *= 2;
  if (r < 4) r = 4;
  if (r > 512) r = 512;
  return r;
}

/* src/params.h:76-148. */
int orc_validate(const orc_script* sp, int width, int height, int bits, orc_internal* out,
                 const char** err) {
  memset(out, 0, sizeof(*out));
  out->sigma = sp->sigma / 255.0f;
  out->beta = sp->beta;
  out->bm = sp->bt;
  if (out->bm < -1 || out->bm > 4) {
    *err = "FFT3DCL: valid modes for bt are -1,0,1,2,3,4";
    return 0;
  }
  out->bt = sp->bt > 0 ? sp->bt : 1;
  out->sharpen = sp->sharpen;
  out->sharpenOn = sp->sharpen != 0.0f;
  out->mode = (sp->mode < 0 || sp->mode > 2) ? 0 : sp->mode;
  out->bordersize = sp->bordersize;
  out->degrid = sp->degrid;
  out->degridOn = sp->degrid != 0.0f;
  out->scutoff = sp->scutoff;
  out->svr = sp->svr;
  out->smin = sp->smin / 255.0f;
  out->smax = sp->smax / 255.0f;
  out->kratio = sp->kratio;
  out->wintype = (sp->wintype < 0 || sp->wintype > 2) ? 0 : sp->wintype;
  out->interlaced = sp->interlaced;
  out->sigma2 = sp->sigma2;
  out->sigma3 = sp->sigma3;
  out->sigma4 = sp->sigma4;
  out->usePattern = (sp->sigma != sp->sigma2 || sp->sigma3 != sp->sigma4 || sp->sigma != sp->sigma3);
  out->oldfft = sp->oldfft;
  out->current = (out->bm == 1) ? 0 : (out->bm == 4) ? 2 : 1;

  out->bw = orc_round_pow2_clamp(sp->bw);
  out->bh = orc_round_pow2_clamp(sp->bh);
  out->ow = (int)(sp->ow * (float)out->bw / (float)sp->bw);
  out->ow &= (int)0xFFFFFFFE;
  out->oh = (int)(sp->oh * (float)out->bh / (float)sp->bh);
  if (out->oh * 2 > out->bh) out->oh = out->bh / 2;
  if (out->ow * 2 > out->bw) out->ow = out->bw / 2;

  int precision = sp->precision;
  if (bits > 8) precision = 2;
  if (precision < 0 || precision > 2) precision = 1;
  out->precision = precision;

  if (out->mode == 2) {
    *err = "FFT3DCL: mode 2 is not implemented yet in this port (use mode 0 or 1)";
    return 0;
  }
  if (out->interlaced) {
    *err = "FFT3DCL: interlaced is not implemented yet in this port";
    return 0;
  }
  if (out->mode == 0 || (out->bw == out->ow * 2 && out->bh == out->oh * 2)) {
    out->ny = (int)((height + out->bh * 1.5 - 1) / out->bh);
    out->nx = (int)((width + out->bw * 1.5 - 1) / out->bw);
  } else {
    out->nx = ((width + out->ow - 1) / (out->bw - out->ow) + 1 + 1) / 2;
    out->ny = ((height + out->oh - 1) / (out->bh - out->oh) + 1 + 1) / 2;
  }
  out->totw = out->nx * out->bw;
  out->toth = out->ny * out->bh;

  out->wienerS = out->sigma * out->sigma * out->bt * out->bw * out->bh;
  out->betaX = (out->beta - 1.0f) / out->beta;
  out->kalmanCN = out->sigma * out->sigma * out->bw * out->bh;
  out->ssmin = out->smin * out->smin * out->bw * out->bh * out->bt;
  out->ssmax = out->smax * out->smax * out->bw * out->bh * out->bt;
  return 1;
}

/* ------------------------------------------------------------------ tables */

static const double kPi = 3.14159265358979323846;

/* src/cpu/tables.cpp:8-19 */
static void build_window_mode0(int bw, int bh, float* w) {
  int off = 0;
  for (double n1 = -bh / 2.0 + 0.5; n1 < bh / 2.0; n1++) {
    double cosy = cos(n1 * kPi / bh);
    for (double n2 = -bw / 2.0 + 0.5; n2 < bw / 2.0; n2++) {
      double cosx = cos(n2 * kPi / bw);
      w[off++] = (float)sqrt(0.5 * (cosx * cosx + cosy * cosy));
    }
  }
}

/* src/cpu/tables.cpp:21-105 */
static void build_windows_mode1(int bw, int bh, int ow, int oh, int wintype, float* A,
                                float* SynA, float* SynB) {
  double* wanx = (double*)calloc(ow > 0 ? ow : 1, sizeof(double));
  double* wany = (double*)calloc(oh > 0 ? oh : 1, sizeof(double));
  double* wsynx = (double*)calloc(ow > 0 ? ow : 1, sizeof(double));
  double* wsyny = (double*)calloc(oh > 0 ? oh : 1, sizeof(double));
  switch (wintype) {
  case 1:
    for (int i = 0; i < ow; i++) {
      wanx[i] = sqrt(cos(kPi * (i - ow + 0.5) / (ow * 2.0)));
      wsynx[i] = wanx[i] * wanx[i] * wanx[i];
    }
    for (int i = 0; i < oh; i++) {
      wany[i] = sqrt(cos(kPi * (i - oh + 0.5) / (oh * 2.0)));
      wsyny[i] = wany[i] * wany[i] * wany[i];
    }
    break;
  case 2:
    for (int i = 0; i < ow; i++) {
      wanx[i] = 1.0;
      wsynx[i] = cos(kPi * (i - ow + 0.5) / (ow * 2.0));
      wsynx[i] *= wsynx[i];
    }
    for (int i = 0; i < oh; i++) {
      wany[i] = 1.0;
      wsyny[i] = cos(kPi * (i - oh + 0.5) / (oh * 2.0));
      wsyny[i] *= wsyny[i];
    }
    break;
  default:
    for (int i = 0; i < ow; i++) {
      wanx[i] = cos(kPi * (i - ow + 0.5) / (ow * 2.0));
      wsynx[i] = wanx[i];
    }
    for (int i = 0; i < oh; i++) {
      wany[i] = cos(kPi * (i - oh + 0.5) / (oh * 2.0));
      wsyny[i] = wany[i];
    }
    break;
  }
  double* wxA = (double*)calloc(bw, sizeof(double));
  double* wxSA = (double*)calloc(bw, sizeof(double));
  double* wxSB = (double*)calloc(bw, sizeof(double));
  double* wyA = (double*)calloc(bh, sizeof(double));
  double* wySA = (double*)calloc(bh, sizeof(double));
  for (int lx = 0; lx < ow; lx++) {
    wxA[lx] = wanx[lx];
    wxSA[lx] = wsynx[lx];
    wxSB[lx] = wsynx[(ow - 1) - lx];
  }
  for (int lx = ow; lx < bw - ow; lx++) {
    wxA[lx] = 1.0;
    wxSA[lx] = 1.0;
    wxSB[lx] = 1.0;
  }
  for (int lx = bw - ow; lx < bw; lx++) {
    int x = lx - (bw - ow);
    wxA[lx] = wanx[(ow - 1) - x];
    wxSA[lx] = wsynx[(ow - 1) - x];
    wxSB[lx] = wsynx[x];
  }
  for (int ly = 0; ly < oh; ly++) {
    wyA[ly] = wany[ly];
    wySA[ly] = wsyny[ly];
  }
  for (int ly = oh; ly < bh - oh; ly++) {
    wyA[ly] = 1.0;
    wySA[ly] = 1.0;
  }
  for (int ly = bh - oh; ly < bh; ly++) {
    wyA[ly] = wany[bh - 1 - ly];
    wySA[ly] = wsyny[bh - 1 - ly];
  }
  for (int ly = 0; ly < bh; ly++)
    for (int lx = 0; lx < bw; lx++) {
      A[ly * bw + lx] = (float)(wxA[lx] * wyA[ly]);
      SynA[ly * bw + lx] = (float)(wxSA[lx] * wySA[ly]);
      SynB[ly * bw + lx] = (float)(wxSB[lx] * wySA[ly]);
    }
  free(wanx); free(wany); free(wsynx); free(wsyny);
  free(wxA); free(wxSA); free(wxSB); free(wyA); free(wySA);
}

static float pattern_scale(float s, int bt, int bw, int bh) {
  double v = s / 255.0;
  return (float)(v * v * bt * bw * bh);
}

/* src/cpu/tables.cpp:107-137 (caller checks usePattern). */
static void build_pattern(int bw, int bh, int bt, float sigma, float sigma2, float sigma3,
                          float sigma4, float* pat) {
  const int fbw = bw / 2 + 1;
  float s = pattern_scale(sigma, bt, bw, bh), s2 = pattern_scale(sigma2, bt, bw, bh);
  float s3 = pattern_scale(sigma3, bt, bw, bh), s4 = pattern_scale(sigma4, bt, bw, bh);
  float ft2 = (float)(sqrt(0.5) / 2.0);
  float ft3 = (float)(sqrt(0.5) / 4.0);
  for (int h = 0; h < bh; h++) {
    float fy = (bh - 2.0f * fabsf(h - bh / 2.0f)) / bh;
    for (int w = 0; w < fbw; w++) {
      float fx = w * 1.0f / fbw;
      float f = sqrtf((fx * fx + fy * fy) * 0.5f);
      float p;
      if (f < ft3)
        p = s4 + (s3 - s4) * f / ft3;
      else if (f < ft2)
        p = s3 + (s2 - s3) * (f - ft3) / (ft2 - ft3);
      else
        p = s + (s2 - s) * (1.0f - f) / (1.0f - ft2);
      pat[h * fbw + w] = p;
    }
  }
}

/* src/cpu/tables.cpp:139-156 (caller checks strength != 0). */
static void build_sharpen_lut(int bw, int bh, float strength, float svr, float scutoff,
                              float* lut) {
  const int fbw = bw / 2 + 1;
  for (int j = 0; j < bh; j++) {
    int dj = j;
    if (j >= bh / 2) dj = bh - j;
    float d2v = (float)(dj * dj) * (svr * svr) / ((bh / 2.0f) * (bh / 2.0f));
    for (int i = 0; i < fbw; i++) {
      float d2 = d2v + (float)(i * i) / ((bw / 2.0f) * (bw / 2.0f));
      lut[j * fbw + i] = strength * (1.0f - expf(-d2 / (2.0f * scutoff * scutoff)));
    }
  }
}

/* --------------------------------------------------------------------- FFT */

/* src/cpu/fft.cpp:8-25: twiddles exp(-2 pi i k/n) in double, stored in the
 * arithmetic type; bit-reversal table. */
typedef struct {
  int n;
  real* twr;
  real* twi;
  int* rev;
} fft_plan;

static void plan_init(fft_plan* p, int n) {
  p->n = n;
  p->twr = (real*)malloc(sizeof(real) * (n / 2 > 0 ? n / 2 : 1));
  p->twi = (real*)malloc(sizeof(real) * (n / 2 > 0 ? n / 2 : 1));
  for (int k = 0; k < n / 2; k++) {
    double a = -2.0 * 3.14159265358979323846 * k / n;
    p->twr[k] = (real)cos(a);
    p->twi[k] = (real)sin(a);
  }
  p->rev = (int*)malloc(sizeof(int) * n);
  int bits = 0;
  while ((1 << bits) < n) bits++;
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int b = 0; b < bits; b++)
      if (i & (1 << b)) r |= 1 << (bits - 1 - b);
    p->rev[i] = r;
  }
}

static void plan_free(fft_plan* p) {
  free(p->twr);
  free(p->twi);
  free(p->rev);
}

/* src/cpu/fft.cpp:28-58: iterative radix-2 DIT. Complex product w*b is
 * (wr*br - wi*bi, wr*bi + wi*br), the expansion g++ emits for std::complex.
 * in/out interleaved (re, im); out may alias in. */
static void transform(const fft_plan* p, const real* in, real* out, int inverse, real* a) {
  const int n = p->n;
  for (int i = 0; i < n; i++) {
    a[2 * i] = in[2 * p->rev[i]];
    a[2 * i + 1] = in[2 * p->rev[i] + 1];
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    const int step = n / len;
    for (int i = 0; i < n; i += len)
      for (int j = 0; j < half; j++) {
        real wr = p->twr[j * step];
        real wi = p->twi[j * step];
        if (inverse) wi = -wi;
        real br = a[2 * (i + j + half)], bi = a[2 * (i + j + half) + 1];
        real tr = wr * br - wi * bi;
        real ti = wr * bi + wi * br;
        real ur = a[2 * (i + j)], ui = a[2 * (i + j) + 1];
        a[2 * (i + j)] = ur + tr;
        a[2 * (i + j) + 1] = ui + ti;
        a[2 * (i + j + half)] = ur - tr;
        a[2 * (i + j + half) + 1] = ui - ti;
      }
  }
  if (inverse) {
    const real invn = (real)1.0f / n;
    for (int i = 0; i < 2 * n; i++) out[i] = a[i] * invn;
  } else {
    for (int i = 0; i < 2 * n; i++) out[i] = a[i];
  }
}

/* src/cpu/fft.cpp:68-86: rows then columns, full complex spectrum. */
static void fft2d_real_forward(const real* block, int bw, int bh, const fft_plan* rp,
                               const fft_plan* cp, real* full, real* scratch) {
  int m = bw > bh ? bw : bh;
  real* row = scratch;           /* 2*m */
  real* col = scratch + 2 * m;   /* 2*m */
  real* colout = col + 2 * m;    /* 2*m */
  real* a = colout + 2 * m;      /* 2*m */
  for (int r = 0; r < bh; r++) {
    for (int c = 0; c < bw; c++) {
      row[2 * c] = block[r * bw + c];
      row[2 * c + 1] = 0;
    }
    transform(rp, row, full + 2 * r * bw, 0, a);
  }
  for (int c = 0; c < bw; c++) {
    for (int r = 0; r < bh; r++) {
      col[2 * r] = full[2 * (r * bw + c)];
      col[2 * r + 1] = full[2 * (r * bw + c) + 1];
    }
    transform(cp, col, colout, 0, a);
    for (int r = 0; r < bh; r++) {
      full[2 * (r * bw + c)] = colout[2 * r];
      full[2 * (r * bw + c) + 1] = colout[2 * r + 1];
    }
  }
}

/* src/cpu/fft.cpp:88-119: Hermitian rebuild, inverse columns then rows, real
 * part kept. half is fbw x bh interleaved complex; full is scratch bw*bh. */
static void fft2d_real_inverse_half(const real* half, int bw, int bh, int fbw,
                                    const fft_plan* rp, const fft_plan* cp, real* block,
                                    real* full, real* scratch) {
  int m = bw > bh ? bw : bh;
  real* col = scratch;
  real* colout = col + 2 * m;
  real* rowout = colout + 2 * m;
  real* a = rowout + 2 * m;
  for (int r = 0; r < bh; r++)
    for (int c = 0; c < fbw; c++) {
      full[2 * (r * bw + c)] = half[2 * (r * fbw + c)];
      full[2 * (r * bw + c) + 1] = half[2 * (r * fbw + c) + 1];
    }
  for (int r = 0; r < bh; r++) {
    int rm = (bh - r) % bh;
    for (int c = fbw; c < bw; c++) {
      int cm = bw - c;
      full[2 * (r * bw + c)] = full[2 * (rm * bw + cm)];
      full[2 * (r * bw + c) + 1] = -full[2 * (rm * bw + cm) + 1];
    }
  }
  for (int c = 0; c < bw; c++) {
    for (int r = 0; r < bh; r++) {
      col[2 * r] = full[2 * (r * bw + c)];
      col[2 * r + 1] = full[2 * (r * bw + c) + 1];
    }
    transform(cp, col, colout, 1, a);
    for (int r = 0; r < bh; r++) {
      full[2 * (r * bw + c)] = colout[2 * r];
      full[2 * (r * bw + c) + 1] = colout[2 * r + 1];
    }
  }
  for (int r = 0; r < bh; r++) {
    transform(rp, full + 2 * r * bw, rowout, 1, a);
    for (int c = 0; c < bw; c++) block[r * bw + c] = rowout[2 * c];
  }
}

void orc_fft1d(int n, const real* in, real* out, int inverse) {
  fft_plan p;
  plan_init(&p, n);
  real* a = (real*)malloc(sizeof(real) * 2 * n);
  transform(&p, in, out, inverse, a);
  free(a);
  plan_free(&p);
}

void orc_fft2d_real_forward(const real* block, int bw, int bh, real* full) {
  fft_plan rp, cp;
  plan_init(&rp, bw);
  plan_init(&cp, bh);
  int m = bw > bh ? bw : bh;
  real* s = (real*)malloc(sizeof(real) * 8 * m);
  fft2d_real_forward(block, bw, bh, &rp, &cp, full, s);
  free(s);
  plan_free(&rp);
  plan_free(&cp);
}

void orc_fft2d_real_inverse_half(const real* half, int bw, int bh, real* block) {
  fft_plan rp, cp;
  plan_init(&rp, bw);
  plan_init(&cp, bh);
  int m = bw > bh ? bw : bh;
  real* s = (real*)malloc(sizeof(real) * 8 * m);
  real* full = (real*)malloc(sizeof(real) * 2 * bw * bh);
  fft2d_real_inverse_half(half, bw, bh, bw / 2 + 1, &rp, &cp, block, full, s);
  free(full);
  free(s);
  plan_free(&rp);
  plan_free(&cp);
}

/* ------------------------------------------------------------------- plane */

struct orc_plane {
  orc_internal p;
  int width, height, bits, sample_type, chroma;
  int fbw, nb, ncols;
  int* bx;
  int* by;
  uint8_t* par;
  float* winAna;
  float* winSynA;
  float* winSynB;
  float* pattern; /* NULL when off */
  float* sharpen; /* NULL when off */
  float chromaNorm;
  fft_plan rowPlan, colPlan;
  /* degrid: one planeSize copy (identical for every block) */
  float* gsRe;
  float* gsIm;
  int degridReady;
  /* Kalman state */
  real *kLastRe, *kLastIm, *kCovar, *kProc;
  int kalmanFirst;
  /* scratch */
  real* blockReal;
  real* fullSpec;
  real* fftScratch;
  real* fullTmp;
  real *filtRe, *filtIm, *corrRe, *corrIm;
};

/* src/cpu/reference.cpp:10-17 (Direct3D MIRROR addressing). */
static int mirror_i(int x, int w) {
  int p = x % (2 * w);
  if (p < 0) p += 2 * w;
  if (p >= w) p = 2 * w - 1 - p;
  return p;
}

/* src/cpu/reference.cpp:19-80 */
orc_plane* orc_plane_new(const orc_internal* p, int width, int height, int bits, int sample_type,
                         int chroma) {
  orc_plane* pl = (orc_plane*)calloc(1, sizeof(orc_plane));
  pl->p = *p;
  pl->width = width;
  pl->height = height;
  pl->bits = bits;
  pl->sample_type = sample_type;
  pl->chroma = chroma;
  const int bw = p->bw, bh = p->bh;
  pl->fbw = bw / 2 + 1;
  plan_init(&pl->rowPlan, bw);
  plan_init(&pl->colPlan, bh);
  pl->winAna = (float*)calloc(bw * bh, sizeof(float));
  pl->winSynA = (float*)calloc(bw * bh, sizeof(float));
  pl->winSynB = (float*)calloc(bw * bh, sizeof(float));
  if (p->mode == 0) {
    pl->nb = 2 * p->nx * p->ny;
    pl->bx = (int*)malloc(sizeof(int) * pl->nb);
    pl->by = (int*)malloc(sizeof(int) * pl->nb);
    pl->par = (uint8_t*)calloc(pl->nb, 1);
    int k = 0;
    for (int grid = 0; grid < 2; grid++)
      for (int j = 0; j < p->ny; j++)
        for (int i = 0; i < p->nx; i++) {
          int off = grid ? -bw / 2 : 0, offy = grid ? -bh / 2 : 0;
          pl->bx[k] = i * bw + off;
          pl->by[k] = j * bh + offy;
          k++;
        }
    pl->ncols = 0;
    build_window_mode0(bw, bh, pl->winAna);
  } else {
    int ncols = 2 * p->nx, nrows = 2 * p->ny;
    pl->ncols = ncols;
    pl->nb = ncols * nrows;
    pl->bx = (int*)malloc(sizeof(int) * pl->nb);
    pl->by = (int*)malloc(sizeof(int) * pl->nb);
    pl->par = (uint8_t*)calloc(pl->nb, 1);
    int ystep = bh - p->oh, k = 0;
    for (int m = 0; m < nrows; m++)
      for (int n = 0; n < ncols; n++) {
        pl->bx[k] = -p->ow + n * (bw - p->ow);
        pl->by[k] = -p->oh + m * ystep;
        pl->par[k] = (uint8_t)(n & 1);
        k++;
      }
    float* synB = (float*)calloc(bw * bh, sizeof(float));
    build_windows_mode1(bw, bh, p->ow, p->oh, p->wintype, pl->winAna, pl->winSynA, synB);
    /* reference.cpp:54-57: SynB sampled (bw-ow) to the right. */
    for (int ly = 0; ly < bh; ly++)
      for (int lx = 0; lx < bw; lx++)
        pl->winSynB[ly * bw + lx] = synB[ly * bw + ((lx + bw - p->ow) % bw)];
    free(synB);
  }
  const int ps = pl->fbw * bh;
  if (p->usePattern) {
    pl->pattern = (float*)malloc(sizeof(float) * ps);
    build_pattern(bw, bh, p->bt, p->sigma * 255.0f, p->sigma2, p->sigma3, p->sigma4, pl->pattern);
  }
  if (p->sharpen != 0.0f) {
    pl->sharpen = (float*)malloc(sizeof(float) * ps);
    build_sharpen_lut(bw, bh, p->sharpen, p->svr, p->scutoff, pl->sharpen);
  }
  /* reference.cpp:70-79: CHROMANORM */
  if (p->mode == 1 && chroma) {
    if (sample_type == 1)
      pl->chromaNorm = 0.0f;
    else if (bits <= 8)
      pl->chromaNorm = 128.0f / 255.0f;
    else
      pl->chromaNorm = 32768.0f / 65535.0f;
  }
  int m = bw > bh ? bw : bh;
  pl->blockReal = (real*)malloc(sizeof(real) * bw * bh);
  pl->fullSpec = (real*)malloc(sizeof(real) * 2 * bw * bh);
  pl->fullTmp = (real*)malloc(sizeof(real) * 2 * bw * bh);
  pl->fftScratch = (real*)malloc(sizeof(real) * 8 * m);
  pl->kalmanFirst = 1;
  return pl;
}

void orc_plane_free(orc_plane* pl) {
  if (!pl) return;
  free(pl->bx); free(pl->by); free(pl->par);
  free(pl->winAna); free(pl->winSynA); free(pl->winSynB);
  free(pl->pattern); free(pl->sharpen);
  free(pl->gsRe); free(pl->gsIm);
  free(pl->kLastRe); free(pl->kLastIm); free(pl->kCovar); free(pl->kProc);
  free(pl->blockReal); free(pl->fullSpec); free(pl->fullTmp); free(pl->fftScratch);
  free(pl->filtRe); free(pl->filtIm); free(pl->corrRe); free(pl->corrIm);
  plan_free(&pl->rowPlan);
  plan_free(&pl->colPlan);
  free(pl);
}

int orc_plane_nblocks(const orc_plane* pl) { return pl->nb; }
int orc_plane_fbw(const orc_plane* pl) { return pl->fbw; }
int orc_plane_ncols(const orc_plane* pl) { return pl->ncols; }
float orc_plane_chroma_norm(const orc_plane* pl) { return pl->chromaNorm; }
int orc_plane_bx(const orc_plane* pl, int* d) { memcpy(d, pl->bx, sizeof(int) * pl->nb); return pl->nb; }
int orc_plane_by(const orc_plane* pl, int* d) { memcpy(d, pl->by, sizeof(int) * pl->nb); return pl->nb; }
int orc_plane_par(const orc_plane* pl, uint8_t* d) { memcpy(d, pl->par, pl->nb); return pl->nb; }
int orc_plane_win(const orc_plane* pl, int which, float* d) {
  const float* s = which == 0 ? pl->winAna : which == 1 ? pl->winSynA : pl->winSynB;
  memcpy(d, s, sizeof(float) * pl->p.bw * pl->p.bh);
  return pl->p.bw * pl->p.bh;
}
int orc_plane_pattern(const orc_plane* pl, float* d) {
  if (!pl->pattern) return 0;
  memcpy(d, pl->pattern, sizeof(float) * pl->fbw * pl->p.bh);
  return pl->fbw * pl->p.bh;
}
int orc_plane_sharpen(const orc_plane* pl, float* d) {
  if (!pl->sharpen) return 0;
  memcpy(d, pl->sharpen, sizeof(float) * pl->fbw * pl->p.bh);
  return pl->fbw * pl->p.bh;
}
int orc_plane_degrid(const orc_plane* pl, float* d) {
  if (!pl->degridReady) return 0;
  int ps = pl->fbw * pl->p.bh;
  for (int i = 0; i < ps; i++) {
    d[2 * i] = pl->gsRe[i];
    d[2 * i + 1] = pl->gsIm[i];
  }
  return ps;
}

/* src/cpu/reference.cpp:82-105 */
void orc_upload(const orc_plane* pl, const void* src, ptrdiff_t pitch, float* dst) {
  const int W = pl->width, H = pl->height;
  const uint8_t* s = (const uint8_t*)src;
  if (pl->sample_type == 1) {
    for (int y = 0; y < H; y++) {
      const float* row = (const float*)(s + y * pitch);
      for (int x = 0; x < W; x++) dst[y * W + x] = row[x];
    }
  } else if (pl->bits <= 8) {
    for (int y = 0; y < H; y++) {
      const uint8_t* row = s + y * pitch;
      for (int x = 0; x < W; x++) dst[y * W + x] = row[x] * (1.0f / 255.0f);
    }
  } else {
    const float norm = (float)(1u << (16 - pl->bits)) / 65535.0f;
    for (int y = 0; y < H; y++) {
      const uint16_t* row = (const uint16_t*)(s + y * pitch);
      for (int x = 0; x < W; x++) dst[y * W + x] = row[x] * norm;
    }
  }
}

/* src/cpu/reference.cpp:107-138 */
void orc_download(const orc_plane* pl, const float* src, void* dst, ptrdiff_t pitch) {
  const int W = pl->width, H = pl->height;
  uint8_t* d = (uint8_t*)dst;
  if (pl->sample_type == 1) {
    for (int y = 0; y < H; y++) {
      float* row = (float*)(d + y * pitch);
      for (int x = 0; x < W; x++) row[x] = src[y * W + x];
    }
  } else if (pl->bits <= 8) {
    for (int y = 0; y < H; y++) {
      uint8_t* row = d + y * pitch;
      for (int x = 0; x < W; x++) {
        float v = src[y * W + x];
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        row[x] = (uint8_t)(v * 255.0f + 0.5f);
      }
    }
  } else {
    const float invNorm = 1.0f / (float)(1u << (16 - pl->bits));
    const float maxPix = (float)((1u << pl->bits) - 1);
    for (int y = 0; y < H; y++) {
      uint16_t* row = (uint16_t*)(d + y * pitch);
      for (int x = 0; x < W; x++) {
        float v = src[y * W + x] * invNorm;
        v = v < maxPix ? v : maxPix;
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        row[x] = (uint16_t)(v * 65535.0f + 0.5f);
      }
    }
  }
}

/* src/cpu/reference.cpp:140-170 */
void orc_compute_spectra(orc_plane* pl, const float* plane, real* re, real* im) {
  const int nb = pl->nb, fbw = pl->fbw, bw = pl->p.bw, bh = pl->p.bh;
  const int W = pl->width, H = pl->height;
  const int center = pl->chromaNorm != 0.0f;
  for (int b = 0; b < nb; b++) {
    const int x0 = pl->bx[b], y0 = pl->by[b];
    for (int ly = 0; ly < bh; ly++) {
      const float* prow = &plane[mirror_i(y0 + ly, H) * W];
      const float* wrow = &pl->winAna[ly * bw];
      real* brow = &pl->blockReal[ly * bw];
      for (int lx = 0; lx < bw; lx++) {
        float v = prow[mirror_i(x0 + lx, W)];
        if (center) v -= pl->chromaNorm;
        brow[lx] = (real)v * wrow[lx];
      }
    }
    fft2d_real_forward(pl->blockReal, bw, bh, &pl->rowPlan, &pl->colPlan, pl->fullSpec,
                       pl->fftScratch);
    real* dRe = &re[(size_t)b * fbw * bh];
    real* dIm = &im[(size_t)b * fbw * bh];
    for (int r = 0; r < bh; r++)
      for (int c = 0; c < fbw; c++) {
        dRe[r * fbw + c] = pl->fullSpec[2 * (r * bw + c)];
        dIm[r * fbw + c] = pl->fullSpec[2 * (r * bw + c) + 1];
      }
  }
}

/* src/cpu/reference.cpp:172-198. The sample is identical for every block, so
 * one planeSize copy is kept (the reference replicates it per block). */
void orc_build_degrid(orc_plane* pl) {
  const int fbw = pl->fbw, bw = pl->p.bw, bh = pl->p.bh, ps = fbw * bh;
  for (int i = 0; i < bw * bh; i++) pl->blockReal[i] = (real)pl->winAna[i];
  fft2d_real_forward(pl->blockReal, bw, bh, &pl->rowPlan, &pl->colPlan, pl->fullSpec,
                     pl->fftScratch);
  free(pl->gsRe);
  free(pl->gsIm);
  pl->gsRe = (float*)malloc(sizeof(float) * ps);
  pl->gsIm = (float*)malloc(sizeof(float) * ps);
  const real dc = pl->fullSpec[0];
  for (int r = 0; r < bh; r++)
    for (int c = 0; c < fbw; c++) {
      real zr = pl->fullSpec[2 * (r * bw + c)], zi = pl->fullSpec[2 * (r * bw + c) + 1];
      pl->gsRe[r * fbw + c] = (float)(zr / dc * pl->p.degrid);
      pl->gsIm[r * fbw + c] = (float)(zi / dc * pl->p.degrid);
    }
  pl->degridReady = 1;
}

/* src/cpu/reference.cpp:201-209 (Wiener, epsilon 1e-37f). */
static inline void wienner(real* re, real* im, real sigmaX, real sigmaY, real betaX) {
  real psd = *re * *re + *im * *im + (real)0.0000000000000000000000000000000000001f;
  real psdInv = (real)1.0f / psd;
  real mulFac = betaX;
  if ((real)1.0f / sigmaX > psdInv) mulFac = (real)1.0f - sigmaY * psdInv;
  *re *= mulFac;
  *im *= mulFac;
}

/* src/cpu/reference.cpp:211-437 */
void orc_filter_synth(orc_plane* pl, const real* const* wre, const real* const* wim, int nwin,
                      real* out) {
  const orc_internal* p = &pl->p;
  const int nb = pl->nb, fbw = pl->fbw, bw = p->bw, bh = p->bh;
  const int W = pl->width, H = pl->height;
  const int ps = fbw * bh;
  const size_t total = (size_t)nb * ps;
  (void)nwin;
  for (int i = 0; i < W * H; i++) out[i] = 0;
  if (!pl->filtRe) {
    pl->filtRe = (real*)malloc(sizeof(real) * total);
    pl->filtIm = (real*)malloc(sizeof(real) * total);
    pl->corrRe = (real*)malloc(sizeof(real) * total);
    pl->corrIm = (real*)malloc(sizeof(real) * total);
  }
  real *filtRe = pl->filtRe, *filtIm = pl->filtIm, *corrRe = pl->corrRe, *corrIm = pl->corrIm;
  const int ci = p->bm >= 2 ? p->current : 0;
  const real* curRe = wre[ci];
  const real* curIm = wim[ci];

  /* reference.cpp:229-239: FImgDegrid = current DC * gridsample. */
  const int dg = p->degridOn && pl->degridReady;
  if (dg) {
    for (int b = 0; b < nb; b++) {
      const real dc = curRe[(size_t)b * ps];
      for (int i = 0; i < ps; i++) {
        corrRe[(size_t)b * ps + i] = dc * pl->gsRe[i];
        corrIm[(size_t)b * ps + i] = dc * pl->gsIm[i];
      }
    }
  }
  const real sigmaX = p->beta * p->wienerS;
  const real sigmaY = p->wienerS;
  const real betaX = p->betaX;
  const int pat = pl->pattern != NULL;

  if (p->bm >= 2) {
    const real K = (real)0.8660254f; /* reference.cpp:247 (truncated) */
    for (int r = 0; r < bh; r++)
      for (int c = 0; c < fbw; c++) {
        real sX = sigmaX, sY = sigmaY;
        if (pat) {
          float pv = pl->pattern[r * fbw + c];
          sX = pv * p->beta;
          sY = pv;
        }
        for (int b = 0; b < nb; b++) {
          size_t idx = (size_t)b * ps + r * fbw + c;
          real dRe = 0, dIm = 0;
          if (dg) {
            dRe = corrRe[idx];
            dIm = corrIm[idx];
          }
          if (p->bm == 2) { /* reference.cpp:261-272 */
            real pRe = wre[0][idx], pIm = wim[0][idx];
            real cRe = wre[1][idx], cIm = wim[1][idx];
            real aRe = cRe - pRe, aIm = cIm - pIm;
            real bRe = cRe + pRe, bIm = cIm + pIm;
            if (dg) {
              bRe -= (real)2.0f * dRe;
              bIm -= (real)2.0f * dIm;
            }
            wienner(&aRe, &aIm, sX, sY, betaX);
            wienner(&bRe, &bIm, sX, sY, betaX);
            filtRe[idx] = (aRe + bRe) * (real)0.5f + (dg ? dRe : (real)0.0f);
            filtIm[idx] = (aIm + bIm) * (real)0.5f + (dg ? dIm : (real)0.0f);
          } else if (p->bm == 3) { /* reference.cpp:273-292 */
            real pRe = wre[0][idx], pIm = wim[0][idx];
            real cRe = wre[1][idx], cIm = wim[1][idx];
            real nRe = wre[2][idx], nIm = wim[2][idx];
            real t1Re = pRe + nRe, t1Im = pIm + nIm;
            real y1Re = cRe - (real)0.5f * t1Re, y1Im = cIm - (real)0.5f * t1Im;
            real dnRe = nRe - pRe, dnIm = nIm - pIm;
            real y2Re = K * dnIm, y2Im = -K * dnRe;
            real d0Re = y1Re + y2Re, d0Im = y1Im + y2Im;
            real d1Re = cRe + t1Re, d1Im = cIm + t1Im;
            real d2Re = y1Re - y2Re, d2Im = y1Im - y2Im;
            if (dg) {
              d1Re -= (real)3.0f * dRe;
              d1Im -= (real)3.0f * dIm;
            }
            wienner(&d0Re, &d0Im, sX, sY, betaX);
            wienner(&d1Re, &d1Im, sX, sY, betaX);
            wienner(&d2Re, &d2Im, sX, sY, betaX);
            filtRe[idx] = (d0Re + d1Re + d2Re) * ((real)1.0f / (real)3.0f);
            filtIm[idx] = (d0Im + d1Im + d2Im) * ((real)1.0f / (real)3.0f);
            if (dg) {
              filtRe[idx] += dRe;
              filtIm[idx] += dIm;
            }
          } else { /* bm == 4, reference.cpp:293-315 */
            real p2Re = wre[0][idx], p2Im = wim[0][idx];
            real p1Re = wre[1][idx], p1Im = wim[1][idx];
            real cRe = wre[2][idx], cIm = wim[2][idx];
            real nRe = wre[3][idx], nIm = wim[3][idx];
            real fp1r = -p2Re + p1Im + cRe - nIm;
            real fp1i = -p2Im - p1Re + cIm + nRe;
            real fcr = p2Re + p1Re + cRe + nRe;
            real fci = p2Im + p1Im + cIm + nIm;
            real fnr = -p2Re - p1Im + cRe + nIm;
            real fni = -p2Im + p1Re + cIm - nRe;
            real fp2r = p2Re - p1Re + cRe - nRe;
            real fp2i = p2Im - p1Im + cIm - nIm;
            if (dg) {
              fp1r -= (real)4.0f * dRe;
              fp1i -= (real)4.0f * dIm;
            }
            wienner(&fp2r, &fp2i, sX, sY, betaX);
            wienner(&fp1r, &fp1i, sX, sY, betaX);
            wienner(&fcr, &fci, sX, sY, betaX);
            wienner(&fnr, &fni, sX, sY, betaX);
            filtRe[idx] = (fp2r + fp1r + fcr + fnr) * (real)0.25f;
            filtIm[idx] = (fp2i + fp1i + fci + fni) * (real)0.25f;
            if (dg) {
              filtRe[idx] += dRe;
              filtIm[idx] += dIm;
            }
          }
        }
      }
  } else if (p->bm == 1) { /* reference.cpp:318-333 */
    for (size_t i = 0; i < total; i++) {
      real re = curRe[i], im = curIm[i];
      real dRe = 0, dIm = 0;
      if (dg) {
        dRe = corrRe[i];
        dIm = corrIm[i];
        re -= dRe;
        im -= dIm;
      }
      real sX = sigmaX, sY = sigmaY;
      if (pat) {
        float pv = pl->pattern[i % ps];
        sX = pv * p->beta;
        sY = pv;
      }
      wienner(&re, &im, sX, sY, betaX);
      if (dg) {
        re += dRe;
        im += dIm;
      }
      filtRe[i] = re;
      filtIm[i] = im;
    }
  } else if (p->bm == 0) { /* Kalman, reference.cpp:334-383 */
    if (pl->kalmanFirst) {
      pl->kLastRe = (real*)malloc(sizeof(real) * total);
      pl->kLastIm = (real*)malloc(sizeof(real) * total);
      pl->kCovar = (real*)malloc(sizeof(real) * total);
      pl->kProc = (real*)malloc(sizeof(real) * total);
      memcpy(pl->kLastRe, curRe, sizeof(real) * total);
      memcpy(pl->kLastIm, curIm, sizeof(real) * total);
      for (size_t i = 0; i < total; i++) {
        real cn0 = pat ? (real)pl->pattern[i % ps] : (real)p->kalmanCN;
        pl->kCovar[i] = cn0;
        pl->kProc[i] = cn0;
      }
      pl->kalmanFirst = 0;
    }
    const real kr2 = (real)p->kratio * (real)p->kratio;
    for (size_t i = 0; i < total; i++) {
      real re = curRe[i], im = curIm[i];
      real lRe = pl->kLastRe[i], lIm = pl->kLastIm[i];
      real dRe = re - lRe, dIm = im - lIm;
      real d2r = dRe * dRe, d2i = dIm * dIm;
      real mask = (d2r < d2i) ? d2i : d2r; /* std::max(a, b) */
      real cn = pat ? (real)pl->pattern[i % ps] : (real)p->kalmanCN;
      real s = pl->kCovar[i] + pl->kProc[i];
      real gain = s / (s + cn);
      real newProc = gain * gain * cn;
      real newCovar = ((real)1.0f - gain) * s;
      real dstRe = gain * re + ((real)1.0f - gain) * lRe;
      real dstIm = gain * im + ((real)1.0f - gain) * lIm;
      if (mask > kr2 * cn) {
        dstRe = re;
        dstIm = im;
        newCovar = cn;
        newProc = cn;
      }
      filtRe[i] = dstRe;
      filtIm[i] = dstIm;
      pl->kCovar[i] = newCovar;
      pl->kProc[i] = newProc;
      pl->kLastRe[i] = dstRe;
      pl->kLastIm[i] = dstIm;
    }
  } else { /* bm == -1 */
    for (size_t i = 0; i < total; i++) {
      filtRe[i] = curRe[i];
      filtIm[i] = curIm[i];
    }
  }

  /* Phase 2: sharpen, inverse FFT, windowed overlap-add (reference.cpp:391-431). */
  real* half = pl->fullTmp; /* reused: ps complex fits in bw*bh complex */
  real* full = (real*)malloc(sizeof(real) * 2 * bw * bh);
  for (int b = 0; b < nb; b++) {
    if (p->sharpenOn) {
      for (int r = 0; r < bh; r++)
        for (int c = 0; c < fbw; c++) {
          size_t idx = (size_t)b * ps + r * fbw + c;
          real re = filtRe[idx], im = filtIm[idx];
          real dRe = 0, dIm = 0;
          if (dg) {
            dRe = corrRe[idx];
            dIm = corrIm[idx];
            re -= dRe;
            im -= dIm;
          }
          real psd = re * re + im * im + (real)0.0000000000000000000000000000000000001f;
          real factor = pl->sharpen[r * fbw + c];
          real g = (real)1.0f + factor * RSQRT_(psd * (real)p->ssmax /
                                               ((psd + (real)p->ssmin) * (psd + (real)p->ssmax)));
          re *= g;
          im *= g;
          if (dg) {
            re += dRe;
            im += dIm;
          }
          filtRe[idx] = re;
          filtIm[idx] = im;
        }
    }
    for (int i = 0; i < ps; i++) {
      half[2 * i] = filtRe[(size_t)b * ps + i];
      half[2 * i + 1] = filtIm[(size_t)b * ps + i];
    }
    fft2d_real_inverse_half(half, bw, bh, fbw, &pl->rowPlan, &pl->colPlan, pl->blockReal, full,
                            pl->fftScratch);
    const int x0 = pl->bx[b], y0 = pl->by[b];
    const float* win = (p->mode == 1) ? (pl->par[b] ? pl->winSynB : pl->winSynA) : pl->winAna;
    for (int ly = 0; ly < bh; ly++) {
      int dy = y0 + ly;
      if (dy < 0 || dy >= H) continue;
      real* orow = &out[(size_t)dy * W];
      const float* wrow = &win[ly * bw];
      const real* brow = &pl->blockReal[ly * bw];
      for (int lx = 0; lx < bw; lx++) {
        int dx = x0 + lx;
        if (dx < 0 || dx >= W) continue;
        orow[dx] += brow[lx] * wrow[lx];
      }
    }
  }
  free(full);
  if (pl->chromaNorm != 0.0f)
    for (int i = 0; i < W * H; i++) out[i] += pl->chromaNorm;
}

/* ------------------------------------------------------------- clip driver */

static int clamp_frame(int n, int numFrames) { /* src/fft3dcl.cpp:303-309 */
  if (n < 0) return 0;
  if (n > numFrames - 1) return numFrames - 1;
  return n;
}

typedef struct {
  int fn;
  real* re;
  real* im;
} cache_entry;

int orc_run_clip(const orc_script* sp, int color_family, int width, int height, int ssw, int ssh,
                 int bits, int sample_type, int nframes, const void* const* planes_in,
                 void* const* planes_out, real* const* float_out, const int* order, int norder,
                 const char** err) {
  const int nplanes = color_family == 1 ? 1 : 3;
  const int bps = sample_type == 1 ? 4 : (bits <= 8 ? 1 : 2);
  /* src/fft3dcl.cpp:144-157 format checks */
  if (sample_type == 1 && bits != 32) {
    *err = "FFT3DCL: only 32-bit float clips supported (16-bit float not ported yet)";
    return -1;
  }
  if (sample_type == 0 && (bits < 8 || bits > 16)) {
    *err = "FFT3DCL: integer bit depths 8-16 only";
    return -1;
  }
  /* src/fft3dcl.cpp:200-225 plane routing */
  int plane = sp->plane;
  if (color_family == 2) plane = 4;
  if (color_family == 1) {
    if (plane == 1) {
      *err = "FFT3DCL: cannot set chroma processing for a greyscale clip";
      return -1;
    }
    if (plane == 4) plane = 0;
  }
  if (plane == 2 || plane == 3) plane = 1;
  if (plane != 0 && plane != 1 && plane != 4) {
    *err = "FFT3DCL: plane must be 0 (luma), 1 (chroma) or 4 (all)";
    return -1;
  }
  int planes[3], np = 0;
  if (color_family == 1 || plane == 0) {
    planes[np++] = 0;
  } else if (plane == 1) {
    planes[np++] = 1;
    planes[np++] = 2;
  } else {
    planes[np++] = 0;
    planes[np++] = 1;
    planes[np++] = 2;
  }
  orc_internal ips[3];
  orc_plane* refs[3] = {0, 0, 0};
  int pw[3], ph[3];
  for (int k = 0; k < np; k++) {
    int q = planes[k];
    pw[k] = width >> (q ? ssw : 0);
    ph[k] = height >> (q ? ssh : 0);
    if (!orc_validate(sp, pw[k], ph[k], bits, &ips[k], err)) {
      for (int j = 0; j < k; j++) orc_plane_free(refs[j]);
      return -1;
    }
    refs[k] = orc_plane_new(&ips[k], pw[k], ph[k], bits, sample_type,
                            color_family == 3 && q > 0);
  }
  const orc_internal* ip = &ips[np - 1]; /* representative, as the reference's inst->ip */
  int degridDone[3] = {0, 0, 0};
  cache_entry* cache[3];
  int ncache[3] = {0, 0, 0};
  const int maxCache = nframes + 8;
  for (int k = 0; k < np; k++) cache[k] = (cache_entry*)calloc(maxCache, sizeof(cache_entry));
  float* planeBuf = (float*)malloc(sizeof(float) * (size_t)width * height);
  real* outBuf = (real*)malloc(sizeof(real) * (size_t)width * height);
  float* outF = (float*)malloc(sizeof(float) * (size_t)width * height);

  for (int oi = 0; oi < norder; oi++) {
    const int n = order[oi];
    /* untouched planes: straight copy (src/fft3dcl.cpp:347-363) */
    for (int q = 0; q < nplanes; q++) {
      int processed = 0;
      for (int k = 0; k < np; k++) processed |= planes[k] == q;
      if (processed) continue;
      int w = width >> (q ? ssw : 0), h = height >> (q ? ssh : 0);
      memcpy(planes_out[(size_t)n * nplanes + q], planes_in[(size_t)n * nplanes + q],
             (size_t)w * h * bps);
    }
    for (int k = 0; k < np; k++) {
      const int q = planes[k];
      orc_plane* ref = refs[k];
      const int ps = ref->fbw * ips[k].bh;
      const size_t total = (size_t)ref->nb * ps;
      if (ip->degridOn && !degridDone[k]) {
        orc_build_degrid(ref);
        degridDone[k] = 1;
      }
      const int nwin = ip->bm >= 2 ? ip->bt : 1;
      const real* wre[4];
      const real* wim[4];
      for (int i = 0; i < nwin; i++) {
        int fn = clamp_frame(ip->bm >= 2 ? n - ip->current + i : n, nframes);
        cache_entry* e = NULL;
        for (int c = 0; c < ncache[k]; c++)
          if (cache[k][c].fn == fn) e = &cache[k][c];
        if (!e) {
          e = &cache[k][ncache[k]++];
          e->fn = fn;
          e->re = (real*)malloc(sizeof(real) * total);
          e->im = (real*)malloc(sizeof(real) * total);
          orc_upload(ref, planes_in[(size_t)fn * nplanes + q], (ptrdiff_t)pw[k] * bps, planeBuf);
          orc_compute_spectra(ref, planeBuf, e->re, e->im);
        }
        wre[i] = e->re;
        wim[i] = e->im;
      }
      orc_filter_synth(ref, wre, wim, nwin, outBuf);
      if (float_out && float_out[(size_t)n * nplanes + q])
        memcpy(float_out[(size_t)n * nplanes + q], outBuf, sizeof(real) * (size_t)pw[k] * ph[k]);
      for (size_t i = 0; i < (size_t)pw[k] * ph[k]; i++) outF[i] = (float)outBuf[i];
      orc_download(ref, outF, planes_out[(size_t)n * nplanes + q], (ptrdiff_t)pw[k] * bps);
      /* prune (src/fft3dcl.cpp:511-521) */
      if (ip->bm >= 2) {
        int lo = n - ip->current - 1, hi = n - ip->current + ip->bt;
        int w = 0;
        for (int c = 0; c < ncache[k]; c++) {
          if (cache[k][c].fn < lo || cache[k][c].fn > hi) {
            free(cache[k][c].re);
            free(cache[k][c].im);
          } else {
            cache[k][w++] = cache[k][c];
          }
        }
        ncache[k] = w;
      } else if (ip->bm != 0) {
        for (int c = 0; c < ncache[k]; c++) {
          free(cache[k][c].re);
          free(cache[k][c].im);
        }
        ncache[k] = 0;
      } else if (ncache[k] >= maxCache - 1) {
        /* bm == 0 never prunes in the reference; entries are immutable, so
         * dropping them only costs a recompute. Bound the memory here. */
        for (int c = 0; c < ncache[k]; c++) {
          free(cache[k][c].re);
          free(cache[k][c].im);
        }
        ncache[k] = 0;
      }
    }
  }
  for (int k = 0; k < np; k++) {
    for (int c = 0; c < ncache[k]; c++) {
      free(cache[k][c].re);
      free(cache[k][c].im);
    }
    free(cache[k]);
    orc_plane_free(refs[k]);
  }
  free(planeBuf);
  free(outBuf);
  free(outF);
  return 0;
}
