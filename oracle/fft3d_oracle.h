/*
 * FFT3D denoise oracle: a plain-C restatement of the reference's CPU backend.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (fft3dcl_amd/, include/)
 * may call, link or load this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * What it restates (reference files, read as text, not copied):
 *   src/params.h:62-148            parameter validation and derived constants
 *   src/cpu/reference.cpp:10-437   block grid, upload/download, spectra,
 *                                   degrid sample, temporal filter, synthesis
 *   src/cpu/tables.cpp:8-156       windows, sigma pattern, sharpen LUT
 *   src/cpu/fft.cpp:8-119          radix-2 DIT FFT, 2-D real forward / inverse
 *   src/fft3dcl.cpp:303-331,471-521 clamped temporal window and spectra cache
 *
 * The single-precision build (liboracle.so) follows the reference's FP32
 * operation order term by term so that it is bit-identical to the reference
 * built with g++ -O2 -ffp-contract=off. The same source compiled with
 * -DORC_REAL=double (liboracle64.so) is the FP64 twin: identical tables and
 * constants, every spectral/filter/synthesis operation in double.
 */
#ifndef FFT3D_ORACLE_H
#define FFT3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic type of spectra, filter and synthesis: float in liboracle.so,
 * double in the FP64 twin liboracle64.so. */
#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL orc_real;
/* sizeof(orc_real) of the compiled library (4 or 8). */
int orc_real_size(void);

/* Script-level parameters (src/params.h:8-27), user units. */
typedef struct {
  float sigma, beta;
  int bw, bh, bt;
  float sharpen;
  int plane, mode, bordersize, precision;
  float degrid, scutoff, svr, smin, smax, kratio;
  int ow, oh, wintype, interlaced;
  float sigma2, sigma3, sigma4;
  int oldfft;
} orc_script;

/* Validated parameters (src/params.h:30-59). */
typedef struct {
  float sigma, beta;
  int bw, bh, bm, bt;
  float sharpen;
  int sharpenOn;
  int mode, bordersize, precision;
  float degrid;
  int degridOn;
  float scutoff, svr, smin, smax, kratio;
  int ow, oh, wintype, interlaced;
  float sigma2, sigma3, sigma4;
  int usePattern, current;
  int nx, ny, totw, toth, oldfft;
  float wienerS, betaX, kalmanCN, ssmin, ssmax;
} orc_internal;

/* Fills the reference defaults (src/fft3dcl.cpp:171-195 with no user keys). */
void orc_script_defaults(orc_script* sp);
/* src/params.h:62-72 */
int orc_round_pow2_clamp(int v);
/* src/params.h:76-148. Returns 1 on success, 0 with *err set on a hard error. */
int orc_validate(const orc_script* sp, int width, int height, int bits, orc_internal* out,
                 const char** err);

typedef struct orc_plane orc_plane;

/* PlaneRef ctor (src/cpu/reference.cpp:19-80). sample_type 0 int, 1 float. */
orc_plane* orc_plane_new(const orc_internal* p, int width, int height, int bits,
                         int sample_type, int chroma);
void orc_plane_free(orc_plane* pl);

int orc_plane_nblocks(const orc_plane* pl);
int orc_plane_fbw(const orc_plane* pl);
int orc_plane_ncols(const orc_plane* pl);
float orc_plane_chroma_norm(const orc_plane* pl);
/* Copies of the tables (float, bw*bh or fbw*bh); return element count. */
int orc_plane_bx(const orc_plane* pl, int* dst);
int orc_plane_by(const orc_plane* pl, int* dst);
int orc_plane_par(const orc_plane* pl, uint8_t* dst);
int orc_plane_win(const orc_plane* pl, int which /*0 ana, 1 synA, 2 synB*/, float* dst);
int orc_plane_pattern(const orc_plane* pl, float* dst); /* 0 when off */
int orc_plane_sharpen(const orc_plane* pl, float* dst); /* 0 when off */
/* Degrid sample gs (one planeSize copy, re/im interleaved); 0 before build. */
int orc_plane_degrid(const orc_plane* pl, float* dst);

/* PlaneRef::upload / download (src/cpu/reference.cpp:82-138). */
void orc_upload(const orc_plane* pl, const void* src, ptrdiff_t pitch, float* dst);
void orc_download(const orc_plane* pl, const float* src, void* dst, ptrdiff_t pitch);

/* PlaneRef::computeSpectra (src/cpu/reference.cpp:140-170). re/im hold
 * nblocks*fbw*bh values laid out [block][row][col]. */
void orc_compute_spectra(orc_plane* pl, const float* plane, orc_real* re, orc_real* im);
/* PlaneRef::buildDegridSample (src/cpu/reference.cpp:172-198). */
void orc_build_degrid(orc_plane* pl);
/* PlaneRef::filterAndSynthesize (src/cpu/reference.cpp:211-437). win_re/im
 * hold nwin spectra (bt of them for bm>=2, else one). out is W*H float. */
void orc_filter_synth(orc_plane* pl, const orc_real* const* win_re,
                      const orc_real* const* win_im, int nwin, orc_real* out);

/* 1-D and 2-D FFT helpers (src/cpu/fft.cpp). Complex data interleaved. */
void orc_fft1d(int n, const orc_real* in, orc_real* out, int inverse);
void orc_fft2d_real_forward(const orc_real* block, int bw, int bh, orc_real* full);
void orc_fft2d_real_inverse_half(const orc_real* half, int bw, int bh, orc_real* block);

/* Whole-clip driver: the reference's CPU getFrame path (src/fft3dcl.cpp:
 * 200-258 plane routing, 319-331 window, 375-386 degrid, 471-521 cache) for a
 * constant-format clip. planes_in/planes_out point at nplanes planes per frame
 * (frame-major: frame f plane q at index f*nplanes+q), each packed (pitch =
 * width*bytes). color_family: 1 gray, 2 RGB, 3 YUV. ssw/ssh = chroma
 * subsampling log2. order lists the requested output frames in call order.
 * float_out, when not NULL, receives each processed plane's synthesized value
 * before the pixel conversion (same indexing as planes_out, W*H orc_real;
 * NULL entries skipped). Returns 0, or -1 with *err set (reference strings). */
int orc_run_clip(const orc_script* sp, int color_family, int width, int height, int ssw,
                 int ssh, int bits, int sample_type, int nframes, const void* const* planes_in,
                 void* const* planes_out, orc_real* const* float_out, const int* order,
                 int norder, const char** err);

#ifdef __cplusplus
}
#endif
#endif
