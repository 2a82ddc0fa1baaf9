/*
 * fft3d_hip.h - C ABI of the MI355X (gfx950) backend of the FFT3D denoise path.
 *
 * This is the operator boundary that replaces the reference's OpenCL backend
 * class GpuRef (src/cl/gpu.h:16-88) behind the VapourSynth filter. Every entry
 * point names the reference interface it stands in for. Conventions:
 *   - plain C types only; no C++ exceptions cross the boundary;
 *   - functions return 0 (FFT3D_OK) or a negative FFT3D_E* code; a plane
 *     object keeps the reason as a sticky string (fft3d_plane_error), the
 *     reference's bool + err_ convention;
 *   - a plane object is used by one thread at a time (the reference holds its
 *     instance mutex around every call, src/fft3dcl.cpp:366);
 *   - `stream` is a hipStream_t (NULL = the plane's own stream). All device
 *     work of a call is ordered on that stream; calls with host buffers
 *     return after the host side is complete. A plane's calls share its
 *     buffers: a call on another stream than the plane's previous call first
 *     makes its stream wait for that call (an event), so callers may switch
 *     streams; fft3d_plane_destroy waits for the plane's last call only.
 */
#ifndef FFT3D_HIP_H
#define FFT3D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFT3D_ABI_VERSION 1

enum {
  FFT3D_OK = 0,
  FFT3D_EINVAL = -1,   /* bad argument or parameter (message says which) */
  FFT3D_EDEVICE = -2,  /* no usable HIP device */
  FFT3D_ENOMEM = -3,   /* device allocation failed */
  FFT3D_ELAUNCH = -4,  /* kernel launch or HIP runtime failure */
  FFT3D_EMISSING = -5, /* a window frame has no spectrum */
  FFT3D_EUNSUPPORTED = -6 /* parameter combination the GPU path does not run */
};

/* Script parameters of FFT3DCL(clip, ...), user units
 * (ScriptParams, src/params.h:8-27; defaults src/fft3dcl.cpp:171-195). */
typedef struct fft3d_params {
  float sigma, beta;
  int bw, bh, bt;
  float sharpen;
  int plane, mode, bordersize, precision;
  float degrid, scutoff, svr, smin, smax, kratio;
  int ow, oh, wintype, interlaced;
  float sigma2, sigma3, sigma4;
  int oldfft;
} fft3d_params;

/* One plane of a clip (PlaneGeom, src/cpu/reference.h:17-22). */
typedef struct fft3d_plane_desc {
  int width, height;
  int bits;        /* 8..16 for integer samples, 32 for float */
  int sample_type; /* 0 integer, 1 float */
  int chroma;      /* YUV chroma plane (mode-1 centring) */
} fft3d_plane_desc;

typedef struct fft3d_plane fft3d_plane;

/* Derived geometry and footprint of a plane object. */
typedef struct fft3d_plane_info {
  int nb;            /* blocks per frame */
  int bw, bh, ow, oh, mode, fbw;
  int bm;            /* temporal mode -1..4 (the script's bt) */
  int bt;            /* frames in the window (max(bm,1)) */
  int current;       /* index of frame n inside the window */
  int ncols, nrows;  /* block grid */
  int blocks_per_wg; /* blocks one 64-lane workgroup transforms */
  size_t spectrum_bytes; /* one frame's spectrum in the ring */
  size_t device_bytes;   /* all device memory the plane holds */
  int cached_frames;     /* spectra currently in the ring */
  int dft;               /* DFT form (FFT3D_OPT_DFT) */
  int spectra_fp16;      /* 1: the ring stores FP16 spectra */
  int band_rows;         /* fused sweep: block rows per band (0 off the sweep) */
  int sweep_occupancy;   /* fused sweep: workgroups per CU it is built for */
  int tuned;             /* 1: band rows / occupancy come from the tuning table */
} fft3d_plane_info;

/* Fills the reference's defaults for a clip with no keys set (degrid 1, ow/oh
 * 16 for bw/bh 32, sigma2/3/4 = sigma). */
void fft3d_params_default(fft3d_params* p);

/* Selects and probes HIP device `device` (GpuContext::init,
 * src/cl/cl.cpp:138-181). Returns FFT3D_EDEVICE with a reason when there is
 * no such device or it is not a gfx950 (its gcnArchName is checked). */
int fft3d_dev_init(int device, char* err, size_t errlen);

/* Validates parameters for this plane (validateParams, src/params.h:76-148)
 * and builds the plane object: tables, block grid, Kalman state, staging
 * (GpuRef::GpuRef, src/cl/gpu.cpp:59-245). On failure *out is NULL and err
 * holds the reference's message ("FFT3DCL: ..."). */
int fft3d_plane_create(const fft3d_params* params, const fft3d_plane_desc* desc,
                       fft3d_plane** out, char* err, size_t errlen);

/* Releases everything the plane holds (GpuRef::~GpuRef, src/cl/gpu.cpp:247-270)
 * after the plane's last call and its upload stream are done; work of other
 * planes is not waited for. */
void fft3d_plane_destroy(fft3d_plane* p);

/* Last error of the plane (GpuRef::error, src/cl/gpu.h:23). */
const char* fft3d_plane_error(const fft3d_plane* p);

/* 1 when frame fn's spectrum is in the ring (GpuRef::has, src/cl/gpu.h:35). */
int fft3d_plane_has(const fft3d_plane* p, int fn);

/* Analysis of frame fn: window + 2-D R2C FFT of every block into the ring
 * (GpuRef::forwardRaw / forward, src/cl/gpu.cpp:325-372). src holds the plane
 * as packed samples, width*height*bytes_per_sample bytes (u8, u16 or f32);
 * src_on_device = 1 reads it in place from device memory. */
int fft3d_plane_forward(fft3d_plane* p, int fn, const void* src, size_t bytes,
                        int src_on_device, void* stream);

/* Filter + inverse + overlap-add synthesis of the output frame for the
 * temporal window `window` (clamped frame numbers in temporal order: bt
 * entries for bt >= 2, else one), written to dst as packed samples
 * (GpuRef::processRaw / process, src/cl/gpu.cpp:474-505). */
int fft3d_plane_process(fft3d_plane* p, const int* window, int nwin, void* dst,
                        size_t bytes, int dst_on_device, void* stream);

/* fft3d_plane_forward(fn, src) followed by fft3d_plane_process(window, dst),
 * as one call: the per-frame sequence of the reference's scheduler, which
 * forwards each window frame missing from the cache and then processes the
 * window (src/fft3dcl.cpp:396-442). Same arguments, errors, ring contents and
 * output as the two calls. With FFT3D_OPT_FUSED_ANALYSIS set, fn in the
 * window and the geometry on the fused sweep, frame fn is transformed inside
 * the sweep (its spectrum still lands in the ring) instead of by its own
 * launch, which saves one spectrum read per frame. Otherwise it runs the two
 * calls. */
int fft3d_plane_forward_process(fft3d_plane* p, int fn, const void* src, size_t src_bytes,
                                int src_on_device, const int* window, int nwin, void* dst,
                                size_t dst_bytes, int dst_on_device, void* stream);

/* fft3d_plane_process with a device output, allocating nothing: no
 * counterpart in the reference. It is for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The kernels' scratch (the fused sweep's four partial
 * planes, or the blocks buffer off the sweep path) comes from a static device
 * array of 128 MiB per GPU, held until the process ends; fft3d_dev_init looks
 * up its address on each device (else the first call on a device does).
 * Limit: planes whose scratch exceeds 128 MiB (4 x W x H floats on the sweep
 * path, nb x bw x bh floats off it; 3840x2160 fits on the sweep path) get
 * FFT3D_EINVAL. Every call on one GPU shares the array and it keeps the last
 * call's contents: calls that overlap on one GPU (other streams, other
 * threads) race, so the caller serialises them. The plane's device must be
 * current. Same window, errors and output as fft3d_plane_process; no
 * profiling events, and a pending prefetch request is left for the next
 * fft3d_plane_process call. */
int fft3d_plane_process_noalloc(fft3d_plane* p, const int* window, int nwin, void* dst,
                                size_t bytes, void* stream);

/* forwardRaw(fn, bytes) issued ahead of need (src/cl/gpu.h, forwardRaw; the
 * host step src/fft3dcl.cpp:396-448 forwards each window frame in the call
 * that first needs it). Requests that plane fn go into the ring; the next
 * fft3d_plane_process or fft3d_plane_forward_process call on the plane
 * carries it out after launching its own kernels and before its readback,
 * on the plane's own upload stream:
 *   host src   : the host copy (pageable buffers go through pinned chunks
 *                that hold the whole plane) and the upload overlap that
 *                call's kernels;
 *   device src : the source is read after the work the caller queued on
 *                that call's stream before the call, so src may be written
 *                there just before.
 * In both cases the analysis starts after the call's sweep (or filter
 * kernel) and runs beside its seam combine (or synthesis) and readback.
 * src must stay valid, and a device src unchanged, until the prefetch is
 * done: until the carrying call returns for a host src, until the next call
 * whose window holds fn for a device src. One request at a time (a new one
 * replaces it); a no-op when fn is already in the ring; dropped by a prune
 * that excludes fn, by fft3d_plane_clear, and by a call that fails before its
 * readback. If the prefetch itself fails the frame is simply not in the ring
 * (fft3d_plane_has reports it). Ordering: the upload stream also waits for
 * the plane's previous call; a later process / forward_process whose window
 * holds fn waits for the prefetch on its own stream, and a forward that
 * rewrites fn or reuses a pooled spectrum waits for the latest prefetch. */
int fft3d_plane_prefetch(fft3d_plane* p, int fn, const void* src, size_t bytes,
                         int src_on_device);

/* Drops ring entries outside [lo, hi]; a negative bound is open
 * (GpuRef::pruneCache, src/cl/gpu.cpp:507-517). Retired spectra go to a
 * bounded reuse pool. */
void fft3d_plane_prune(fft3d_plane* p, int lo, int hi);

/* Drops every ring entry (the intent of pruneCache(-1,-2) at
 * src/fft3dcl.cpp:465, which in the reference keeps everything). */
void fft3d_plane_clear(fft3d_plane* p);

/* The serving path's batched step (no counterpart call in the reference: it
 * runs the reference's per-frame host step, src/fft3dcl.cpp:319-331 and
 * :396-467, for many planes at once). One work item is one output frame of
 * one plane. For each item, in order, the library:
 *   - builds the clamped temporal window of `frame` (bt >= 2: frames
 *     frame - current + i, i < bt, clamped to [0, num_frames - 1]; else the
 *     frame itself);
 *   - forwards the window frames missing from the ring, asking `source` for
 *     each one's plane (`bytes` long; on the device if src_on_device);
 *   - processes the window into `dst` on `stream` (null: the plane's own);
 *   - prunes the ring (bt >= 2: keep [frame - current - 1,
 *     frame - current + bt], as the reference does; bt 1, 0 and -1: clear.
 *     Kalman state is not held in the ring; the reference never prunes a
 *     bt=0 ring, src/fft3dcl.cpp:458-466).
 * Items of one plane must come in the order their frames are wanted (Kalman
 * frames in order). With device outputs nothing blocks: the items of
 * different planes overlap on their streams. A host output is complete when
 * the call returns. `status` receives each item's result. The return value
 * is FFT3D_OK, or the code of the first failing item (the plane's error
 * string has the reason); later items still run. */
typedef const void* (*fft3d_source_fn)(void* user, int frame);
typedef struct fft3d_work_item {
  fft3d_plane* plane;
  int frame;
  int num_frames;
  fft3d_source_fn source;
  void* user;
  int src_on_device;
  void* dst;
  size_t bytes;
  int dst_on_device;
  void* stream;
  int status;
} fft3d_work_item;
int fft3d_step(fft3d_work_item* items, int n);

/* Process-wide live plane objects and device bytes
 * (GpuRef::liveInstances / liveBytes, src/cl/gpu.h:27-28). */
int fft3d_live_instances(void);
size_t fft3d_live_bytes(void);

/* Per-kernel device time, measured with HIP events recorded on the launch
 * stream around every kernel while profiling is on (the reference's
 * FFT3DCL_TIME stage table, src/timing.h:12-65, without draining the queue). */
typedef struct fft3d_kernel_time {
  char name[32];
  int launches;
  double total_ms;
} fft3d_kernel_time;
int fft3d_plane_set_profiling(fft3d_plane* p, int on);
/* Waits for the recorded events, then returns per-kernel totals since the
 * last call (up to max entries) and resets them. */
int fft3d_plane_get_profile(fft3d_plane* p, fft3d_kernel_time* out, int max);

/* Execution options of a plane (tests and tuning; results never depend on
 * them beyond floating-point summation order):
 *   FFT3D_OPT_SWEEP      1 (default): fused filter/inverse/overlap-add sweep
 *                        where the geometry allows it (mode 1, ow = bw/2,
 *                        oh = bh/2); 0: separate filter/inverse and
 *                        synthesis kernels through a blocks buffer.
 *   FFT3D_OPT_BAND_ROWS  block rows per sweep band (0 = automatic).
 *   FFT3D_OPT_FUSED_ANALYSIS  0 (default): fft3d_plane_forward_process runs
 *                        the analysis launch then the sweep; 1: the sweep
 *                        transforms the new frame itself (bit-identical,
 *                        one launch and one spectrum read fewer, but slower
 *                        on gfx950 so far: the sweep is latency-bound).
 *   FFT3D_OPT_SPECTRA_FP16  1: stores the spectra ring as FP16 pairs (round
 *                        to nearest even; arithmetic stays FP32), half the
 *                        ring's bytes; drops the ring. With bt=0 on integer
 *                        samples the Kalman state (last value, covariance,
 *                        process noise) is stored as FP16 too, 8 instead of
 *                        16 bytes per bin; the recursion starts again at the
 *                        next frame. Square blocks of 16,
 *                        32 and 64 only (FFT3D_EUNSUPPORTED otherwise). The
 *                        reference's precision=0 storage; not exact.
 *   FFT3D_OPT_RESERVE_SPECTRA  allocates that many spectra into the reuse
 *                        pool now (FFT3D_ENOMEM if they do not fit), so that
 *                        a serving loop makes no device allocation later.
 *   FFT3D_OPT_HOST_STAGING  1 (default): host buffers in pageable memory go
 *                        through the plane's pinned chunks, copied by host
 *                        threads while the DMA engine moves the neighbouring
 *                        chunks; 0: handed to hipMemcpyAsync as they are.
 *                        Page-locked and device buffers never use the chunks.
 *   FFT3D_OPT_STAGING_CHUNK  bytes per pinned chunk (0 = 4 MiB; else a
 *                        multiple of 64 from 256 to 256 MiB).
 *   FFT3D_OPT_DFT        0 (default): the block DFTs as register FFTs in
 *                        FP32; 1: as dense matrix products on the matrix
 *                        cores (MFMA) with FP16 operands split hi + lo and
 *                        FP32 accumulation; 2: the same with BF16 operands.
 *                        Applies to the analysis and to the fused sweep's
 *                        inverse; 32x32 blocks on the fused sweep only
 *                        (FFT3D_EUNSUPPORTED otherwise). Replaces the
 *                        reference's radix-2 fft_batch stages
 *                        (src/cl/kernels.h:84-129); not exact (error stated
 *                        per configuration by the tests and the bench).
 *   FFT3D_OPT_SWEEP_OCC  0 (default): the tuning table's choice, else 2;
 *                        2..4: the register-FFT sweep built for that many
 *                        resident workgroups per CU (square 8 and 16 blocks
 *                        have 3, 8x8 also 4; FFT3D_EUNSUPPORTED otherwise).
 * FFT3D_OPT_BAND_ROWS 0 and FFT3D_OPT_SWEEP_OCC 0 take the plane's entry in
 * the tuning table (block size, plane size, ring precision, DFT form), read
 * once per process from $FFT3D_TUNING or tuning.json beside the library and
 * written by measurement (python -m fft3dcl_amd.tune). The reference sizes
 * its work-groups by formula (src/cl/gpu.cpp:70-81). Results do not depend
 * on these choices beyond floating-point summation order at band seams.
 * Returns FFT3D_EINVAL for an unknown option or value. */
enum {
  FFT3D_OPT_SWEEP = 1,
  FFT3D_OPT_BAND_ROWS = 2,
  FFT3D_OPT_FUSED_ANALYSIS = 3,
  FFT3D_OPT_RESERVE_SPECTRA = 4,
  FFT3D_OPT_SPECTRA_FP16 = 5,
  FFT3D_OPT_HOST_STAGING = 6,
  FFT3D_OPT_STAGING_CHUNK = 7,
  FFT3D_OPT_DFT = 8,
  FFT3D_OPT_SWEEP_OCC = 9
};
int fft3d_plane_set_option(fft3d_plane* p, int option, int value);

/* Introspection for tests and tools. */
int fft3d_plane_get_info(const fft3d_plane* p, fft3d_plane_info* info);
/* Host copy of a constant table: 0 analysis window, 1/2 synthesis windows
 * (even/odd block columns), 3 sigma pattern, 4 sharpen gains, 5 degrid
 * sample (re,im interleaved); mode 1 only: 6 row factors {shY, gyRe, gyIm,
 * strength} per row, 7 column factors {shX, gxRe, gxIm, 0} per column
 * (sharpen = strength - shY*shX, degrid = gx*gy). Returns the float count
 * (0 when off). */
int fft3d_plane_get_table(const fft3d_plane* p, int which, float* dst);
/* Frame fn's ring spectrum unpacked to the reference layout: nb blocks x bh
 * rows x (bw/2+1) columns of (re, im), i.e. PlaneRef::Spectra
 * (src/cpu/reference.h:30-33) interleaved. */
int fft3d_plane_read_spectrum(fft3d_plane* p, int fn, float* dst);

#ifdef __cplusplus
}
#endif
#endif
