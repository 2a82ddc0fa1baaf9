// C ABI of the MI355X backend (include/fft3d_hip.h): plane objects, the
// device-resident spectra ring, and the launches of the gfx950 kernels.
//
// Replaces the reference's GpuRef host class (src/cl/gpu.cpp). Differences
// by design: tables and grid are built here from the validated parameters
// (no CPU backend object is needed); the Kalman first-frame init is folded
// into the filter launch; spectra are stored packed (fk_kernels.h).
#define FK_COMMON_KERNELS  // k_synthesize and k_combine_seams live in this unit
#include "fk_set.h"
#include "fft3d_params.h"
#include "fft3d_hip.h"
#include "fk_xfer.h"

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

// Scratch of fft3d_plane_process_noalloc: the sweep's four partial planes,
// or the windowed blocks off the sweep path. One copy per GPU, held until the
// process ends. Host code never passes `scratch` itself to a kernel (in host
// code the name is a host-side stand-in): it uses the device address that
// hipGetSymbolAddress returns for the current device, kept in g_scratch.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

namespace {

std::atomic<int> g_live{0};
std::atomic<size_t> g_bytes{0};

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch[kMaxDevices];  // device address of `scratch` per device

// Looks up (once per device) the address of `scratch` on device `device`,
// which must be current. Returns null on failure.
void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  void* ptr = g_scratch[device].load();
  if (ptr) return ptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  g_scratch[device].store(ptr);
  return ptr;
}

using fk::AnalysisFn;
using fk::CombineFn;
using fk::FilterFn;
using fk::KernelSet;
using fk::SweepFn;
using fk::SynthFn;

template <int BW>
KernelSet kernels_for_bh(int bh) {
  switch (bh) {
  case 4: return fk::kernel_set<BW, 4>();
  case 8: return fk::kernel_set<BW, 8>();
  case 16: return fk::kernel_set<BW, 16>();
  case 32: return fk::kernel_set<BW, 32>();
  case 64: return fk::kernel_set<BW, 64>();
  }
  return KernelSet{};
}

KernelSet kernels_for(int bw, int bh) {
  switch (bw) {
  case 4: if (bh <= 64) return kernels_for_bh<4>(bh); break;
  case 8: if (bh <= 64) return kernels_for_bh<8>(bh); break;
  case 16: if (bh <= 64) return kernels_for_bh<16>(bh); break;
  case 32: if (bh <= 64) return kernels_for_bh<32>(bh); break;
  case 64: if (bh <= 64) return kernels_for_bh<64>(bh); break;
  }
  // a side above 64, bw * bh <= 4096 (the reference GPU's 32 KB of block
  // LDS, src/cl/gpu.cpp:82-87)
  return fk::kernel_set_large(bw, bh);
}

void set_err(char* err, size_t n, const std::string& s) {
  if (err && n) {
    std::snprintf(err, n, "%s", s.c_str());
  }
}

// IEEE binary16 -> float (host side of the FP16 spectra readback).
float half_to_float(uint32_t h) {
  const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
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

// True when the separable factors reproduce the sharpen and degrid tables to
// float rounding (the fused sweep builds the per-bin constants from them). A
// zero scutoff gives NaN entries that no product of factors reproduces.
bool sepMatches(const fft3d::Tables& t, const fft3d::Params& p) {
  const int fbw = p.bw / 2 + 1;
  if (t.sepRow.size() != 4u * p.bh || t.sepCol.size() != 4u * fbw) return false;
  for (int r = 0; r < p.bh; r++)
    for (int c = 0; c < fbw; c++) {
      const float* R = &t.sepRow[4 * r];
      const float* C = &t.sepCol[4 * c];
      const int i = r * fbw + c;
      if (p.sharpenOn) {
        const double v = (double)R[3] - (double)R[0] * C[0], want = t.sharpen[i];
        if (!(std::fabs(v - want) <= 1e-6 * (1.0 + std::fabs(want)))) return false;
      }
      if (p.degridOn) {
        const double re = (double)C[1] * R[1] - (double)C[2] * R[2];
        const double im = (double)C[1] * R[2] + (double)C[2] * R[1];
        const double scale = 1e-6 * (1.0 + std::fabs((double)p.degrid));
        if (!(std::fabs(re - t.degrid[2 * i]) <= scale) ||
            !(std::fabs(im - t.degrid[2 * i + 1]) <= scale))
          return false;
      }
    }
  return true;
}

// ---- tuning table ----
// Launch parameters chosen by measurement (fft3dcl_amd/tune.py) per block
// shape, plane size, ring precision and DFT form: sweep band rows and sweep
// workgroups per CU. Read once from $FFT3D_TUNING, else tuning.json next to
// this library; planes without an entry use the heuristic.
struct TuneKey {
  int bw, bh, w, h, half, dft;
  bool operator<(const TuneKey& o) const {
    return std::tie(bw, bh, w, h, half, dft) < std::tie(o.bw, o.bh, o.w, o.h, o.half, o.dft);
  }
};
struct TuneVal {
  int bandRows, occ;
};
std::map<TuneKey, TuneVal> g_tune;
std::string g_tunePath;
std::once_flag g_tuneOnce;

// Integer value of "key" in one flat JSON object, or def.
int json_int(const std::string& obj, const char* key, int def) {
  const std::string k = std::string("\"") + key + "\"";
  size_t i = obj.find(k);
  if (i == std::string::npos) return def;
  i = obj.find(':', i + k.size());
  if (i == std::string::npos) return def;
  char* end = nullptr;
  const long v = std::strtol(obj.c_str() + i + 1, &end, 10);
  return end == obj.c_str() + i + 1 ? def : (int)v;
}

void load_tuning() {
  const char* env = std::getenv("FFT3D_TUNING");
  std::string path;
  if (env && *env) {
    path = env;
  } else {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&load_tuning), &info) && info.dli_fname) {
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
             "/tuning.json";
    }
  }
  FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "rb");
  if (!f) return;
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, n);
  std::fclose(f);
  const size_t list = txt.find("\"entries\"");
  if (list == std::string::npos) return;
  size_t i = txt.find('[', list);
  while (i != std::string::npos) {
    const size_t a = txt.find('{', i), b = a == std::string::npos ? a : txt.find('}', a);
    if (b == std::string::npos) break;
    const std::string obj = txt.substr(a, b - a + 1);
    TuneKey k{json_int(obj, "bw", 0), json_int(obj, "bh", 0), json_int(obj, "width", 0),
              json_int(obj, "height", 0), json_int(obj, "half", 0), json_int(obj, "dft", 0)};
    TuneVal v{json_int(obj, "band_rows", 0), json_int(obj, "occupancy", 0)};
    if (k.bw > 0 && k.bh > 0 && k.w > 0 && k.h > 0) g_tune[k] = v;
    i = b + 1;
  }
  g_tunePath = path;
}

} // namespace

struct fft3d_plane {
  fft3d::Params p{};
  fft3d::Tables t;
  fft3d_plane_desc d{};
  int nb = 0, ncols = 0, nrows = 0, fbw = 0, bps = 4, pix = 0;
  int ncu = 0;  // compute units of the plane's device
  int mfmaPerCU = 0;  // resident MFMA analysis workgroups per CU (0: unknown)
  size_t specBytes = 0;
  KernelSet ks;
  fk::PlaneArgs args{};
  hipStream_t own = nullptr;
  std::string err;
  const char* errStatic = nullptr;  // set instead of err by fail_static
  int dev = 0;                      // the device current at create
  size_t bytes = 0;

  float* dTables = nullptr;  // windows, pattern, sharpen, degrid
  void* dStage = nullptr;    // host-upload / host-readback staging (W*H*bps)
  void* dStageIn = nullptr;  // host upload of a fused analysis (W*H*bps, on first use)
  fx::Staging stg;           // pinned chunks for pageable host buffers (on first use)
  bool stagingOn = true;     // option

  // fft3d_plane_prefetch: its own stream, upload buffers, and the events that
  // order it against the plane's other calls (all created on first use)
  hipStream_t up = nullptr;
  void* dStageUp = nullptr;
  fx::Staging stgUp;
  hipEvent_t lastCall = nullptr;      // end of the plane's latest call (any stream)
  hipStream_t lastStream = nullptr;   // the stream that call used
  bool anyCall = false;
  hipEvent_t upDone = nullptr;        // end of the latest prefetch
  std::map<int, hipEvent_t> upReady;  // prefetched spectrum -> end of its analysis
  std::vector<hipEvent_t> evSpare;
  int pendFn = -1;                    // requested prefetch, run by the next process call
  const void* pendSrc = nullptr;
  bool pendOnDevice = false;
  hipEvent_t callStart = nullptr;     // start of the call that runs the prefetch
  bool callStarted = false;
  // after that call's filter kernel (the sweep): the prefetch's analysis
  // starts there, beside the seam combine or synthesis and the readback,
  // instead of taking CU slots from the sweep
  hipEvent_t filterDone = nullptr;
  bool filterMarked = false;
  float* dBlocks = nullptr;  // windowed blocks awaiting overlap-add (unfused path)
  float4* dKalman = nullptr; // bm == 0 state (FP16 pairs with the FP16 ring)
  size_t kalmanBytes = 0;
  bool kalmanFirst = true;

  // fused sweep path
  bool sweepOk = false;      // geometry allows it
  bool sweepOn = true;       // option
  int bandRows = 0;          // option, 0 = automatic
  bool fuseOn = false;       // option: forward_process analyses inside the sweep
  bool half = false;         // option: spectra stored as FP16 pairs
  int dft = 0;               // option: 0 register FFTs, 1/2 MFMA DFT with FP16/BF16 operands
  int occOpt = 0;            // option: sweep workgroups per CU (0: tuned or 2)
  void* dDft = nullptr;      // the MFMA form's constant operands (fk_mfma.h)
  float* dPart = nullptr;    // 4 partial planes
  fk::SeamTile* dSeams = nullptr;  // border tiles of the current band layout
  int nseams = 0, seamR = 0;
  fk::SweepArgs sw{};

  std::map<int, fk::c32*> ring;  // frame number -> spectrum
  std::vector<fk::c32*> pool;    // retired spectra kept for reuse

  // profiling: events around each launch, resolved in get_profile
  enum { K_ANALYSIS, K_FILTER, K_SYNTH, K_SWEEP, K_COMBINE, K_FSWEEP, K_COUNT };
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  bool prof = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  hipEvent_t ev() {
    hipEvent_t e = nullptr;
    if (!evpool.empty()) {
      e = evpool.back();
      evpool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      e = nullptr;
    }
    return e;
  }
  // brackets a launch when profiling: returns the start event (or null)
  hipEvent_t prof_begin(hipStream_t s) {
    if (!prof) return nullptr;
    hipEvent_t a = ev();
    if (a) (void)hipEventRecord(a, s);
    return a;
  }
  void prof_end(hipEvent_t a, int kind, hipStream_t s) {
    if (!a) return;
    hipEvent_t b = ev();
    if (!b) {
      evpool.push_back(a);
      return;
    }
    (void)hipEventRecord(b, s);
    recs.push_back(Rec{kind, a, b});
  }

  ~fft3d_plane() {
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : evpool) (void)hipEventDestroy(e);
    for (auto& kv : ring) (void)hipFree(kv.second);
    for (auto* m : pool) (void)hipFree(m);
    (void)hipFree(dTables);
    (void)hipFree(dStage);
    (void)hipFree(dStageIn);
    stg.release();
    (void)hipFree(dStageUp);
    stgUp.release();
    for (auto& kv : upReady) (void)hipEventDestroy(kv.second);
    for (auto e : evSpare) (void)hipEventDestroy(e);
    if (lastCall) (void)hipEventDestroy(lastCall);
    if (callStart) (void)hipEventDestroy(callStart);
    if (filterDone) (void)hipEventDestroy(filterDone);
    if (upDone) (void)hipEventDestroy(upDone);
    if (up) (void)hipStreamDestroy(up);
    (void)hipFree(dBlocks);
    (void)hipFree(dKalman);
    (void)hipFree(dDft);
    (void)hipFree(dPart);
    (void)hipFree(dSeams);
    if (own) (void)hipStreamDestroy(own);
    g_bytes -= bytes;
  }

  bool alloc(void** ptr, size_t n) {
    if (hipMalloc(ptr, n) != hipSuccess) {
      *ptr = nullptr;
      (void)hipGetLastError();
      return false;
    }
    bytes += n;
    g_bytes += n;
    return true;
  }

  hipStream_t stream(void* s) const { return s ? static_cast<hipStream_t>(s) : own; }

  // Host side of the plane transfers: pageable buffers through the pinned
  // chunks (fk_xfer.h), others straight to the DMA engine. If the chunks
  // cannot be allocated the copy goes to hipMemcpyAsync as it is.
  bool staged(const void* host) {
    if (!stagingOn || !fx::pageable(host)) return false;
    if (!stg.ready() && stg.init() != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return true;
  }
  hipError_t upload(void* dDst, const void* src, size_t n, hipStream_t s) {
    if (staged(src)) return stg.upload(dDst, src, n, s);
    return hipMemcpyAsync(dDst, src, n, hipMemcpyHostToDevice, s);
  }
  // returns once the plane is in dst
  hipError_t download(void* dst, const void* dSrc, size_t n, hipStream_t s) {
    if (staged(dst)) return stg.download(dst, dSrc, n, s);
    const hipError_t e = stagingOn && !fx::pageable(dst)
                             ? fx::copy_to_host(dst, dSrc, n, s)
                             : hipMemcpyAsync(dst, dSrc, n, hipMemcpyDeviceToHost, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
  }

  // The ring entry of frame fn, taken from the pool or allocated if new, for
  // a call that writes it on stream s. Once prefetch is in use, s first waits
  // for a prefetch still writing that entry, or for the latest prefetch when
  // the buffer comes from the pool (it may have held a pruned prefetched
  // frame). A prefetch itself (s == up) is ordered by its own wait.
  fk::c32* spectrum(int fn, hipStream_t s) {
    auto it = ring.find(fn);
    if (it != ring.end()) {
      auto r = upReady.find(fn);
      if (r != upReady.end() && s != up) {
        if (hipStreamWaitEvent(s, r->second, 0) != hipSuccess) return nullptr;
        forget(fn);
      }
      return it->second;
    }
    fk::c32* m = nullptr;
    if (!pool.empty()) {
      if (up && s != up && hipStreamWaitEvent(s, upDone, 0) != hipSuccess) return nullptr;
      m = pool.back();
      pool.pop_back();
    } else if (!alloc(reinterpret_cast<void**>(&m), specBytes)) {
      return nullptr;
    }
    ring[fn] = m;
    return m;
  }

  // drops the prefetch event of frame fn (its ring entry is going away)
  void forget(int fn) {
    auto r = upReady.find(fn);
    if (r == upReady.end()) return;
    evSpare.push_back(r->second);
    upReady.erase(r);
  }
  void forget_all() {
    for (auto& kv : upReady) evSpare.push_back(kv.second);
    upReady.clear();
  }
  // s waits for the prefetches of the window's frames
  bool wait_prefetched(const int* win, int nwin, hipStream_t s) {
    for (int i = 0; i < nwin; i++) {
      auto r = upReady.find(win[i]);
      if (r != upReady.end() && hipStreamWaitEvent(s, r->second, 0) != hipSuccess) return false;
    }
    return true;
  }
  // Calls of one plane share its staging buffer, ring, pool and partial
  // planes. A call on another stream than the previous call first waits for
  // that call's end, so callers may switch streams between calls.
  bool order_call(hipStream_t s) {
    if (!anyCall || s == lastStream) return true;
    if (hipStreamWaitEvent(s, lastCall, 0) == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
  }
  // marks the end of a call on stream s (for the next call on another
  // stream, the next prefetch, and destroy)
  void end_call(hipStream_t s) {
    if (hipEventRecord(lastCall, s) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    lastStream = s;
    anyCall = true;
  }

  void release(fk::c32* m) {
    (void)hipFree(m);
    bytes -= specBytes;
    g_bytes -= specBytes;
  }

  void retire(fk::c32* m) {
    // keep at most bt+2 spare spectra (enough for sequential access); free the rest
    if (pool.size() < (size_t)(p.bt + 2)) {
      pool.push_back(m);
    } else if (holdFrees) {
      held.push_back(m);  // a planned launch may still read it
    } else {
      (void)hipFree(m);
      bytes -= specBytes;
      g_bytes -= specBytes;
    }
  }
  // While a batched step plans its rounds, spectra that would be freed are
  // held until its launches are queued.
  bool holdFrees = false;
  std::vector<fk::c32*> held;
  void release_held() {
    holdFrees = false;
    for (auto* m : held) {
      (void)hipFree(m);
      bytes -= specBytes;
      g_bytes -= specBytes;
    }
    held.clear();
  }

  // Chooses the band height and rebuilds the border-tile list for it.
  // The tuning table's entry for the current options, if any.
  const TuneVal* tuned() const {
    std::call_once(g_tuneOnce, load_tuning);
    auto it = g_tune.find(TuneKey{p.bw, p.bh, d.width, d.height, half ? 1 : 0, dft});
    return it == g_tune.end() ? nullptr : &it->second;
  }
  // Sweep workgroups per CU: the option, else the table, else 2 (register
  // FFTs only; the MFMA sweep has one variant).
  int occupancy() const {
    int o = occOpt;
    if (o <= 0) {
      const TuneVal* t = tuned();
      o = t && t->occ > 0 ? t->occ : 2;
    }
    if (dft || o > ks.maxOcc || o < 2 || !ks.sweepOcc[o - 2][half ? 1 : 0]) o = 2;
    return o;
  }
  // Block rows per sweep band for the current options.
  int bandRowsFor() const {
    const int nstrips = (ncols + ks.nbw - 1) / ks.nbw;
    int R = bandRows;
    if (R <= 0) {
      const TuneVal* t = tuned();
      if (t && t->bandRows > 0) R = t->bandRows;
    }
    if (R <= 0) {
      // one round of workgroups: as many as the device's CUs hold at once,
      // at least 2 rows per band
      const int occ = occupancy();
      const int perCU = occ > 2 ? occ : ks.sweepPerCU;
      const int slots = (ncu > 0 ? ncu : 256) * std::max(1, perCU);
      int nbands = (slots + nstrips - 1) / nstrips;
      R = (nrows + nbands - 1) / nbands;
      if (R < 2) R = 2;
    }
    return R > nrows ? nrows : R;
  }
  bool layoutSweep() {
    const int nbw = ks.nbw;
    const int nstrips = (ncols + nbw - 1) / nbw;
    const int R = bandRowsFor();
    if (R == seamR && (dSeams || nseams == 0)) return true;
    const int nbands = (nrows + R - 1) / R;
    const int sx = p.bw / 2, sy = p.bh / 2;
    const int ntx = (d.width + sx - 1) / sx, nty = (d.height + sy - 1) / sy;
    std::vector<char> rowIn(nty, 0), colIn(ntx, 0);
    for (int i = 0; i < nbands; i++)
      for (int u = i * R; u <= std::min(i * R + R, nrows) - 2; u++)
        if (u >= 0 && u < nty) rowIn[u] = 1;
    for (int st = 0; st < nstrips; st++) {
      const int nlive = std::min(nbw, ncols - st * nbw);
      for (int t = st * nbw; t <= st * nbw + nlive - 2; t++)
        if (t >= 0 && t < ntx) colIn[t] = 1;
    }
    // each border tile with the partial planes of the workgroups that touch
    // it, in the order the combine adds them (bands, then strips)
    std::vector<fk::SeamTile> seams;
    for (int u = 0; u < nty; u++)
      for (int t = 0; t < ntx; t++) {
        if (rowIn[u] && colIn[t]) continue;
        // band i covers tile rows [i*R - 1, min(i*R + R, nrows) - 1]
        int bands[2], nbd = 0;
        for (int i = u / R; i <= (u + 1) / R && nbd < 2; i++)
          if (i < nbands && i * R - 1 <= u && u <= std::min(i * R + R, nrows) - 1)
            if (nbd == 0 || bands[nbd - 1] != i) bands[nbd++] = i;
        // strip st covers tile columns [st*nbw - 1, st*nbw + nlive - 1]
        int strips[2], nst = 0;
        for (int st = t / nbw; st <= (t + 1) / nbw && nst < 2; st++) {
          if (st >= nstrips) continue;
          const int nlive = std::min(nbw, ncols - st * nbw);
          if (st * nbw - 1 <= t && t <= st * nbw + nlive - 1)
            if (nst == 0 || strips[nst - 1] != st) strips[nst++] = st;
        }
        fk::SeamTile tile{t, u, 0, 0};
        for (int a = 0; a < nbd; a++)
          for (int c = 0; c < nst; c++)
            tile.planes |= ((bands[a] & 1) * 2 + (strips[c] & 1)) << (2 * tile.np++);
        seams.push_back(tile);
      }
    if (dSeams) {
      (void)hipFree(dSeams);
      bytes -= (size_t)nseams * sizeof(fk::SeamTile);
      g_bytes -= (size_t)nseams * sizeof(fk::SeamTile);
      dSeams = nullptr;
    }
    nseams = (int)seams.size();
    const size_t seamBytes = seams.size() * sizeof(fk::SeamTile);
    if (nseams && (!alloc(reinterpret_cast<void**>(&dSeams), seamBytes) ||
                   hipMemcpy(dSeams, seams.data(), seamBytes, hipMemcpyHostToDevice) !=
                       hipSuccess)) {
      (void)hipGetLastError();
      return false;
    }
    seamR = R;
    sw.R = R;
    sw.nstrips = nstrips;
    sw.nbands = nbands;
    sw.ntx = ntx;
    sw.nty = nty;
    sw.part = dPart;
    sw.ppitch = (d.width + 3) & ~3;
    sw.seams = dSeams;
    sw.nseams = nseams;
    return true;
  }

  int fail(int code, const std::string& why) {
    err = why;
    errStatic = nullptr;
    return code;
  }

  // Window spectra of a process call (errors as fft3d_plane_process).
  // A failure whose message is a string literal: sets no std::string, so
  // the no-alloc entry point allocates nothing on the host either.
  int fail_static(int code, const char* why) {
    err.clear();
    errStatic = why;
    return code;
  }
  const char* error() const { return errStatic ? errStatic : err.c_str(); }

  // Window spectra of a process call (errors as fft3d_plane_process).
  int window_args(const int* win, int nwin, size_t dstBytes, fk::WindowArgs& w) {
    const int want = p.bm >= 2 ? p.bt : 1;
    if (nwin != want) return fail_static(FFT3D_EINVAL, "window length does not match bt");
    const size_t need = (size_t)d.width * d.height * bps;
    if (dstBytes < need) return fail_static(FFT3D_EINVAL, "output buffer too small");
    w = fk::WindowArgs{};
    w.nwin = nwin;
    w.cur = p.bm >= 2 ? p.current : 0;
    for (int i = 0; i < 4; i++) {
      auto it = ring.find(win[i < nwin ? i : 0]);
      if (it == ring.end()) return fail_static(FFT3D_EMISSING, "missing spectra in window");
      w.s[i] = it->second;
    }
    return FFT3D_OK;
  }
  int window(const int* win, int nwin, size_t dstBytes, fk::WindowArgs& w) {
    return window_args(win, nwin, dstBytes, w);
  }

  // The kernels of a process call: the sweep and seam combine with partial
  // planes `part`, or filter/inverse and synthesis through `blocks`. Returns
  // null, or the launch that failed (the HIP error is in lastLaunchErr).
  const char* launch(const fk::WindowArgs& w, void* outp, hipStream_t s, const void* fwdPlane,
                     fk::c32* fwdSpec, int fwdMask, float* part, float* blocks, bool profOk);
  hipError_t lastLaunchErr = hipSuccess;

  // Filter + synthesis of window w into dst. With fwdMask != 0 (sweep only)
  // the sweep also transforms fwdPlane into fwdSpec, the window slots in the
  // mask.
  int run(const fk::WindowArgs& w, void* dst, int dst_on_device, hipStream_t s,
          const void* fwdPlane, fk::c32* fwdSpec, int fwdMask);
  void run_prefetch();
  // the upload stream and its events (created on first use, after a device
  // synchronisation so that everything queued before comes first)
  bool ensure_up() {
    if (up) return true;
    bool ok = hipDeviceSynchronize() == hipSuccess &&
              hipStreamCreateWithFlags(&up, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&upDone, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&callStart, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&filterDone, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventRecord(upDone, up) == hipSuccess;
    if (!ok) {
      (void)hipGetLastError();
      if (up) (void)hipStreamDestroy(up);
      up = nullptr;
    }
    return ok;
  }
  // At the start of a process call on stream s with a prefetch pending: marks
  // what the caller queued on s before the call (a device source is ready
  // after it), without this call's own launches.
  void begin_call(hipStream_t s) {
    callStarted = false;
    if (pendFn < 0 || !ensure_up()) return;
    if (hipEventRecord(callStart, s) == hipSuccess) callStarted = true;
    else (void)hipGetLastError();
  }

  // Analysis of `plane` into `spec` on stream s (register FFTs, or the MFMA
  // form: 256-thread workgroups, one 4-block chunk each while they fit one
  // round on the device, else one round of them with longer runs; 32 blocks
  // each when the resident count is unknown).
  void launch_analysis(const void* plane, fk::c32* spec, hipStream_t s) {
    hipEvent_t t0 = prof_begin(s);
    if (dft) {
      AnalysisFn k = ks.analysisM[dft - 1][half ? 1 : 0];
      int grid = (nb + 31) / 32;
      if (mfmaPerCU > 0 && ncu > 0) grid = std::min((nb + 3) / 4, mfmaPerCU * ncu);
      FK_LAUNCH(k, dim3(grid), dim3(256), s, args, plane, spec);
    } else {
      const int grid = (nb + ks.G - 1) / ks.G;
      AnalysisFn k = half ? ks.analysisH : ks.analysis;
      FK_LAUNCH(k, dim3(grid), dim3(ks.fftThreads), s, args, plane, spec);
    }
    prof_end(t0, K_ANALYSIS, s);
  }

  // with a prefetch pending, marks the end of the call's filter kernel on s
  void mark_filter(hipStream_t s) {
    filterMarked = false;
    if (pendFn < 0 || !up) return;
    if (hipEventRecord(filterDone, s) == hipSuccess) filterMarked = true;
    else (void)hipGetLastError();
  }

  bool launched(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      fail(FFT3D_ELAUNCH, std::string(what) + " launch failed: " + hipGetErrorString(e));
      return false;
    }
    return true;
  }
};

const char* fft3d_plane::launch(const fk::WindowArgs& w, void* outp, hipStream_t s,
                                const void* fwdPlane, fk::c32* fwdSpec, int fwdMask, float* part,
                                float* blocks, bool profOk) {
  const int kfirst = (p.bm == 0 && kalmanFirst) ? 1 : 0;
  const bool pr = prof && profOk;
  auto begin = [&]() { return pr ? prof_begin(s) : nullptr; };
  auto ok = [&]() {
    lastLaunchErr = hipGetLastError();
    return lastLaunchErr == hipSuccess;
  };
  if (sweepOk && sweepOn) {
    fk::SweepArgs a = sw;
    a.part = part;
    a.out = outp;
    a.fwdPlane = fwdPlane;
    a.fwdSpec = fwdSpec;
    a.fwdMask = fwdMask;
    SweepFn ksw = fwdMask ? ks.sweepFwd : ks.sweepOcc[occupancy() - 2][half ? 1 : 0];
    if (dft && !fwdMask) ksw = ks.sweepM[dft - 1][half ? 1 : 0];
    hipEvent_t t0 = begin();
    FK_LAUNCH(ksw, dim3(a.nstrips, a.nbands), dim3(ks.sweepThreads), s, args, w, dKalman, kfirst,
              a);
    prof_end(t0, fwdMask ? K_FSWEEP : K_SWEEP, s);
    if (!ok()) return fwdMask ? "analysis/filter/synthesis sweep" : "filter/synthesis sweep";
    mark_filter(s);
    if (p.bm == 0) kalmanFirst = false;
    if (a.nseams) {
      CombineFn kc = &fk::k_combine_seams;
      t0 = begin();
      FK_LAUNCH(kc, dim3((a.nseams + 3) / 4), dim3(256), s, args, a);
      prof_end(t0, K_COMBINE, s);
      if (!ok()) return "seam combine";
    }
  } else {
    const int grid = (nb + ks.G - 1) / ks.G;
    FilterFn kf = half ? ks.filterH : ks.filter_inverse;
    hipEvent_t t0 = begin();
    FK_LAUNCH(kf, dim3(grid), dim3(ks.fftThreads), s, args, w, dKalman, kfirst, blocks);
    prof_end(t0, K_FILTER, s);
    if (!ok()) return "filter/inverse";
    mark_filter(s);
    if (p.bm == 0) kalmanFirst = false;
    SynthFn ksy = &fk::k_synthesize;
    t0 = begin();
    FK_LAUNCH(ksy, dim3((d.width + 15) / 16, (d.height + 15) / 16), dim3(256), s, args, blocks,
              outp);
    prof_end(t0, K_SYNTH, s);
    if (!ok()) return "synthesize";
  }
  return nullptr;
}

int fft3d_plane::run(const fk::WindowArgs& w, void* dst, int dst_on_device, hipStream_t s,
                     const void* fwdPlane, fk::c32* fwdSpec, int fwdMask) {
  const size_t need = (size_t)d.width * d.height * bps;
  void* outp = dst_on_device ? dst : dStage;
  if (sweepOk && sweepOn) {
    if (!layoutSweep()) return fail(FFT3D_ENOMEM, "GPU alloc failed (seam list)");
  } else if (fwdMask) {
    return fail(FFT3D_EINVAL, "fused analysis needs the sweep path");
  }
  if (const char* what = launch(w, outp, s, fwdPlane, fwdSpec, fwdMask, dPart, dBlocks, true))
    return fail(FFT3D_ELAUNCH,
                std::string(what) + " launch failed: " + hipGetErrorString(lastLaunchErr));
  if (pendFn >= 0) run_prefetch();
  if (!dst_on_device && download(dst, dStage, need, s) != hipSuccess) {
    (void)hipGetLastError();
    return fail(FFT3D_ELAUNCH, "readback failed");
  }
  return FFT3D_OK;
}

// The pending prefetch: frame pendFn from plane pendSrc into the ring on the
// upload stream. Runs inside a process call, after its launches, so that the
// host copies and the upload overlap that call's kernels, and the analysis
// its seam combine (or synthesis) and readback: the analysis waits for the
// call's sweep, whose CU slots it would otherwise take. A failure only drops
// the request (the frame is then forwarded when a window first needs it).
void fft3d_plane::run_prefetch() {
  const int fn = pendFn;
  const void* src = pendSrc;
  const bool onDevice = pendOnDevice;
  const bool started = callStarted, marked = filterMarked;
  pendFn = -1;
  pendSrc = nullptr;
  callStarted = false;
  filterMarked = false;
  if (fn < 0 || ring.count(fn) || !ensure_up()) return;
  const size_t need = (size_t)d.width * d.height * bps;
  // the plane's earlier calls may still read a pooled spectrum; a device
  // source is complete once the caller's earlier work on its stream is
  if (hipStreamWaitEvent(up, lastCall, 0) != hipSuccess ||
      (started && hipStreamWaitEvent(up, callStart, 0) != hipSuccess)) {
    (void)hipGetLastError();
    return;
  }
  if (onDevice && !started) return;  // no event to order the source by
  if (!onDevice && !dStageUp && !alloc(&dStageUp, need)) return;
  fk::c32* spec = spectrum(fn, up);
  if (!spec) {
    (void)hipGetLastError();
    return;
  }
  auto drop = [&]() {
    (void)hipGetLastError();
    ring.erase(fn);
    retire(spec);
  };
  bool staged = !onDevice && stagingOn && fx::pageable(src);
  if (staged && !stgUp.ready()) {
    // pinned chunks for the whole plane: the host copies never wait for a DMA
    stgUp.chunk = stg.chunk;
    const size_t n = (need + stgUp.chunk - 1) / stgUp.chunk;
    if (stgUp.init((int)std::min<size_t>(n, fx::Staging::MAXBUF)) != hipSuccess) {
      (void)hipGetLastError();
      staged = false;
    }
  }
  const hipError_t e =
      onDevice ? hipSuccess
               : (staged ? stgUp.upload(dStageUp, src, need, up)
                         : hipMemcpyAsync(dStageUp, src, need, hipMemcpyHostToDevice, up));
  if (e != hipSuccess) return drop();
  // the upload overlaps the call's kernels; the analysis waits for its sweep
  if (marked && hipStreamWaitEvent(up, filterDone, 0) != hipSuccess) return drop();
  const void* plane = onDevice ? src : dStageUp;
  launch_analysis(plane, spec, up);
  if (hipGetLastError() != hipSuccess) return drop();
  hipEvent_t ev = nullptr;
  if (!evSpare.empty()) {
    ev = evSpare.back();
    evSpare.pop_back();
  } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    ev = nullptr;
  }
  if (!ev || hipEventRecord(ev, up) != hipSuccess || hipEventRecord(upDone, up) != hipSuccess) {
    (void)hipGetLastError();
    if (ev) evSpare.push_back(ev);
    // no event to order the readers by: finish it here instead
    if (hipStreamSynchronize(up) != hipSuccess) drop();
    return;
  }
  upReady[fn] = ev;
}

namespace {

// ---- the serving step (fft3d_step) ----

bool step_item_valid(const fft3d_work_item& it) {
  return it.plane && it.source && it.dst && it.num_frames > 0 && it.frame >= 0 &&
         it.frame < it.num_frames;
}

// The clamped temporal window of the item's frame and the window frames not
// in the ring (src/fft3dcl.cpp:319-331, :424-430).
void step_window(const fft3d_work_item& it, int* window, int& nwin, int* missing, int& nmiss) {
  const fft3d_plane* p = it.plane;
  const int bm = p->p.bm, bt = p->p.bt, cur = p->p.current, nf = it.num_frames;
  auto clampf = [nf](int f) { return f < 0 ? 0 : (f > nf - 1 ? nf - 1 : f); };
  nwin = bm >= 2 ? bt : 1;
  nmiss = 0;
  for (int i = 0; i < nwin; i++) {
    const int fn = clampf(bm >= 2 ? it.frame - cur + i : it.frame);
    window[i] = fn;
    bool dup = false;
    for (int j = 0; j < nmiss; j++) dup = dup || missing[j] == fn;
    if (!dup && !p->ring.count(fn)) missing[nmiss++] = fn;
  }
}

// The ring rule after an output frame: bt >= 2 keeps [n - current - 1,
// n - current + bt] (src/fft3dcl.cpp:462); the one-frame modes and Kalman
// (whose state is not in the ring) keep nothing.
void step_prune(fft3d_plane* p, int frame) {
  if (p->p.bm >= 2) fft3d_plane_prune(p, frame - p->p.current - 1, frame - p->p.current + p->p.bt);
  else fft3d_plane_clear(p);
}

// One item through the plane calls: forwards of the missing frames, the last
// one with the filter pass, then the prune.
int step_item(fft3d_work_item& it) {
  fft3d_plane* p = it.plane;
  if (!step_item_valid(it)) {
    if (p) p->fail(FFT3D_EINVAL, "bad work item");
    return FFT3D_EINVAL;
  }
  int window[4], missing[4], nwin = 0, nmiss = 0, rc = FFT3D_OK;
  step_window(it, window, nwin, missing, nmiss);
  const void* src[4] = {};
  for (int j = 0; j < nmiss && rc == FFT3D_OK; j++)
    if (!(src[j] = it.source(it.user, missing[j])))
      rc = p->fail(FFT3D_EINVAL, "work item source gave no plane");
  for (int j = 0; j + 1 < nmiss && rc == FFT3D_OK; j++)
    rc = fft3d_plane_forward(p, missing[j], src[j], it.bytes, it.src_on_device, it.stream);
  if (rc == FFT3D_OK)
    rc = nmiss ? fft3d_plane_forward_process(p, missing[nmiss - 1], src[nmiss - 1], it.bytes,
                                             it.src_on_device, window, nwin, it.dst, it.bytes,
                                             it.dst_on_device, it.stream)
               : fft3d_plane_process(p, window, nwin, it.dst, it.bytes, it.dst_on_device,
                                     it.stream);
  if (rc == FFT3D_OK) step_prune(p, it.frame);
  return rc;
}

// Items that go through the batched launches: device sources and outputs, a
// plane on the fused sweep of the current device with no pending prefetch.
bool step_batchable(const fft3d_work_item& it, int dev) {
  const fft3d_plane* p = it.plane;
  const int hf = p->half ? 1 : 0;
  return it.src_on_device && it.dst_on_device && p->dev == dev && p->sweepOk && p->sweepOn &&
         p->pendFn < 0 && p->ks.analysisB[p->dft][hf] && p->ks.sweepB[p->dft][hf] &&
         (p->dft || p->ks.sweepBOcc[p->occupancy() - 2][hf]);
}

// Process-wide state of the batched step: its stream, the descriptor buffer
// in device memory with rotating pinned staging, ordering events, and the
// hipGraphs of the launch sequences seen so far (a launch sequence is
// replayed when the same kernels, grids and descriptor offsets recur; the
// descriptors themselves are uploaded before every launch).
struct StepEngine {
  std::mutex mu;
  int dev = -1;
  hipStream_t bs = nullptr;
  hipEvent_t done = nullptr;
  fk::BatchItem* dItems = nullptr;
  size_t capItems = 0;
  static constexpr int NSTAGE = 4;
  fk::BatchItem* hItems[NSTAGE] = {};
  size_t hCap[NSTAGE] = {};
  hipEvent_t hEv[NSTAGE] = {};
  bool hUsed[NSTAGE] = {};
  int nextStage = 0;
  std::vector<hipEvent_t> evs;
  std::map<std::vector<uintptr_t>, hipGraphExec_t> graphs;
  int graphsOn = -1;  // FFT3D_STEP_GRAPHS=0 launches directly
  long graphHits = 0, graphBuilds = 0;
  std::vector<const void*> loaded;  // kernels whose code object is loaded

  bool init(int device) {
    if (dev == device && bs) return true;
    if (bs) return false;  // one device per process for the batched step
    if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      bs = nullptr;
      return false;
    }
    for (int i = 0; i < NSTAGE; i++)
      if (hipEventCreateWithFlags(&hEv[i], hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
    dev = device;
    if (graphsOn < 0) {
      const char* e = std::getenv("FFT3D_STEP_GRAPHS");
      graphsOn = (e && e[0] == '0') ? 0 : 1;
#if defined(FFT3D_CPU_EMU)
      graphsOn = 0;  // the emulation runs launches at once: no capture
#endif
    }
    return true;
  }
  // streams of the kernel classes of a step call and their join events
  std::vector<hipStream_t> cls;
  std::vector<hipEvent_t> join;
  bool class_streams(size_t n) {
    while (cls.size() < n) {
      hipStream_t st = nullptr;
      hipEvent_t ev = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
      cls.push_back(st);
      join.push_back(ev);
    }
    return true;
  }
  // ordering event i (created on first use); false if it cannot be created
  bool event(size_t i, hipEvent_t& e) {
    while (evs.size() <= i) {
      hipEvent_t x = nullptr;
      if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
      evs.push_back(x);
    }
    e = evs[i];
    return true;
  }
  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  // A descriptor buffer of at least n entries for a step call (before its
  // first upload; a new buffer drops the graphs, whose kernel arguments
  // point at the old one).
  bool reserve(size_t n) {
    if (n <= capItems) return true;
    if (hipStreamSynchronize(bs) != hipSuccess) return false;
    for (hipStream_t c : cls)
      if (hipStreamSynchronize(c) != hipSuccess) return false;
    drop_graphs();
    (void)hipFree(dItems);
    dItems = nullptr;
    capItems = std::max<size_t>(256, 2 * n);
    if (hipMalloc(reinterpret_cast<void**>(&dItems), capItems * sizeof(fk::BatchItem)) !=
        hipSuccess) {
      capItems = 0;
      return false;
    }
    return true;
  }
  // Device copy of descriptors into [at, at + n) on bs, through a pinned
  // stage (the previous step's kernels are done with the buffer by the time
  // the copy runs: bs waits for them).
  bool upload(const std::vector<fk::BatchItem>& v, size_t at) {
    const size_t n = v.size();
    if (at + n > capItems) return false;
    const int k = nextStage;
    nextStage = (nextStage + 1) % NSTAGE;
    if (hUsed[k] && hipEventSynchronize(hEv[k]) != hipSuccess) return false;
    if (hCap[k] < n) {
      if (hItems[k]) (void)hipHostFree(hItems[k]);
      hItems[k] = nullptr;
      hCap[k] = std::max<size_t>(256, 2 * n);
      if (hipHostMalloc(reinterpret_cast<void**>(&hItems[k]), hCap[k] * sizeof(fk::BatchItem),
                        hipHostMallocDefault) != hipSuccess) {
        hCap[k] = 0;
        return false;
      }
    }
    std::memcpy(hItems[k], v.data(), n * sizeof(fk::BatchItem));
    if (hipMemcpyAsync(dItems + at, hItems[k], n * sizeof(fk::BatchItem),
                       hipMemcpyHostToDevice, bs) != hipSuccess ||
        hipEventRecord(hEv[k], bs) != hipSuccess)
      return false;
    hUsed[k] = true;
    return true;
  }
  // the event that marks chunk i's descriptors in place (created on first use)
  std::vector<hipEvent_t> chunkEvs;
  bool chunk_event(size_t i, hipEvent_t& e) {
    while (chunkEvs.size() <= i) {
      hipEvent_t x = nullptr;
      if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
      chunkEvs.push_back(x);
    }
    e = chunkEvs[i];
    return true;
  }
};
StepEngine g_step;

// One batched launch: kernel, workgroup size, descriptors [first, first + n)
// and total workgroups.
struct BatchLaunch {
  fk::BatchFn fn;
  int threads;
  size_t first, n;
  int grid;
};
// The launches of one kernel class (block shape, ring precision, DFT form)
// in one round of a step call: analysis, sweep, seam combine, in that order.
struct ClassRound {
  BatchLaunch l[3];
  std::vector<fk::BatchItem> d[3];
};

void batch_add(ClassRound& cr, int which, fk::BatchFn fn, int threads, fk::BatchItem bi) {
  BatchLaunch& l = cr.l[which];
  if (cr.d[which].empty()) l = BatchLaunch{fn, threads, 0, 0, 0};
  bi.wg0 = (l.grid + 7) & ~7;  // items start on a multiple of 8 workgroups
  l.grid = bi.wg0 + bi.nwg;
  cr.d[which].push_back(bi);
}

// The batchable items of a step call, round by round (round r: each plane's
// r-th item). Each kernel class runs its rounds in order on a stream of its
// own, started after the step stream and joined back, so classes overlap;
// the step stream starts after every item's stream and each plane's previous
// call, and every item's stream continues after the step. The rounds go to
// the GPU in chunks of 1, 2, 4, ... rounds, so that the host plans a chunk
// while the GPU runs the ones before it: each chunk's descriptors are copied
// into a range of their own, and each class's share of the chunk follows
// them on the class stream. With graphs on, that share is a hipGraph, built
// once per shape and replayed.
void step_batch(fft3d_work_item* items, const std::vector<std::vector<int>>& rounds) {
  StepEngine& E = g_step;
  std::lock_guard<std::mutex> lk(E.mu);
  std::vector<int> all_items;
  for (auto& r : rounds) all_items.insert(all_items.end(), r.begin(), r.end());
  if (all_items.empty()) return;
  auto fail_all = [&](int code, const char* why) {
    for (int k : all_items)
      if (items[k].status == FFT3D_OK) items[k].status = items[k].plane->fail_static(code, why);
  };
  for (int k : all_items) items[k].status = FFT3D_OK;
  if (!E.init(items[all_items[0]].plane->dev)) {
    // no step stream: the plane calls serve the items one by one
    for (auto& r : rounds)
      for (int k : r) items[k].status = step_item(items[k]);
    return;
  }
  hipStream_t bs = E.bs;
  std::vector<hipStream_t> streams;
  for (int k : all_items) {
    fft3d_plane* p = items[k].plane;
    hipStream_t s = p->stream(items[k].stream);
    if (std::find(streams.begin(), streams.end(), s) == streams.end()) streams.push_back(s);
    if (!p->order_call(bs)) return fail_all(FFT3D_ELAUNCH, "stream wait failed");
  }
  for (size_t i = 0; i < streams.size(); i++) {
    hipEvent_t e = nullptr;
    if (!E.event(i, e) || hipEventRecord(e, streams[i]) != hipSuccess ||
        hipStreamWaitEvent(bs, e, 0) != hipSuccess) {
      (void)hipGetLastError();
      return fail_all(FFT3D_ELAUNCH, "step stream ordering failed");
    }
  }
  // host side of every round: windows, sources, ring entries, descriptors
  for (int k : all_items) items[k].plane->holdFrees = true;
  struct Release {  // frees held spectra once the launches are queued (or failed)
    fft3d_work_item* items;
    const std::vector<int>& ks;
    ~Release() {
      for (int k : ks) items[k].plane->release_held();
    }
  } release{items, all_items};
  std::vector<std::pair<fk::BatchFn, int>> classes;  // sweep kernel, plane size range
  std::vector<std::vector<ClassRound>> plan;   // [class][round of the chunk]
  struct Added {  // a ring entry an analysis fills, and the chunk that launches it
    fft3d_plane* p;
    int fn, chunk;
  };
  std::vector<Added> added;
  std::vector<char> clsUsed;  // class streams that received work
  // descriptors for every chunk of the call, each chunk in its own range
  size_t maxDesc = 0;
  for (int k : all_items) maxDesc += 6;  // <= 4 analyses, a sweep, a combine
  bool good = E.reserve(maxDesc);
  size_t base = 0;
  int chunk = 0, failedChunk = -1;
  // Rounds go to the GPU in chunks of 1, 2, 4, ... rounds: the host plans a
  // chunk while the GPU runs the ones before it. Each class's share of a
  // chunk goes onto the class's stream after the chunk's descriptors.
  auto flush = [&]() -> bool {
    std::vector<fk::BatchItem> desc;
    std::vector<std::vector<BatchLaunch>> seq(plan.size());
    for (size_t c = 0; c < plan.size(); c++)
      for (ClassRound& cr : plan[c])
        for (int q = 0; q < 3; q++) {
          if (cr.d[q].empty()) continue;
          cr.l[q].first = base + desc.size();
          cr.l[q].n = cr.d[q].size();
          desc.insert(desc.end(), cr.d[q].begin(), cr.d[q].end());
          seq[c].push_back(cr.l[q]);
        }
    for (auto& pc : plan) pc.clear();
    if (desc.empty()) return true;
    hipEvent_t ready = nullptr;
    if (!E.class_streams(seq.size()) || !E.chunk_event(chunk, ready) ||
        !E.upload(desc, base) || hipEventRecord(ready, bs) != hipSuccess)
      return false;
    base += desc.size();
    if (clsUsed.size() < seq.size()) clsUsed.resize(seq.size(), 0);
    for (size_t c = 0; c < seq.size(); c++) {
      if (seq[c].empty()) continue;
      hipStream_t cs = E.cls[c];
      if (hipStreamWaitEvent(cs, ready, 0) != hipSuccess) return false;
      clsUsed[c] = 1;
      auto launch_class = [&]() {
        for (const BatchLaunch& l : seq[c])
          FK_LAUNCH(l.fn, dim3(l.grid), dim3(l.threads), cs, E.dItems + l.first, (int)l.n);
        return hipGetLastError() == hipSuccess;
      };
      if (!E.graphsOn) {
        if (!launch_class()) return false;
        continue;
      }
      // a capture must not be the first use of a kernel (its code object is
      // loaded then): query each kernel once beforehand
      std::vector<uintptr_t> key{reinterpret_cast<uintptr_t>(E.dItems), ~(uintptr_t)c};
      for (const BatchLaunch& l : seq[c]) {
        const void* f = reinterpret_cast<const void*>(l.fn);
        if (std::find(E.loaded.begin(), E.loaded.end(), f) == E.loaded.end()) {
          hipFuncAttributes attr;
          if (hipFuncGetAttributes(&attr, f) != hipSuccess) (void)hipGetLastError();
          else E.loaded.push_back(f);
        }
        key.insert(key.end(), {reinterpret_cast<uintptr_t>(l.fn), (uintptr_t)l.grid,
                               (uintptr_t)l.threads, (uintptr_t)l.first, (uintptr_t)l.n});
      }
      auto g = E.graphs.find(key);
      if (g == E.graphs.end()) {
        if (E.graphs.size() >= 256) E.drop_graphs();
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        bool built = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (built) {
          built = launch_class();
          built = hipStreamEndCapture(cs, &graph) == hipSuccess && built;
        }
        built = built && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!built) return false;
        g = E.graphs.emplace(key, exec).first;
        E.graphBuilds++;
      } else {
        E.graphHits++;
      }
      if (hipGraphLaunch(g->second, cs) != hipSuccess) return false;
    }
    return true;
  };
  size_t r0 = 0, span = 1;
  for (size_t r = 0; r < rounds.size() && good; r++) {
    for (int k : rounds[r]) {
      fft3d_work_item& it = items[k];
      fft3d_plane* p = it.plane;
      int window[4], missing[4], nwin = 0, nmiss = 0;
      step_window(it, window, nwin, missing, nmiss);
      const void* src[4] = {};
      bool ok = true;
      for (int j = 0; j < nmiss && ok; j++)
        if (!(src[j] = it.source(it.user, missing[j]))) {
          it.status = p->fail_static(FFT3D_EINVAL, "work item source gave no plane");
          ok = false;
        }
      const size_t need = (size_t)p->d.width * p->d.height * p->bps;
      if (ok && it.bytes < need) {
        it.status = p->fail_static(FFT3D_EINVAL, "plane buffer too small");
        ok = false;
      }
      if (ok && !p->layoutSweep()) {
        it.status = p->fail_static(FFT3D_ENOMEM, "GPU alloc failed (seam list)");
        ok = false;
      }
      fk::c32* spec[4] = {};
      for (int j = 0; j < nmiss && ok; j++)
        if (!(spec[j] = p->spectrum(missing[j], bs))) {
          it.status = p->fail_static(FFT3D_ENOMEM, "GPU spectra alloc failed");
          ok = false;
        }
      fk::WindowArgs w{};
      if (ok && (it.status = p->window_args(window, nwin, it.bytes, w)) != FFT3D_OK) ok = false;
      if (ok && !p->wait_prefetched(window, nwin, bs)) {
        (void)hipGetLastError();
        it.status = p->fail_static(FFT3D_ELAUNCH, "wait for prefetched spectra failed");
        ok = false;
      }
      if (!ok) {
        for (int j = 0; j < nmiss; j++)
          if (spec[j]) {
            p->ring.erase(missing[j]);
            p->retire(spec[j]);
          }
        continue;
      }
      const int hf = p->half ? 1 : 0;
      const fk::BatchFn sfn =
          p->dft ? p->ks.sweepB[p->dft][hf] : p->ks.sweepBOcc[p->occupancy() - 2][hf];
      // a class is one sweep kernel and one plane size range: a round of a
      // class ends with its largest plane, so large and small planes go on
      // separate streams and small ones are not held back by large ones
      const long px = (long)p->d.width * p->d.height;
      const int bucket = px <= (1l << 20) ? 0 : (px <= (3l << 20) ? 1 : 2);
      const std::pair<fk::BatchFn, int> key{sfn, bucket};
      size_t c = std::find(classes.begin(), classes.end(), key) - classes.begin();
      if (c == classes.size()) {
        classes.push_back(key);
        plan.emplace_back();
      }
      if (plan[c].size() <= r - r0) plan[c].resize(r - r0 + 1);
      ClassRound& cr = plan[c][r - r0];
      for (int j = 0; j < nmiss; j++) {
        added.push_back({p, missing[j], chunk});
        fk::BatchItem bi{};
        bi.P = p->args;
        bi.plane = src[j];
        bi.spec = spec[j];
        // MFMA: two 4-block chunks per workgroup, which on the served mix
        // beat one chunk by 1.4 % and eight by 3.2 %
        bi.nwg = p->dft ? (p->nb + 7) / 8 : (p->nb + p->ks.G - 1) / p->ks.G;
        batch_add(cr, 0, p->ks.analysisB[p->dft][hf], p->dft ? 256 : 64, bi);
      }
      fk::BatchItem bi{};
      bi.P = p->args;
      bi.Wn = w;
      bi.A = p->sw;
      bi.A.part = p->dPart;
      bi.A.out = it.dst;
      bi.A.fwdPlane = nullptr;
      bi.A.fwdSpec = nullptr;
      bi.A.fwdMask = 0;
      bi.kstate = p->dKalman;
      bi.kfirst = (p->p.bm == 0 && p->kalmanFirst) ? 1 : 0;
      bi.nwg = bi.A.nstrips * bi.A.nbands;
      batch_add(cr, 1, sfn, p->ks.sweepThreads, bi);
      if (bi.A.nseams) {
        bi.nwg = (bi.A.nseams + 3) / 4;
        batch_add(cr, 2, &fk::k_combine_batch, 256, bi);
      }
      // the ring and Kalman bookkeeping of this item, before the plane's next round
      if (p->p.bm == 0) p->kalmanFirst = false;
      step_prune(p, it.frame);
    }
    if (r + 1 == r0 + span || r + 1 == rounds.size()) {
      if (!flush()) {
        (void)hipGetLastError();
        failedChunk = chunk;
        good = false;
      }
      chunk++;
      r0 = r + 1;
      span *= 2;
    }
  }
  if (!good) {
    (void)hipGetLastError();
    // the analyses of the failed chunk and after did not run: their ring
    // entries hold nothing
    for (const Added& a : added) {
      if (failedChunk >= 0 && a.chunk < failedChunk) continue;
      auto r = a.p->ring.find(a.fn);
      if (r != a.p->ring.end()) {
        a.p->retire(r->second);
        a.p->ring.erase(r);
      }
    }
  }
  // the step stream continues after every class stream's work
  for (size_t c = 0; c < clsUsed.size(); c++)
    if (clsUsed[c] && (hipEventRecord(E.join[c], E.cls[c]) != hipSuccess ||
                       hipStreamWaitEvent(bs, E.join[c], 0) != hipSuccess)) {
      (void)hipGetLastError();
      good = false;
    }
  if (!good) fail_all(FFT3D_ELAUNCH, "batched step launch failed");

  // later work on each item's stream comes after the step
  if (hipEventRecord(E.done, bs) != hipSuccess) {
    (void)hipGetLastError();
    return fail_all(FFT3D_ELAUNCH, "step stream ordering failed");
  }
  for (hipStream_t s : streams)
    if (hipStreamWaitEvent(s, E.done, 0) != hipSuccess) {
      (void)hipGetLastError();
      return fail_all(FFT3D_ELAUNCH, "step stream ordering failed");
    }
  // one end-of-call event per plane (a plane has one item per round)
  std::vector<fft3d_plane*> ended;
  for (int k : all_items) {
    fft3d_plane* p = items[k].plane;
    if (items[k].status != FFT3D_OK || std::find(ended.begin(), ended.end(), p) != ended.end())
      continue;
    ended.push_back(p);
    p->end_call(bs);
  }
}

}  // namespace

extern "C" {

void fft3d_params_default(fft3d_params* p) {
  fft3d::Script s;
  p->sigma = s.sigma; p->beta = s.beta; p->bw = s.bw; p->bh = s.bh; p->bt = s.bt;
  p->sharpen = s.sharpen; p->plane = s.plane; p->mode = s.mode; p->bordersize = s.bordersize;
  p->precision = s.precision; p->degrid = s.degrid; p->scutoff = s.scutoff; p->svr = s.svr;
  p->smin = s.smin; p->smax = s.smax; p->kratio = s.kratio; p->ow = s.ow; p->oh = s.oh;
  p->wintype = s.wintype; p->interlaced = s.interlaced; p->sigma2 = s.sigma2;
  p->sigma3 = s.sigma3; p->sigma4 = s.sigma4; p->oldfft = s.oldfft;
}

int fft3d_dev_init(int device, char* err, size_t errlen) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    set_err(err, errlen, "FFT3DCL: no HIP GPU device available");
    return FFT3D_EDEVICE;
  }
  if (device < 0 || device >= n) {
    set_err(err, errlen, "FFT3DCL: HIP device index out of range");
    return FFT3D_EDEVICE;
  }
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    set_err(err, errlen, "FFT3DCL: hipSetDevice failed");
    return FFT3D_EDEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    (void)hipGetLastError();
    set_err(err, errlen, "FFT3DCL: hipGetDeviceProperties failed");
    return FFT3D_EDEVICE;
  }
  // set-up of the no-alloc entry point's scratch on this device
  if (!scratch_address(device)) {
    set_err(err, errlen, "FFT3DCL: hipGetSymbolAddress of the scratch array failed");
    return FFT3D_EDEVICE;
  }
#if !defined(FFT3D_CPU_EMU)
  // the code objects are built for gfx950 only (gcnArchName "gfx950:sramecc+:xnack-")
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0 ||
      (prop.gcnArchName[6] != '\0' && prop.gcnArchName[6] != ':')) {
    set_err(err, errlen, std::string("FFT3DCL: HIP device ") + std::to_string(device) + " is " +
                             prop.gcnArchName + "; this library is built for gfx950 (MI355X)");
    return FFT3D_EDEVICE;
  }
#endif
  return FFT3D_OK;
}

int fft3d_plane_create(const fft3d_params* sp, const fft3d_plane_desc* desc, fft3d_plane** out,
                       char* err, size_t errlen) {
  if (out) *out = nullptr;
  if (!sp || !desc || !out) {
    set_err(err, errlen, "FFT3DCL: null argument");
    return FFT3D_EINVAL;
  }
  if (desc->width <= 0 || desc->height <= 0) {
    set_err(err, errlen, "FFT3DCL: empty plane");
    return FFT3D_EINVAL;
  }
  if (desc->sample_type == 1 ? desc->bits != 32 : (desc->bits < 8 || desc->bits > 16)) {
    set_err(err, errlen, desc->sample_type == 1
                             ? "FFT3DCL: only 32-bit float clips supported (16-bit float not ported yet)"
                             : "FFT3DCL: integer bit depths 8-16 only");
    return FFT3D_EINVAL;
  }
  fft3d::Script s;
  s.sigma = sp->sigma; s.beta = sp->beta; s.bw = sp->bw; s.bh = sp->bh; s.bt = sp->bt;
  s.sharpen = sp->sharpen; s.plane = sp->plane; s.mode = sp->mode; s.bordersize = sp->bordersize;
  s.precision = sp->precision; s.degrid = sp->degrid; s.scutoff = sp->scutoff; s.svr = sp->svr;
  s.smin = sp->smin; s.smax = sp->smax; s.kratio = sp->kratio; s.ow = sp->ow; s.oh = sp->oh;
  s.wintype = sp->wintype; s.interlaced = sp->interlaced; s.sigma2 = sp->sigma2;
  s.sigma3 = sp->sigma3; s.sigma4 = sp->sigma4; s.oldfft = sp->oldfft;

  auto pl = new fft3d_plane();
  pl->d = *desc;
  std::string why = fft3d::validate(s, desc->width, desc->height, desc->bits, pl->p);
  if (!why.empty()) {
    delete pl;
    set_err(err, errlen, why);
    return FFT3D_EINVAL;
  }
  const fft3d::Params& p = pl->p;
  pl->ks = kernels_for(p.bw, p.bh);
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess) pl->dev = dev;
    if (hipGetDeviceProperties(&prop, pl->dev) == hipSuccess)
      pl->ncu = prop.multiProcessorCount;
    // the MFMA analysis instances share one resource footprint
    int n = 0;
    if (pl->ks.analysisM[0][0] &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(pl->ks.analysisM[0][0]), 256, 0) == hipSuccess)
      pl->mfmaPerCU = n;
    (void)hipGetLastError();
  }
  if (!pl->ks.analysis) {
    delete pl;
    set_err(err, errlen,
            "FFT3DCL: GPU setup failed: block too large for device local memory (max 64x64)");
    return FFT3D_EUNSUPPORTED;
  }
  ++g_live;
  const int W = desc->width, H = desc->height, bw = p.bw, bh = p.bh;
  fft3d::Grid grid = fft3d::Grid::of(p);
  pl->nb = grid.nb;
  pl->ncols = grid.ncols;
  pl->nrows = grid.nrows;
  pl->fbw = bw / 2 + 1;
  pl->bps = desc->sample_type == 1 ? 4 : (desc->bits <= 8 ? 1 : 2);
  pl->pix = desc->sample_type == 1 ? 0 : (desc->bits <= 8 ? 8 : 16);
  // precision 0 and 1 (8-bit clips only: deeper ones are forced to 2,
  // src/params.h:116-120) store the spectra ring as FP16, the original
  // FFT3dGPU's storage at those settings; 2 keeps FP32. Arithmetic is FP32
  // either way. Block shapes without FP16-ring kernels keep FP32.
  pl->half = p.precision < 2 && pl->ks.analysisH != nullptr;
  pl->specBytes = (pl->half ? 4 : 8) * (size_t)pl->nb * bh * (bw / 2);
  pl->t = fft3d::buildTables(p);

  float chromaNorm = 0.0f;
  if (p.mode == 1 && desc->chroma) {
    if (desc->sample_type == 1)
      chromaNorm = 0.0f;
    else if (desc->bits <= 8)
      chromaNorm = 128.0f / 255.0f;
    else
      chromaNorm = 32768.0f / 65535.0f;
  }

  // one device table block: 3 windows, pattern, sharpen, degrid, separable factors
  const size_t nwin = (size_t)bw * bh, nbin = (size_t)pl->fbw * bh;
  const size_t sep = 3 * nwin + 4 * nbin;
  const size_t fac = (sep + 2 * bw + 2 * bh + 3) / 4 * 4;  // float4 aligned
  // twiddles of the LDS FFTs (sides above 64): (cos, sin)(2 pi k / nmax)
  const int nmax = std::max(bw, bh);
  const size_t twOff = fac + 4 * (size_t)bh + 4 * (size_t)pl->fbw;
  std::vector<float> host(twOff + (size_t)nmax, 0.0f);
  for (int k = 0; k < nmax / 2; k++) {
    const double a = 2.0 * 3.14159265358979323846 * k / nmax;
    host[twOff + 2 * k] = (float)std::cos(a);
    host[twOff + 2 * k + 1] = (float)std::sin(a);
  }
  std::copy(pl->t.sepRow.begin(), pl->t.sepRow.end(), host.begin() + fac);
  std::copy(pl->t.sepCol.begin(), pl->t.sepCol.end(), host.begin() + fac + 4 * bh);
  std::copy(pl->t.anaX.begin(), pl->t.anaX.end(), host.begin() + sep);
  std::copy(pl->t.anaY.begin(), pl->t.anaY.end(), host.begin() + sep + bw);
  std::copy(pl->t.synX.begin(), pl->t.synX.end(), host.begin() + sep + bw + bh);
  std::copy(pl->t.synY.begin(), pl->t.synY.end(), host.begin() + sep + 2 * bw + bh);
  std::copy(pl->t.winAna.begin(), pl->t.winAna.end(), host.begin());
  std::copy(pl->t.winSynA.begin(), pl->t.winSynA.end(), host.begin() + nwin);
  std::copy(pl->t.winSynB.begin(), pl->t.winSynB.end(), host.begin() + 2 * nwin);
  std::copy(pl->t.pattern.begin(), pl->t.pattern.end(), host.begin() + 3 * nwin);
  std::copy(pl->t.sharpen.begin(), pl->t.sharpen.end(), host.begin() + 3 * nwin + nbin);
  std::copy(pl->t.degrid.begin(), pl->t.degrid.end(), host.begin() + 3 * nwin + 2 * nbin);

  bool ok = hipStreamCreateWithFlags(&pl->own, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&pl->lastCall, hipEventDisableTiming) == hipSuccess;
  ok = ok && pl->alloc(reinterpret_cast<void**>(&pl->dTables), host.size() * sizeof(float));
  ok = ok && pl->alloc(&pl->dStage, (size_t)W * H * pl->bps);
  ok = ok && pl->alloc(reinterpret_cast<void**>(&pl->dBlocks), sizeof(float) * (size_t)pl->nb * bw * bh);
  if (ok && p.bm == 0) {
    // FP16 pairs with the FP16 ring on integer samples (PlaneArgs::khalf)
    pl->kalmanBytes = (pl->half && pl->pix != 0 ? 8 : 16) * (size_t)pl->nb * nbin;
    ok = pl->alloc(reinterpret_cast<void**>(&pl->dKalman), pl->kalmanBytes);
  }
  pl->sweepOk = pl->ks.sweep && p.mode == 1 && p.ow * 2 == bw && p.oh * 2 == bh &&
                sepMatches(pl->t, p);
  if (ok && pl->sweepOk) {
    ok = pl->alloc(reinterpret_cast<void**>(&pl->dPart),
                   sizeof(float) * 4 * (size_t)((W + 3) & ~3) * H);
    ok = ok && pl->layoutSweep();
  }
  if (!ok) {
    (void)hipGetLastError();
    delete pl;
    --g_live;
    set_err(err, errlen, "FFT3DCL: GPU setup failed: GPU alloc failed");
    return FFT3D_ENOMEM;
  }
  if (hipMemcpy(pl->dTables, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipGetLastError();
    delete pl;
    --g_live;
    set_err(err, errlen, "FFT3DCL: GPU setup failed: table upload failed");
    return FFT3D_ELAUNCH;
  }

  fk::PlaneArgs& a = pl->args;
  a.W = W;
  a.H = H;
  a.bw = bw;
  a.bh = bh;
  a.ow = p.ow;
  a.oh = p.oh;
  a.mode = p.mode;
  a.nx = p.nx;
  a.ny = p.ny;
  a.ncols = pl->ncols;
  a.nrows = pl->nrows;
  a.nb = pl->nb;
  a.pix = pl->pix;
  a.khalf = (pl->half && pl->pix != 0) ? 1 : 0;
  a.inNorm = (pl->pix == 16) ? (float)(1u << (16 - desc->bits)) / 65535.0f : 1.0f;
  a.invNorm = (pl->pix == 16) ? 1.0f / (float)(1u << (16 - desc->bits)) : 1.0f;
  a.maxPix = (pl->pix == 16) ? (float)((1u << desc->bits) - 1) : 0.0f;
  a.chromaNorm = chromaNorm;
  a.bm = p.bm;
  a.dg = p.degridOn ? 1 : 0;
  a.pat = p.usePattern ? 1 : 0;
  a.shp = p.sharpenOn ? 1 : 0;
  a.sigmaX = p.beta * p.wienerS;
  a.thrX = fk::wiener_threshold(a.sigmaX);
  a.sigmaY = p.wienerS;
  a.betaX = p.betaX;
  a.beta = p.beta;
  a.ssmin = p.ssmin;
  a.ssmax = p.ssmax;
  a.kalmanCN = p.kalmanCN;
  a.kr2 = p.kratio * p.kratio;
  a.winAna = pl->dTables;
  a.winSynA = p.mode == 1 ? pl->dTables + nwin : pl->dTables;
  a.winSynB = p.mode == 1 ? pl->dTables + 2 * nwin : pl->dTables;
  a.pattern = pl->dTables + 3 * nwin;
  a.sharpen = pl->dTables + 3 * nwin + nbin;
  a.gs = reinterpret_cast<const fk::c32*>(pl->dTables + 3 * nwin + 2 * nbin);
  a.anaX = pl->dTables + sep;
  a.anaY = pl->dTables + sep + bw;
  a.synX = pl->dTables + sep + bw + bh;
  a.synY = pl->dTables + sep + 2 * bw + bh;
  a.sepRow = reinterpret_cast<const float4*>(pl->dTables + fac);
  a.sepCol = reinterpret_cast<const float4*>(pl->dTables + fac + 4 * bh);
  a.tw = reinterpret_cast<const float2*>(pl->dTables + twOff);
  *out = pl;
  return FFT3D_OK;
}

void fft3d_plane_destroy(fft3d_plane* p) {
  if (!p) return;
  // Work queued on caller streams may still use the plane's buffers. Every
  // call ends with lastCall, and a call on a new stream first waits for the
  // previous one, so lastCall covers all of them; the upload stream and the
  // plane's own stream are drained too. Other streams' work is not waited for.
  if (p->anyCall) (void)hipEventSynchronize(p->lastCall);
  if (p->up) (void)hipStreamSynchronize(p->up);
  if (p->own) (void)hipStreamSynchronize(p->own);
  (void)hipGetLastError();
  --g_live;
  delete p;
}

const char* fft3d_plane_error(const fft3d_plane* p) { return p ? p->error() : ""; }

int fft3d_plane_has(const fft3d_plane* p, int fn) { return p && p->ring.count(fn) ? 1 : 0; }

int fft3d_plane_forward(fft3d_plane* p, int fn, const void* src, size_t bytes, int src_on_device,
                        void* stream) {
  if (!p || !src) return FFT3D_EINVAL;
  const size_t need = (size_t)p->d.width * p->d.height * p->bps;
  if (bytes < need) return p->fail(FFT3D_EINVAL, "plane buffer too small");
  hipStream_t s = p->stream(stream);
  if (!p->order_call(s)) return p->fail(FFT3D_ELAUNCH, "stream wait for the previous call failed");
  fk::c32* spec = p->spectrum(fn, s);
  if (!spec) return p->fail(FFT3D_ENOMEM, "GPU spectra alloc failed");
  const void* plane = src;
  if (!src_on_device) {
    if (p->upload(p->dStage, src, need, s) != hipSuccess) {
      (void)hipGetLastError();
      p->ring.erase(fn);
      p->retire(spec);
      p->end_call(s);
      return p->fail(FFT3D_ELAUNCH, "plane upload failed");
    }
    plane = p->dStage;
  }
  p->launch_analysis(plane, spec, s);
  p->end_call(s);
  if (!p->launched("fft2d forward")) {
    p->ring.erase(fn);
    p->retire(spec);
    return FFT3D_ELAUNCH;
  }
  return FFT3D_OK;
}

int fft3d_plane_prefetch(fft3d_plane* p, int fn, const void* src, size_t bytes,
                         int src_on_device) {
  if (!p || !src) return FFT3D_EINVAL;
  const size_t need = (size_t)p->d.width * p->d.height * p->bps;
  if (bytes < need) return p->fail(FFT3D_EINVAL, "plane buffer too small");
  // carried out by the next process / forward_process call, after its launches
  p->pendFn = p->ring.count(fn) ? -1 : fn;
  p->pendSrc = src;
  p->pendOnDevice = src_on_device != 0;
  return FFT3D_OK;
}

int fft3d_plane_process(fft3d_plane* p, const int* window, int nwin, void* dst, size_t bytes,
                        int dst_on_device, void* stream) {
  if (!p || !window || !dst) return FFT3D_EINVAL;
  fk::WindowArgs w{};
  int rc = p->window(window, nwin, bytes, w);
  if (rc != FFT3D_OK) {
    p->pendFn = -1;
    return rc;
  }
  hipStream_t s = p->stream(stream);
  if (!p->order_call(s)) {
    p->pendFn = -1;
    return p->fail(FFT3D_ELAUNCH, "stream wait for the previous call failed");
  }
  p->begin_call(s);
  if (!p->wait_prefetched(window, nwin, s)) {
    p->pendFn = -1;
    (void)hipGetLastError();
    return p->fail(FFT3D_ELAUNCH, "wait for prefetched spectra failed");
  }
  rc = p->run(w, dst, dst_on_device, s, nullptr, nullptr, 0);
  p->pendFn = -1;  // a request the run did not reach is dropped with its buffer
  p->end_call(s);
  return rc;
}

int fft3d_plane_process_noalloc(fft3d_plane* p, const int* window, int nwin, void* dst,
                                size_t bytes, void* stream) {
  if (!p || !window || !dst) return FFT3D_EINVAL;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != p->dev) {
    (void)hipGetLastError();
    return p->fail_static(FFT3D_EINVAL, "no-alloc process: the plane's device is not current");
  }
  const bool sweep = p->sweepOk && p->sweepOn;
  const size_t need = sweep ? sizeof(float) * 4 * (size_t)p->sw.ppitch * p->d.height
                            : sizeof(float) * (size_t)p->nb * p->p.bw * p->p.bh;
  if (need > kScratchBytes)
    return p->fail_static(FFT3D_EINVAL,
                          "no-alloc process: the plane needs more scratch than the 128 MiB "
                          "array (4 partial planes of W x H floats on the sweep path, "
                          "nb*bw*bh floats off it)");
  if (sweep && p->bandRowsFor() != p->seamR)
    return p->fail_static(FFT3D_EINVAL, "no-alloc process: the sweep's seam list is out of date");
  float* scr = static_cast<float*>(scratch_address(cur));
  if (!scr) return p->fail_static(FFT3D_EDEVICE, "no-alloc process: no scratch address");
  fk::WindowArgs w{};
  int rc = p->window_args(window, nwin, bytes, w);
  if (rc != FFT3D_OK) return rc;
  hipStream_t s = p->stream(stream);
  if (!p->order_call(s))
    return p->fail_static(FFT3D_ELAUNCH, "stream wait for the previous call failed");
  if (!p->wait_prefetched(window, nwin, s)) {
    (void)hipGetLastError();
    return p->fail_static(FFT3D_ELAUNCH, "wait for prefetched spectra failed");
  }
  // a pending prefetch request stays pending for the next fft3d_plane_process
  const char* what = p->launch(w, dst, s, nullptr, nullptr, 0, scr, scr, false);
  p->end_call(s);
  if (what) return p->fail_static(FFT3D_ELAUNCH, "no-alloc process: a kernel launch failed");
  return FFT3D_OK;
}

int fft3d_plane_forward_process(fft3d_plane* p, int fn, const void* src, size_t src_bytes,
                                int src_on_device, const int* window, int nwin, void* dst,
                                size_t dst_bytes, int dst_on_device, void* stream) {
  if (!p || !src || !window || !dst) return FFT3D_EINVAL;
  int mask = 0;
  for (int i = 0; i < nwin && i < 4; i++)
    if (window[i] == fn) mask |= 1 << i;
  // a pending prefetch that this call does not reach is dropped
  auto bail = [p](int rc) {
    p->pendFn = -1;
    return rc;
  };
  if (!mask || !(p->sweepOk && p->sweepOn && p->fuseOn) || p->half || p->dft) {
    const int rc = fft3d_plane_forward(p, fn, src, src_bytes, src_on_device, stream);
    if (rc != FFT3D_OK) return bail(rc);
    return fft3d_plane_process(p, window, nwin, dst, dst_bytes, dst_on_device, stream);
  }
  const size_t need = (size_t)p->d.width * p->d.height * p->bps;
  if (src_bytes < need) return bail(p->fail(FFT3D_EINVAL, "plane buffer too small"));
  if (nwin != (p->p.bm >= 2 ? p->p.bt : 1))
    return bail(p->fail(FFT3D_EINVAL, "window length does not match bt"));
  if (dst_bytes < need) return bail(p->fail(FFT3D_EINVAL, "output buffer too small"));
  for (int i = 0; i < nwin; i++)
    if (window[i] != fn && !p->ring.count(window[i]))
      return bail(p->fail(FFT3D_EMISSING, "missing spectra in window"));
  hipStream_t s = p->stream(stream);
  if (!p->order_call(s)) return bail(p->fail(FFT3D_ELAUNCH, "stream wait for the previous call failed"));
  p->begin_call(s);
  const void* plane = src;
  if (!src_on_device) {
    if (!p->dStageIn && !p->alloc(&p->dStageIn, need))
      return bail(p->fail(FFT3D_ENOMEM, "GPU alloc failed (upload staging)"));
    if (p->upload(p->dStageIn, src, need, s) != hipSuccess) {
      (void)hipGetLastError();
      p->end_call(s);
      return bail(p->fail(FFT3D_ELAUNCH, "plane upload failed"));
    }
    plane = p->dStageIn;
  }
  fk::c32* spec = p->spectrum(fn, s);
  if (!spec) {
    p->end_call(s);
    return bail(p->fail(FFT3D_ENOMEM, "GPU spectra alloc failed"));
  }
  fk::WindowArgs w{};
  int rc = p->window(window, nwin, dst_bytes, w);
  if (rc == FFT3D_OK && !p->wait_prefetched(window, nwin, s)) {
    (void)hipGetLastError();
    rc = p->fail(FFT3D_ELAUNCH, "wait for prefetched spectra failed");
  }
  if (rc == FFT3D_OK) rc = p->run(w, dst, dst_on_device, s, plane, spec, mask);
  p->pendFn = -1;
  p->end_call(s);
  if (rc != FFT3D_OK) {
    // the spectrum may be incomplete: drop it, as a failed forward does
    p->ring.erase(fn);
    p->retire(spec);
  }
  return rc;
}

int fft3d_step(fft3d_work_item* items, int n) {
  if (!items || n < 0) return FFT3D_EINVAL;
  // Round r holds each plane's r-th item. A plane whose items are all
  // device-resident, on the fused sweep of the current device, goes through
  // the batched launches (step_batch); the others run item by item
  // (step_item), in order. The two groups share no plane.
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) (void)hipGetLastError();
  std::map<const fft3d_plane*, std::pair<int, bool>> pl;  // items so far, all batchable
  std::vector<int> round(n, 0);
  for (int k = 0; k < n; k++) {
    const bool b = step_item_valid(items[k]) && step_batchable(items[k], cur);
    auto& e = pl.emplace(items[k].plane, std::make_pair(0, true)).first->second;
    round[k] = e.first++;
    e.second = e.second && b;
  }
  std::vector<std::vector<int>> rounds;
  for (int k = 0; k < n; k++) {
    if (!items[k].plane || !pl[items[k].plane].second) {
      items[k].status = step_item(items[k]);
      continue;
    }
    if ((int)rounds.size() <= round[k]) rounds.resize(round[k] + 1);
    rounds[round[k]].push_back(k);
  }
  if (!rounds.empty()) step_batch(items, rounds);
  int first = FFT3D_OK;
  for (int k = 0; k < n; k++)
    if (items[k].status != FFT3D_OK && first == FFT3D_OK) first = items[k].status;
  return first;
}

void fft3d_plane_prune(fft3d_plane* p, int lo, int hi) {
  if (!p) return;
  if (p->pendFn >= 0 && !((lo < 0 || p->pendFn >= lo) && (hi < 0 || p->pendFn <= hi)))
    p->pendFn = -1;
  for (auto it = p->ring.begin(); it != p->ring.end();) {
    const bool keep = (lo < 0 || it->first >= lo) && (hi < 0 || it->first <= hi);
    if (keep) {
      ++it;
    } else {
      p->forget(it->first);
      p->retire(it->second);
      it = p->ring.erase(it);
    }
  }
}

void fft3d_plane_clear(fft3d_plane* p) {
  if (!p) return;
  for (auto& kv : p->ring) p->retire(kv.second);
  p->ring.clear();
  p->forget_all();
  p->pendFn = -1;
}

int fft3d_plane_set_option(fft3d_plane* p, int option, int value) {
  if (!p) return FFT3D_EINVAL;
  switch (option) {
  case FFT3D_OPT_SWEEP:
    if (value != 0 && value != 1) return p->fail(FFT3D_EINVAL, "sweep option must be 0 or 1");
    p->sweepOn = value == 1;
    return FFT3D_OK;
  case FFT3D_OPT_BAND_ROWS:
    if (value < 0) return p->fail(FFT3D_EINVAL, "band rows must be >= 0");
    p->bandRows = value;
    // the seam list follows at once, so process calls never rebuild it
    if (p->sweepOk && !p->layoutSweep()) return p->fail(FFT3D_ENOMEM, "GPU alloc failed (seam list)");
    return FFT3D_OK;
  case FFT3D_OPT_SPECTRA_FP16: {
    if (value != 0 && value != 1)
      return p->fail(FFT3D_EINVAL, "FP16 spectra option must be 0 or 1");
    if (value == 1 && !p->ks.analysisH)
      return p->fail(FFT3D_EUNSUPPORTED, "FP16 spectra need square blocks of 16, 32 or 64");
    if ((value == 1) == p->half) return FFT3D_OK;
    // the stored spectra change size: drop the ring and the pool
    if (hipStreamSynchronize(p->own) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      (void)hipGetLastError();
      return p->fail(FFT3D_ELAUNCH, "synchronize failed");
    }
    for (auto& kv : p->ring) p->release(kv.second);
    for (auto* m : p->pool) p->release(m);
    p->ring.clear();
    p->pool.clear();
    p->forget_all();
    p->pendFn = -1;
    p->half = value == 1;
    p->specBytes = (p->half ? 4 : 8) * (size_t)p->nb * p->p.bh * (p->p.bw / 2);
    // integer samples keep the Kalman state as FP16 pairs with the FP16 ring
    // (8 instead of 16 bytes per bin; float samples keep FP32 state, whose
    // FP16 rounding would exceed the float tolerance over a long clip)
    p->args.khalf = (p->half && p->pix != 0) ? 1 : 0;
    if (p->dKalman) {
      // the state changes format: the recursion starts again at the next frame
      (void)hipFree(p->dKalman);
      p->bytes -= p->kalmanBytes;
      g_bytes -= p->kalmanBytes;
      p->dKalman = nullptr;
      p->kalmanBytes = (p->args.khalf ? 8 : 16) * (size_t)p->nb * p->fbw * p->p.bh;
      p->kalmanFirst = true;
      if (!p->alloc(reinterpret_cast<void**>(&p->dKalman), p->kalmanBytes))
        return p->fail(FFT3D_ENOMEM, "GPU alloc failed (Kalman state)");
    }
    return FFT3D_OK;
  }
  case FFT3D_OPT_RESERVE_SPECTRA:
    if (value < 0 || value > 64) return p->fail(FFT3D_EINVAL, "reserve count must be 0..64");
    while ((int)(p->pool.size() + p->ring.size()) < value) {
      fk::c32* m = nullptr;
      if (!p->alloc(reinterpret_cast<void**>(&m), p->specBytes))
        return p->fail(FFT3D_ENOMEM, "GPU spectra alloc failed");
      p->pool.push_back(m);
    }
    return FFT3D_OK;
  case FFT3D_OPT_DFT: {
    if (value < 0 || value > 2) return p->fail(FFT3D_EINVAL, "DFT form must be 0, 1 or 2");
    if (value && (!p->ks.analysisM[value - 1][0] || !p->sweepOk))
      return p->fail(FFT3D_EUNSUPPORTED,
                     "the MFMA DFT form needs 32x32 blocks on the fused sweep (mode 1, "
                     "ow = bw/2, oh = bh/2)");
    if (value && !p->dDft) {
      const std::vector<uint16_t> t = fft3d::dft_operands32(1);
      const size_t n = t.size() * sizeof(uint16_t);
      if (!p->alloc(&p->dDft, 2 * n)) return p->fail(FFT3D_ENOMEM, "GPU alloc failed (DFT table)");
      const std::vector<uint16_t> tb = fft3d::dft_operands32(2);
      if (hipMemcpy(p->dDft, t.data(), n, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(static_cast<char*>(p->dDft) + n, tb.data(), n, hipMemcpyHostToDevice) !=
              hipSuccess) {
        (void)hipGetLastError();
        return p->fail(FFT3D_ELAUNCH, "DFT table upload failed");
      }
    }
    p->dft = value;
    const size_t slots = (size_t)fk::mm::kSlots * 64;  // uint4 per table
    p->args.dft = value ? reinterpret_cast<const uint4*>(p->dDft) + (value - 1) * slots : nullptr;
    return FFT3D_OK;
  }
  case FFT3D_OPT_SWEEP_OCC:
    if (value != 0 && (value < 2 || value > 4))
      return p->fail(FFT3D_EINVAL, "sweep occupancy must be 0 (tuned) or 2..4");
    if (value > p->ks.maxOcc)
      return p->fail(FFT3D_EUNSUPPORTED, "no sweep kernel for that occupancy at this block size");
    p->occOpt = value;
    if (p->sweepOk && !p->layoutSweep()) return p->fail(FFT3D_ENOMEM, "GPU alloc failed (seam list)");
    return FFT3D_OK;
  case FFT3D_OPT_FUSED_ANALYSIS:
    if (value != 0 && value != 1)
      return p->fail(FFT3D_EINVAL, "fused analysis option must be 0 or 1");
    p->fuseOn = value == 1;
    return FFT3D_OK;
  case FFT3D_OPT_HOST_STAGING:
    if (value != 0 && value != 1)
      return p->fail(FFT3D_EINVAL, "host staging option must be 0 or 1");
    p->stagingOn = value == 1;
    return FFT3D_OK;
  case FFT3D_OPT_STAGING_CHUNK: {
    if (value != 0 && (value < 256 || value > (256 << 20) || value % 64 != 0))
      return p->fail(FFT3D_EINVAL, "staging chunk must be 0 or a multiple of 64 in 256..256 MiB");
    const size_t chunk = value ? (size_t)value : fx::Staging::kDefaultChunk;
    if (chunk == p->stg.chunk) return FFT3D_OK;
    // queued DMAs may still use the old chunks
    if (p->stg.ready() && hipDeviceSynchronize() != hipSuccess) {
      (void)hipGetLastError();
      return p->fail(FFT3D_ELAUNCH, "synchronize failed");
    }
    p->stg.release();
    p->stgUp.release();
    p->stg.chunk = chunk;
    return FFT3D_OK;
  }
  }
  return p->fail(FFT3D_EINVAL, "unknown option");
}

int fft3d_plane_set_profiling(fft3d_plane* p, int on) {
  if (!p) return FFT3D_EINVAL;
  p->prof = on != 0;
  return FFT3D_OK;
}

int fft3d_plane_get_profile(fft3d_plane* p, fft3d_kernel_time* out, int max) {
  if (!p) return FFT3D_EINVAL;
  static const char* names[fft3d_plane::K_COUNT] = {"analysis", "filter_inverse", "synthesize",
                                                     "filter_synth_sweep", "combine_seams",
                                                     "forward_filter_synth_sweep"};
  double ms[fft3d_plane::K_COUNT] = {};
  int n[fft3d_plane::K_COUNT] = {};
  for (auto& r : p->recs) {
    float t = 0.0f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      ms[r.kind] += t;
      n[r.kind]++;
    } else {
      (void)hipGetLastError();
    }
    p->evpool.push_back(r.a);
    p->evpool.push_back(r.b);
  }
  p->recs.clear();
  int k = 0;
  for (int i = 0; i < fft3d_plane::K_COUNT && k < max; i++) {
    if (!n[i]) continue;
    if (out) {
      std::snprintf(out[k].name, sizeof(out[k].name), "%s", names[i]);
      out[k].launches = n[i];
      out[k].total_ms = ms[i];
    }
    k++;
  }
  return k;
}

int fft3d_live_instances(void) { return g_live.load(); }
size_t fft3d_live_bytes(void) { return g_bytes.load(); }

int fft3d_plane_get_info(const fft3d_plane* p, fft3d_plane_info* i) {
  if (!p || !i) return FFT3D_EINVAL;
  std::memset(i, 0, sizeof(*i));
  i->nb = p->nb;
  i->bw = p->p.bw;
  i->bh = p->p.bh;
  i->ow = p->p.ow;
  i->oh = p->p.oh;
  i->mode = p->p.mode;
  i->fbw = p->fbw;
  i->bm = p->p.bm;
  i->bt = p->p.bt;
  i->current = p->p.current;
  i->ncols = p->ncols;
  i->nrows = p->nrows;
  i->blocks_per_wg = p->ks.G;
  i->spectrum_bytes = p->specBytes;
  i->device_bytes = p->bytes;
  i->cached_frames = (int)p->ring.size();
  i->dft = p->dft;
  i->spectra_fp16 = p->half ? 1 : 0;
  i->band_rows = p->sweepOk ? p->bandRowsFor() : 0;
  i->sweep_occupancy = p->sweepOk ? p->occupancy() : 0;
  i->tuned = p->tuned() ? 1 : 0;
  return FFT3D_OK;
}

int fft3d_plane_get_table(const fft3d_plane* p, int which, float* dst) {
  if (!p) return FFT3D_EINVAL;
  const std::vector<float>* v = nullptr;
  switch (which) {
  case 0: v = &p->t.winAna; break;
  case 1: v = &p->t.winSynA; break;
  case 2: v = &p->t.winSynB; break;
  case 3: v = &p->t.pattern; break;
  case 4: v = &p->t.sharpen; break;
  case 5: v = &p->t.degrid; break;
  case 6: v = &p->t.sepRow; break;
  case 7: v = &p->t.sepCol; break;
  default: return FFT3D_EINVAL;
  }
  if (dst && !v->empty()) std::memcpy(dst, v->data(), v->size() * sizeof(float));
  return (int)v->size();
}

int fft3d_plane_read_spectrum(fft3d_plane* p, int fn, float* dst) {
  if (!p || !dst) return FFT3D_EINVAL;
  auto it = p->ring.find(fn);
  if (it == p->ring.end()) return p->fail(FFT3D_EMISSING, "missing spectra");
  const int bw = p->p.bw, bh = p->p.bh, nc = bw / 2, fbw = p->fbw;
  std::vector<fk::c32> h((size_t)p->nb * bh * nc);
  std::vector<uint32_t> hh(p->half ? h.size() : 0);
  void* host = p->half ? static_cast<void*>(hh.data()) : static_cast<void*>(h.data());
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(host, it->second, p->specBytes, hipMemcpyDeviceToHost) != hipSuccess) {
    (void)hipGetLastError();
    return p->fail(FFT3D_ELAUNCH, "spectrum readback failed");
  }
  for (size_t i = 0; i < hh.size(); i++)
    h[i] = fk::c32{half_to_float(hh[i] & 0xffffu), half_to_float(hh[i] >> 16)};
  for (int b = 0; b < p->nb; b++) {
    const fk::c32* s = &h[(size_t)b * bh * nc];
    float* d = dst + (size_t)b * bh * fbw * 2;
    for (int r = 0; r < bh; r++) {
      const int r2 = (bh - r) % bh;
      for (int c = 1; c < nc; c++) {
        d[2 * (r * fbw + c)] = s[r * nc + c].x;
        d[2 * (r * fbw + c) + 1] = s[r * nc + c].y;
      }
      const fk::c32 wa = s[r * nc], wb = s[r2 * nc];
      d[2 * (r * fbw)] = 0.5f * (wa.x + wb.x);
      d[2 * (r * fbw) + 1] = 0.5f * (wa.y - wb.y);
      d[2 * (r * fbw + nc)] = 0.5f * (wa.y + wb.y);
      d[2 * (r * fbw + nc) + 1] = -0.5f * (wa.x - wb.x);
    }
  }
  return FFT3D_OK;
}

} // extern "C"
