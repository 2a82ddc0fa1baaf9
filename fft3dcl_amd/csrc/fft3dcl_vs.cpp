// VapourSynth API 4 plugin: fft3dcl.FFT3DCL(clip, ...) on the MI355X backend.
//
// Drop-in for the reference plugin's public surface (src/fft3dcl.cpp):
// identifier, namespace, function name and argument string (:540-552),
// argument defaults (:159-198), format checks and plane routing (:143-225),
// the clamped temporal window requested at arInitial (:319-331), spectra
// ring management and pruning (:388-467), untouched-plane copy (:346-363),
// the per-instance mutex (:366), filter mode by bt (:286-292), node ownership
// (:127-142, :528-534), logging to the core and to FFT3DCL_LOG (:67-90).
//
// The compute goes through the C ABI of include/fft3d_hip.h. This build ships
// no CPU backend: device=0 is rejected at create time, and device=2 (GPU with
// CPU fallback in the reference) behaves as device=1: a GPU failure fails the
// frame with the reason instead of falling back.
#include "fft3d_hip.h"
#include "vs4_abi.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace {

std::once_flag g_devOnce;
int g_devRc = FFT3D_EDEVICE;
char g_devErr[256] = "";

void initDevice() { g_devRc = fft3d_dev_init(0, g_devErr, sizeof(g_devErr)); }

struct Inst {
  VSNode* clip = nullptr;
  VSVideoInfo vi{};
  int bm = 3, bt = 3, current = 1;
  int device = 2;
  std::vector<int> planes;              // processed plane indices
  std::vector<fft3d_plane*> gpu;        // parallel to planes
  std::vector<uint8_t> staging;         // input plane whose rows are padded
  std::vector<uint8_t> stagingOut;      // output plane whose rows are padded
  std::vector<uint8_t> stagingNext;     // prefetched plane whose rows are padded
  bool prefetch = false;                // FFT3DCL_PREFETCH=1 (bt >= 2 only)
  std::mutex mtx;
  ~Inst() {
    for (auto* p : gpu) fft3d_plane_destroy(p);
  }
};

void logMsg(VSCore* core, const VSAPI* vsapi, int type, const char* msg) {
  vsapi->logMessage(type, msg, core);
  const char* path = std::getenv("FFT3DCL_LOG");
  if (path && *path) {
    if (FILE* f = std::fopen(path, "a")) {
      char ts[16] = {};
      std::time_t now = std::time(nullptr);
      std::strftime(ts, sizeof(ts), "%H:%M:%S", std::localtime(&now));
      std::fprintf(f, "%s %s\n", ts, msg);
      std::fclose(f);
    }
  }
}

void logUsage(VSCore* core, const VSAPI* vsapi, const char* what) {
  char msg[160];
  std::snprintf(msg, sizeof(msg),
                "FFT3DCL: %s - %d GPU instance(s) alive, %.0f MiB of device buffers", what,
                fft3d_live_instances(), (double)fft3d_live_bytes() / 1048576.0);
  logMsg(core, vsapi, mtInformation, msg);
}

int clampFrame(int n, int numFrames) { return n < 0 ? 0 : (n > numFrames - 1 ? numFrames - 1 : n); }

// With FFT3DCL_PREFETCH=1, the frame a sequential request for n+1 adds to the
// window is requested along with n's window and prefetched (-1: none).
int nextNewFrame(const Inst* d, int n) {
  if (!d->prefetch || d->bm < 2) return -1;
  const int f = n - d->current + d->bt;
  return f >= 0 && f < d->vi.numFrames ? f : -1;
}

const VSFrame* getFrame(int n, int reason, void* instanceData, void** /*frameData*/,
                        VSFrameContext* ctx, VSCore* core, const VSAPI* vsapi) {
  Inst* d = static_cast<Inst*>(instanceData);
  const VSVideoInfo& vi = d->vi;
  if (reason == arInitial) {
    if (d->bm >= 2) {
      const int lo = clampFrame(n - d->current, vi.numFrames);
      const int hi = clampFrame(n - d->current + d->bt - 1, vi.numFrames);
      for (int f = lo; f <= hi; f++) vsapi->requestFrameFilter(f, d->clip, ctx);
      const int nx = nextNewFrame(d, n);
      if (nx > hi) vsapi->requestFrameFilter(nx, d->clip, ctx);
    } else {
      vsapi->requestFrameFilter(clampFrame(n, vi.numFrames), d->clip, ctx);
    }
    return nullptr;
  }
  if (reason != arAllFramesReady) return nullptr;

  const VSFrame* src = vsapi->getFrameFilter(n, d->clip, ctx);
  VSFrame* dst = vsapi->newVideoFrame(&vi.format, vi.width, vi.height, src, core);
  const int bps = vi.format.bytesPerSample;
  for (int p = 0; p < vi.format.numPlanes; p++) {
    bool processed = false;
    for (int q : d->planes) processed |= q == p;
    if (processed) continue;
    const uint8_t* sp = vsapi->getReadPtr(src, p);
    uint8_t* dp = vsapi->getWritePtr(dst, p);
    const ptrdiff_t ss = vsapi->getStride(src, p), ds = vsapi->getStride(dst, p);
    const int h = vsapi->getFrameHeight(src, p);
    const size_t row = (size_t)vsapi->getFrameWidth(src, p) * bps;
    for (int y = 0; y < h; y++) std::memcpy(dp + y * ds, sp + y * ss, row);
  }

  std::lock_guard<std::mutex> lock(d->mtx);
  const int nx = nextNewFrame(d, n);
  const VSFrame* frNext = nullptr;  // held until the process calls have returned
  for (size_t k = 0; k < d->planes.size(); k++) {
    const int p = d->planes[k];
    fft3d_plane* g = d->gpu[k];
    const int pw = vsapi->getFrameWidth(src, p), ph = vsapi->getFrameHeight(src, p);
    const size_t rowBytes = (size_t)pw * bps, planeBytes = rowBytes * ph;
    int rc = FFT3D_OK;
    int window[4], missing[4], nmiss = 0;
    const int nwin = d->bm >= 2 ? d->bt : 1;
    for (int i = 0; i < nwin; i++) {
      const int fn = clampFrame(d->bm >= 2 ? n - d->current + i : n, vi.numFrames);
      window[i] = fn;
      bool seen = false;
      for (int k = 0; k < nmiss; k++) seen = seen || missing[k] == fn;
      if (!seen && !fft3d_plane_has(g, fn)) missing[nmiss++] = fn;
    }
    // plane p of a source frame as packed rows: the frame itself when its
    // stride is the row size (src/fft3dcl.cpp:408-415), else a copy
    auto packed = [&](const VSFrame* fr) -> const uint8_t* {
      const uint8_t* sp = vsapi->getReadPtr(fr, p);
      const ptrdiff_t ss = vsapi->getStride(fr, p);
      if ((size_t)ss == rowBytes) return sp;
      d->staging.resize(planeBytes);
      for (int y = 0; y < ph; y++)
        std::memcpy(d->staging.data() + y * rowBytes, sp + y * ss, rowBytes);
      return d->staging.data();
    };
    // the next sequential request's new frame: the process call below uploads
    // and transforms it alongside its own work
    if (nx >= 0 && !fft3d_plane_has(g, nx)) {
      if (!frNext) frNext = vsapi->getFrameFilter(nx, d->clip, ctx);
      const uint8_t* sp = vsapi->getReadPtr(frNext, p);
      const ptrdiff_t ss = vsapi->getStride(frNext, p);
      if ((size_t)ss != rowBytes) {
        d->stagingNext.resize(planeBytes);
        for (int y = 0; y < ph; y++)
          std::memcpy(d->stagingNext.data() + y * rowBytes, sp + y * ss, rowBytes);
        sp = d->stagingNext.data();
      }
      (void)fft3d_plane_prefetch(g, nx, sp, planeBytes, 0);  // a failure only loses the overlap
    }
    // the missing frames but the last go through forward; the last one is
    // transformed by the filter pass itself (forward_process)
    for (int k = 0; k + 1 < nmiss && rc == FFT3D_OK; k++) {
      const VSFrame* fr = vsapi->getFrameFilter(missing[k], d->clip, ctx);
      rc = fft3d_plane_forward(g, missing[k], packed(fr), planeBytes, 0, nullptr);
      vsapi->freeFrame(fr);
    }
    if (rc == FFT3D_OK) {
      uint8_t* dp = vsapi->getWritePtr(dst, p);
      const ptrdiff_t ds = vsapi->getStride(dst, p);
      const bool direct = (size_t)ds == rowBytes;  // straight into the frame
      if (!direct) d->stagingOut.resize(planeBytes);
      uint8_t* out = direct ? dp : d->stagingOut.data();
      if (nmiss) {
        const int fn = missing[nmiss - 1];
        const VSFrame* fr = vsapi->getFrameFilter(fn, d->clip, ctx);
        rc = fft3d_plane_forward_process(g, fn, packed(fr), planeBytes, 0, window, nwin, out,
                                         planeBytes, 0, nullptr);
        vsapi->freeFrame(fr);
      } else {
        rc = fft3d_plane_process(g, window, nwin, out, planeBytes, 0, nullptr);
      }
      for (int y = 0; rc == FFT3D_OK && !direct && y < ph; y++)
        std::memcpy(dp + y * ds, out + y * rowBytes, rowBytes);
    }
    if (rc != FFT3D_OK) {
      std::string msg = "FFT3DCL: GPU processing failed at frame " + std::to_string(n) + ": " +
                        fft3d_plane_error(g) +
                        " (this build has no CPU backend to fall back to)";
      vsapi->setFilterError(msg.c_str(), ctx);
      if (frNext) vsapi->freeFrame(frNext);
      vsapi->freeFrame(dst);
      vsapi->freeFrame(src);
      return nullptr;
    }
    // keep what the next requests reuse: the current window plus one frame of
    // slack. One-frame modes reuse nothing; Kalman (bt=0) keeps its state in
    // the plane's Kalman buffer, not in the ring, so its ring is dropped too
    // (the reference never prunes it, src/fft3dcl.cpp:458-466, and grows by
    // one spectrum per frame).
    if (d->bm >= 2)
      fft3d_plane_prune(g, n - d->current - 1, n - d->current + d->bt);
    else
      fft3d_plane_clear(g);
  }
  if (frNext) vsapi->freeFrame(frNext);
  vsapi->freeFrame(src);
  return dst;
}

void freeFilter(void* instanceData, VSCore* core, const VSAPI* vsapi) {
  Inst* d = static_cast<Inst*>(instanceData);
  vsapi->freeNode(d->clip);
  delete d;
  logUsage(core, vsapi, "released");
}

void create(const VSMap* in, VSMap* out, void* /*userData*/, VSCore* core, const VSAPI* vsapi) {
  int err = 0;
  VSNode* clip = vsapi->mapGetNode(in, "clip", 0, &err);
  if (err || !clip) {
    vsapi->mapSetError(out, "FFT3DCL: a video clip must be specified as the 'clip' argument");
    return;
  }
  // owns the node reference until createVideoFilter takes it
  struct NodeGuard {
    const VSAPI* api;
    VSNode* node;
    bool owned = true;
    ~NodeGuard() {
      if (owned && node) api->freeNode(node);
    }
  } guard{vsapi, clip};
  const VSVideoInfo* vi = vsapi->getVideoInfo(clip);
  const VSVideoFormat& f = vi->format;
  if (f.colorFamily == cfUndefined || vi->width == 0 || vi->height == 0 ||
      (f.numPlanes != 1 && f.numPlanes != 3)) {
    vsapi->mapSetError(out, "FFT3DCL: only constant-format gray, YUV and planar RGB clips are supported");
    return;
  }
  if (f.sampleType == stFloat && f.bitsPerSample != 32) {
    vsapi->mapSetError(out, "FFT3DCL: only 32-bit float clips supported (16-bit float not ported yet)");
    return;
  }
  if (f.sampleType == stInteger && (f.bitsPerSample < 8 || f.bitsPerSample > 16)) {
    vsapi->mapSetError(out, "FFT3DCL: integer bit depths 8-16 only");
    return;
  }

  auto getF = [&](const char* key, float def) {
    int e = 0;
    double v = vsapi->mapGetFloat(in, key, 0, &e);
    return e ? def : (float)v;
  };
  auto getI = [&](const char* key, int64_t def) {
    int e = 0;
    int64_t v = vsapi->mapGetInt(in, key, 0, &e);
    return e ? def : v;
  };
  fft3d_params sp;
  sp.sigma = getF("sigma", 2.0f);
  sp.beta = getF("beta", 1.0f);
  sp.bw = (int)getI("bw", 32);
  sp.bh = (int)getI("bh", 32);
  sp.bt = (int)getI("bt", 3);
  sp.sharpen = getF("sharpen", 0.0f);
  sp.plane = (int)getI("plane", 0);
  sp.mode = (int)getI("mode", 1);
  sp.bordersize = (int)getI("bordersize", 1);
  sp.precision = (int)getI("precision", 0);
  sp.degrid = getF("degrid", sp.mode == 1 ? 1.0f : 0.0f);
  sp.scutoff = getF("scutoff", 0.3f);
  sp.svr = getF("svr", 1.0f);
  sp.smin = getF("smin", 4.0f);
  sp.smax = getF("smax", 20.0f);
  sp.kratio = getF("kratio", 2.0f);
  sp.ow = (int)getI("ow", sp.bw / 2);
  sp.oh = (int)getI("oh", sp.bh / 2);
  sp.wintype = (int)getI("wintype", 0);
  sp.interlaced = (int)getI("interlaced", 0);
  sp.sigma2 = getF("sigma2", sp.sigma);
  sp.sigma3 = getF("sigma3", sp.sigma);
  sp.sigma4 = getF("sigma4", sp.sigma);
  sp.oldfft = (int)getI("oldfft", -1);
  const int device = (int)getI("device", 2);

  int plane = sp.plane;
  if (f.colorFamily == cfRGB) plane = 4;
  if (f.colorFamily == cfGray) {
    if (plane == 1) {
      vsapi->mapSetError(out, "FFT3DCL: cannot set chroma processing for a greyscale clip");
      return;
    }
    if (plane == 4) plane = 0;
  }
  if (plane == 2 || plane == 3) plane = 1;
  if (plane != 0 && plane != 1 && plane != 4) {
    vsapi->mapSetError(out, "FFT3DCL: plane must be 0 (luma), 1 (chroma) or 4 (all)");
    return;
  }
  std::vector<int> planes;
  if (f.colorFamily == cfGray || plane == 0)
    planes = {0};
  else if (plane == 1)
    planes = {1, 2};
  else
    planes = {0, 1, 2};

  if (device == 0) {
    vsapi->mapSetError(out, "FFT3DCL: device=0 selects the CPU backend, which this HIP build "
                            "does not include (use device=1 or 2)");
    return;
  }
  std::call_once(g_devOnce, initDevice);
  if (g_devRc != FFT3D_OK) {
    vsapi->mapSetError(out, (std::string(g_devErr) + " (device=" + std::to_string(device) + ")").c_str());
    return;
  }

  auto d = std::make_unique<Inst>();
  d->clip = clip;
  d->vi = *vi;
  d->device = device;
  d->planes = planes;
  for (int p : planes) {
    fft3d_plane_desc desc;
    desc.width = vi->width >> (p ? f.subSamplingW : 0);
    desc.height = vi->height >> (p ? f.subSamplingH : 0);
    desc.bits = f.bitsPerSample;
    desc.sample_type = f.sampleType;
    desc.chroma = (f.colorFamily == cfYUV && p > 0) ? 1 : 0;
    fft3d_plane* g = nullptr;
    char e[512] = "";
    if (fft3d_plane_create(&sp, &desc, &g, e, sizeof(e)) != FFT3D_OK) {
      vsapi->mapSetError(out, e);
      return;
    }
    d->gpu.push_back(g);
  }
  fft3d_plane_info info;
  fft3d_plane_get_info(d->gpu[0], &info);
  d->bm = info.bm;
  d->bt = info.bt;
  d->current = info.current;
  // opt-in: request and prefetch the frame a sequential next request adds
  const char* pf = std::getenv("FFT3DCL_PREFETCH");
  d->prefetch = pf && std::strcmp(pf, "1") == 0;

  VSFilterDependency deps[] = {{clip, rpGeneral}};
  // Kalman (bt=0) carries state from frame to frame: frames in order, one at a time.
  const int fm = sp.bt == 0 ? fmFrameState : fmParallelRequests;
  Inst* raw = d.release();
  vsapi->createVideoFilter(out, "FFT3DCL", vi, getFrame, freeFilter, fm, deps, 1, raw, core);
  guard.owned = false;
  logUsage(core, vsapi, "created");
}

const char* const kArgs =
    "clip:vnode;"
    "sigma:float:opt;beta:float:opt;bw:int:opt;bh:int:opt;bt:int:opt;sharpen:float:opt;"
    "plane:int:opt;mode:int:opt;bordersize:int:opt;precision:int:opt;degrid:float:opt;"
    "scutoff:float:opt;svr:float:opt;smin:float:opt;smax:float:opt;kratio:float:opt;"
    "ow:int:opt;oh:int:opt;wintype:int:opt;interlaced:int:opt;sigma2:float:opt;"
    "sigma3:float:opt;sigma4:float:opt;oldfft:int:opt;device:int:opt";

}  // namespace

extern "C" __attribute__((visibility("default"))) void VapourSynthPluginInit2(
    VSPlugin* plugin, const VSPLUGINAPI* vspapi) {
  vspapi->configPlugin("com.github.humbleuser33.fft3dcl", "fft3dcl", "FFT3DCL denoiser (VS port)",
                       VS4_MAKE_VERSION(1, 1), VS4_API_VERSION, 0, plugin);
  vspapi->registerFunction("FFT3DCL", kArgs, "clip:vnode;", create, nullptr, plugin);
}
