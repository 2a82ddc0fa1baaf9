// A minimal single-threaded VapourSynth API 4 host for tests.
//
// Loads a plugin library, calls VapourSynthPluginInit2, invokes the
// registered function with a clip of caller-owned planes, and runs the
// two-phase getFrame protocol (arInitial requests, then arAllFramesReady)
// for a list of output frames, the way the VapourSynth core does with one
// thread. It checks the plugin's side of the protocol: frames fetched with
// getFrameFilter must have been requested, every frame and node reference
// must be released, and the filter must free its dependency node.
//
// Test infrastructure only (no product code uses it).
#include "vs4_abi.h"

#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

struct VSMap {
  std::map<std::string, std::vector<double>> f;
  std::map<std::string, std::vector<int64_t>> i;
  std::map<std::string, std::vector<VSNode*>> nodes;
  std::string error;
};

struct VSNode {
  int refs = 1;
  VSVideoInfo vi{};
  // source clip
  const void* const* planes = nullptr;  // frame-major numFrames * numPlanes
  int pad = 0;                          // extra bytes per row in frames handed out
  // filter
  VSFilterGetFrame getFrame = nullptr;
  VSFilterFree free = nullptr;
  void* instance = nullptr;
  int mode = -1;
  VSNode* dep = nullptr;
};

struct VSFrame {
  int refs = 1;
  VSVideoFormat fmt{};
  int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  ptrdiff_t stride[3] = {0, 0, 0};
  std::vector<uint8_t> data[3];
};

struct VSFrameContext {
  std::set<int> requested;
  std::vector<int> order;
  std::string error;
  std::string violation;
};

struct VSCore {
  std::vector<std::string> log;
};

namespace {

int g_liveFrames = 0;
int g_liveNodes = 0;
std::string g_id, g_ns, g_plugName, g_func, g_args;
VSPublicFunction g_create = nullptr;

int cfg(const char* id, const char* ns, const char* name, int, int, int, VSPlugin*) {
  g_id = id;
  g_ns = ns;
  g_plugName = name;
  return 1;
}
int reg(const char* name, const char* args, const char*, VSPublicFunction f, void*, VSPlugin*) {
  g_func = name;
  g_args = args;
  g_create = f;
  return 1;
}
int apiver() { return VS4_API_VERSION; }

VSFrame* makeFrame(const VSVideoFormat& fmt, int width, int height, int pad) {
  VSFrame* fr = new VSFrame();
  fr->fmt = fmt;
  for (int p = 0; p < fmt.numPlanes; p++) {
    fr->w[p] = width >> (p ? fmt.subSamplingW : 0);
    fr->h[p] = height >> (p ? fmt.subSamplingH : 0);
    fr->stride[p] = (ptrdiff_t)fr->w[p] * fmt.bytesPerSample + pad;
    fr->data[p].assign((size_t)fr->stride[p] * fr->h[p], 0xCD);
  }
  g_liveFrames++;
  return fr;
}

int clampN(int n, int num) { return n < 0 ? 0 : (n >= num ? num - 1 : n); }

void createVideoFilter(VSMap* out, const char*, const VSVideoInfo* vi, VSFilterGetFrame gf,
                       VSFilterFree fr, int mode, const VSFilterDependency* deps, int ndeps,
                       void* inst, VSCore*) {
  VSNode* n = new VSNode();
  g_liveNodes++;
  n->vi = *vi;
  n->getFrame = gf;
  n->free = fr;
  n->instance = inst;
  n->mode = mode;
  n->dep = ndeps > 0 ? deps[0].source : nullptr;
  out->nodes["clip"].push_back(n);
}
void freeNode(VSNode* n) {
  if (n && --n->refs == 0) {
    g_liveNodes--;
    delete n;
  }
}
const VSVideoInfo* getVideoInfo(VSNode* n) { return &n->vi; }
VSFrame* newVideoFrame(const VSVideoFormat* fmt, int w, int h, const VSFrame*, VSCore*) {
  return makeFrame(*fmt, w, h, 32);  // padded rows: the plugin must honour strides
}
void freeFrame(const VSFrame* f) {
  VSFrame* m = const_cast<VSFrame*>(f);
  if (m && --m->refs == 0) {
    g_liveFrames--;
    delete m;
  }
}
ptrdiff_t getStride(const VSFrame* f, int p) { return f->stride[p]; }
const uint8_t* getReadPtr(const VSFrame* f, int p) { return f->data[p].data(); }
uint8_t* getWritePtr(VSFrame* f, int p) { return f->data[p].data(); }
int getFrameWidth(const VSFrame* f, int p) { return f->w[p]; }
int getFrameHeight(const VSFrame* f, int p) { return f->h[p]; }

const VSFrame* getFrameFilter(int n, VSNode* node, VSFrameContext* ctx) {
  n = clampN(n, node->vi.numFrames);
  if (!ctx->requested.count(n)) {
    ctx->violation = "getFrameFilter(" + std::to_string(n) + ") without a request";
    return nullptr;
  }
  const VSVideoFormat& fmt = node->vi.format;
  VSFrame* fr = makeFrame(fmt, node->vi.width, node->vi.height, node->pad);
  for (int p = 0; p < fmt.numPlanes; p++) {
    const uint8_t* src = static_cast<const uint8_t*>(node->planes[(size_t)n * fmt.numPlanes + p]);
    const size_t row = (size_t)fr->w[p] * fmt.bytesPerSample;
    for (int y = 0; y < fr->h[p]; y++) std::memcpy(fr->data[p].data() + y * fr->stride[p], src + y * row, row);
  }
  return fr;
}
void requestFrameFilter(int n, VSNode* node, VSFrameContext* ctx) {
  n = clampN(n, node->vi.numFrames);
  ctx->requested.insert(n);
  ctx->order.push_back(n);
}
void setFilterError(const char* msg, VSFrameContext* ctx) { ctx->error = msg; }
void mapSetError(VSMap* m, const char* msg) { m->error = msg; }
int64_t mapGetInt(const VSMap* m, const char* key, int idx, int* err) {
  auto it = m->i.find(key);
  if (it == m->i.end() || idx >= (int)it->second.size()) {
    if (err) *err = 1;
    return 0;
  }
  if (err) *err = 0;
  return it->second[idx];
}
double mapGetFloat(const VSMap* m, const char* key, int idx, int* err) {
  auto it = m->f.find(key);
  if (it == m->f.end() || idx >= (int)it->second.size()) {
    if (err) *err = 1;
    return 0;
  }
  if (err) *err = 0;
  return it->second[idx];
}
VSNode* mapGetNode(const VSMap* m, const char* key, int idx, int* err) {
  auto it = m->nodes.find(key);
  if (it == m->nodes.end() || idx >= (int)it->second.size()) {
    if (err) *err = 1;
    return nullptr;
  }
  if (err) *err = 0;
  it->second[idx]->refs++;  // the getter hands out a new reference
  return it->second[idx];
}
void logMessage(int, const char* msg, VSCore* core) { core->log.push_back(msg); }

VSAPI makeApi() {
  VSAPI a;
  std::memset(&a, 0, sizeof(a));
  a.createVideoFilter = createVideoFilter;
  a.freeNode = freeNode;
  a.getVideoInfo = getVideoInfo;
  a.newVideoFrame = newVideoFrame;
  a.freeFrame = freeFrame;
  a.getStride = getStride;
  a.getReadPtr = getReadPtr;
  a.getWritePtr = getWritePtr;
  a.getFrameWidth = getFrameWidth;
  a.getFrameHeight = getFrameHeight;
  a.getFrameFilter = getFrameFilter;
  a.requestFrameFilter = requestFrameFilter;
  a.setFilterError = setFilterError;
  a.mapSetError = mapSetError;
  a.mapGetInt = mapGetInt;
  a.mapGetFloat = mapGetFloat;
  a.mapGetNode = mapGetNode;
  a.logMessage = logMessage;
  return a;
}

void copyOut(std::string& dst, const std::string& s) { dst = s; }
void put(char* buf, int len, const std::string& s) {
  if (buf && len > 0) std::snprintf(buf, len, "%s", s.c_str());
}

}  // namespace

extern "C" {

struct fv_clip {
  int colorFamily, sampleType, bits, ssw, ssh, width, height, numFrames, pad;
};

// Registration record of the plugin: "id|namespace|name|function|args".
int fv_plugin_info(const char* path, char* out, int len) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    put(out, len, dlerror());
    return -1;
  }
  auto init = reinterpret_cast<void (*)(VSPlugin*, const VSPLUGINAPI*)>(
      dlsym(h, "VapourSynthPluginInit2"));
  if (!init) {
    put(out, len, "no VapourSynthPluginInit2");
    return -1;
  }
  VSPLUGINAPI papi{apiver, cfg, reg};
  init(nullptr, &papi);
  put(out, len, g_id + "|" + g_ns + "|" + g_plugName + "|" + g_func + "|" + g_args);
  return 0;
}

// Runs the plugin's function on the clip. keys/types/values give the
// arguments (type 0 int, 1 float). requests receives, per output frame, up to
// 8 requested frame numbers (-1 padded). Returns 0 on success, 1 create
// error, 2 frame error, 3 protocol or reference-count violation; err holds
// the message and log the core log lines.
int fv_run(const char* path, const fv_clip* c, const void* const* planesIn, const char* const* keys,
           const int* types, const double* values, int nargs, void* const* planesOut,
           const int* order, int norder, int* requests, int* filterMode, char* err, int errlen,
           char* log, int loglen) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    put(err, errlen, dlerror());
    return 3;
  }
  auto init = reinterpret_cast<void (*)(VSPlugin*, const VSPLUGINAPI*)>(
      dlsym(h, "VapourSynthPluginInit2"));
  VSPLUGINAPI papi{apiver, cfg, reg};
  init(nullptr, &papi);
  VSAPI api = makeApi();
  VSCore core;
  g_liveFrames = 0;
  g_liveNodes = 0;

  VSNode* src = new VSNode();
  g_liveNodes++;
  VSVideoFormat& fmt = src->vi.format;
  fmt.colorFamily = c->colorFamily;
  fmt.sampleType = c->sampleType;
  fmt.bitsPerSample = c->bits;
  fmt.bytesPerSample = c->sampleType == stFloat ? 4 : (c->bits <= 8 ? 1 : 2);
  fmt.subSamplingW = c->colorFamily == cfYUV ? c->ssw : 0;
  fmt.subSamplingH = c->colorFamily == cfYUV ? c->ssh : 0;
  fmt.numPlanes = c->colorFamily == cfGray ? 1 : 3;
  src->vi.fpsNum = 24;
  src->vi.fpsDen = 1;
  src->vi.width = c->width;
  src->vi.height = c->height;
  src->vi.numFrames = c->numFrames;
  src->planes = planesIn;
  src->pad = c->pad;

  VSMap in, out;
  in.nodes["clip"].push_back(src);  // the map holds one reference
  for (int k = 0; k < nargs; k++) {
    if (types[k] == 0)
      in.i[keys[k]].push_back((int64_t)values[k]);
    else
      in.f[keys[k]].push_back(values[k]);
  }
  g_create(&in, &out, nullptr, &core, &api);
  in.nodes.clear();
  freeNode(src);  // drop the map's reference
  std::string logs;
  auto flushLog = [&] {
    for (auto& l : core.log) logs += l + "\n";
    put(log, loglen, logs);
  };
  if (!out.error.empty()) {
    put(err, errlen, out.error);
    flushLog();
    return g_liveNodes == 0 ? 1 : 3;
  }
  VSNode* filt = out.nodes["clip"][0];
  if (filterMode) *filterMode = filt->mode;
  int rc = 0;
  for (int k = 0; k < norder && rc == 0; k++) {
    const int n = order[k];
    VSFrameContext ctx;
    void* fd = nullptr;
    const VSFrame* r = filt->getFrame(n, arInitial, filt->instance, &fd, &ctx, &core, &api);
    if (r) {
      put(err, errlen, "frame returned at arInitial");
      rc = 3;
      break;
    }
    for (int q = 0; q < 8; q++) requests[k * 8 + q] = q < (int)ctx.order.size() ? ctx.order[q] : -1;
    r = filt->getFrame(n, arAllFramesReady, filt->instance, &fd, &ctx, &core, &api);
    if (!ctx.violation.empty()) {
      put(err, errlen, ctx.violation);
      rc = 3;
    } else if (!r) {
      put(err, errlen, ctx.error.empty() ? "no frame and no error" : ctx.error);
      rc = 2;
    } else {
      for (int p = 0; p < fmt.numPlanes; p++) {
        uint8_t* dst = static_cast<uint8_t*>(planesOut[(size_t)n * fmt.numPlanes + p]);
        const size_t row = (size_t)r->w[p] * fmt.bytesPerSample;
        for (int y = 0; y < r->h[p]; y++) std::memcpy(dst + y * row, r->data[p].data() + y * r->stride[p], row);
      }
      freeFrame(r);
    }
  }
  filt->free(filt->instance, &core, &api);
  freeNode(filt);
  flushLog();
  if (rc == 0 && (g_liveFrames != 0 || g_liveNodes != 0)) {
    put(err, errlen, "leak: " + std::to_string(g_liveFrames) + " frame(s), " +
                         std::to_string(g_liveNodes) + " node(s) still referenced");
    rc = 3;
  }
  return rc;
}

}  // extern "C"
