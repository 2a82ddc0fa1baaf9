// C-ABI shim over the reference's own CPU backend, built from the reference's
// sources where they lie (src/params.h, src/cpu/{reference,tables,fft}.cpp)
// into oracle/_ref/libfft3dref.so by oracle/Makefile. TEST INFRASTRUCTURE
// ONLY: tests use it to check the oracle restatement bit for bit, and
// bench.py may time it as the "reference" CPU baseline. Nothing here is
// product code and no reference source is copied: this file only calls the
// reference's classes (PlaneRef, FftPlan, validateParams).
#include "params.h"
#include "cpu/fft.h"
#include "cpu/reference.h"

#include "fft3d_oracle.h"

#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

using namespace fft3dcl;

namespace {
ScriptParams toScript(const orc_script* s) {
  ScriptParams sp;
  sp.sigma = s->sigma; sp.beta = s->beta; sp.bw = s->bw; sp.bh = s->bh; sp.bt = s->bt;
  sp.sharpen = s->sharpen; sp.plane = s->plane; sp.mode = s->mode;
  sp.bordersize = s->bordersize; sp.precision = s->precision; sp.degrid = s->degrid;
  sp.scutoff = s->scutoff; sp.svr = s->svr; sp.smin = s->smin; sp.smax = s->smax;
  sp.kratio = s->kratio; sp.ow = s->ow; sp.oh = s->oh; sp.wintype = s->wintype;
  sp.interlaced = s->interlaced; sp.sigma2 = s->sigma2; sp.sigma3 = s->sigma3;
  sp.sigma4 = s->sigma4; sp.oldfft = s->oldfft;
  return sp;
}

void fromInternal(const InternalParams& i, orc_internal* o) {
  std::memset(o, 0, sizeof(*o));
  o->sigma = i.sigma; o->beta = i.beta; o->bw = i.bw; o->bh = i.bh; o->bm = i.bm; o->bt = i.bt;
  o->sharpen = i.sharpen; o->sharpenOn = i.sharpenOn; o->mode = i.mode;
  o->bordersize = i.bordersize; o->precision = i.precision; o->degrid = i.degrid;
  o->degridOn = i.degridOn; o->scutoff = i.scutoff; o->svr = i.svr; o->smin = i.smin;
  o->smax = i.smax; o->kratio = i.kratio; o->ow = i.ow; o->oh = i.oh; o->wintype = i.wintype;
  o->interlaced = i.interlaced; o->sigma2 = i.sigma2; o->sigma3 = i.sigma3;
  o->sigma4 = i.sigma4; o->usePattern = i.usePattern; o->current = i.current; o->nx = i.nx;
  o->ny = i.ny; o->totw = i.totw; o->toth = i.toth; o->oldfft = i.oldfft;
  o->wienerS = i.wienerS; o->betaX = i.betaX; o->kalmanCN = i.kalmanCN; o->ssmin = i.ssmin;
  o->ssmax = i.ssmax;
}

struct RefPlane {
  InternalParams ip;
  PlaneGeom g;
  std::unique_ptr<PlaneRef> ref;
};

int clampFrame(int n, int numFrames) {
  if (n < 0) return 0;
  if (n > numFrames - 1) return numFrames - 1;
  return n;
}
} // namespace

extern "C" {

int ref_round_pow2_clamp(int v) { return roundPow2Clamp(v); }

int ref_validate(const orc_script* s, int w, int h, int bits, orc_internal* out, const char** err) {
  InternalParams ip;
  bool ok = validateParams(toScript(s), w, h, bits, ip, err);
  if (ok) fromInternal(ip, out);
  return ok ? 1 : 0;
}

void* ref_plane_new(const orc_script* s, int w, int h, int bits, int sampleType, int chroma) {
  auto* rp = new RefPlane();
  const char* err = nullptr;
  if (!validateParams(toScript(s), w, h, bits, rp->ip, &err)) {
    delete rp;
    return nullptr;
  }
  rp->g.width = w; rp->g.height = h; rp->g.bits = bits; rp->g.sampleType = sampleType;
  rp->g.chroma = chroma != 0;
  rp->ref = std::make_unique<PlaneRef>(rp->ip, rp->g);
  return rp;
}

void ref_plane_free(void* p) { delete static_cast<RefPlane*>(p); }

int ref_plane_nblocks(void* p) { return static_cast<RefPlane*>(p)->ref->nblocks(); }
int ref_plane_fbw(void* p) { return static_cast<RefPlane*>(p)->ref->fbw(); }
float ref_plane_chroma_norm(void* p) { return static_cast<RefPlane*>(p)->ref->chromaNorm(); }

int ref_plane_table(void* p, int which, void* dst) {
  PlaneRef& r = *static_cast<RefPlane*>(p)->ref;
  switch (which) {
  case 0: std::memcpy(dst, r.winAna().data(), r.winAna().size() * 4); return (int)r.winAna().size();
  case 1: std::memcpy(dst, r.winSynA().data(), r.winSynA().size() * 4); return (int)r.winSynA().size();
  case 2: std::memcpy(dst, r.winSynB().data(), r.winSynB().size() * 4); return (int)r.winSynB().size();
  case 3: std::memcpy(dst, r.pattern().data(), r.pattern().size() * 4); return (int)r.pattern().size();
  case 4: std::memcpy(dst, r.sharpenLut().data(), r.sharpenLut().size() * 4); return (int)r.sharpenLut().size();
  case 5: std::memcpy(dst, r.bx().data(), r.bx().size() * 4); return (int)r.bx().size();
  case 6: std::memcpy(dst, r.by().data(), r.by().size() * 4); return (int)r.by().size();
  case 7: std::memcpy(dst, r.par().data(), r.par().size()); return (int)r.par().size();
  case 8: { // degrid sample, first planeSize entries interleaved
    int ps = r.fbw() * static_cast<RefPlane*>(p)->ip.bh;
    if ((int)r.gsRe().size() < ps) return 0;
    float* d = static_cast<float*>(dst);
    for (int i = 0; i < ps; i++) { d[2 * i] = r.gsRe()[i]; d[2 * i + 1] = r.gsIm()[i]; }
    return ps;
  }
  }
  return -1;
}

void ref_build_degrid(void* p) { static_cast<RefPlane*>(p)->ref->buildDegridSample(); }

void ref_upload(void* p, const void* src, ptrdiff_t pitch, float* dst) {
  std::vector<float> v;
  static_cast<RefPlane*>(p)->ref->upload(static_cast<const uint8_t*>(src), pitch, v);
  std::memcpy(dst, v.data(), v.size() * 4);
}

void ref_download(void* p, const float* src, void* dst, ptrdiff_t pitch) {
  RefPlane* rp = static_cast<RefPlane*>(p);
  std::vector<float> v(src, src + (size_t)rp->g.width * rp->g.height);
  rp->ref->download(v, static_cast<uint8_t*>(dst), pitch);
}

void ref_compute_spectra(void* p, const float* plane, float* re, float* im) {
  RefPlane* rp = static_cast<RefPlane*>(p);
  std::vector<float> v(plane, plane + (size_t)rp->g.width * rp->g.height);
  auto s = rp->ref->computeSpectra(v);
  std::memcpy(re, s->re.data(), s->re.size() * 4);
  std::memcpy(im, s->im.data(), s->im.size() * 4);
}

void ref_filter_synth(void* p, const float* const* wre, const float* const* wim, int nwin,
                      float* out) {
  RefPlane* rp = static_cast<RefPlane*>(p);
  size_t total = (size_t)rp->ref->nblocks() * rp->ref->fbw() * rp->ip.bh;
  std::vector<std::shared_ptr<PlaneRef::Spectra>> win;
  for (int i = 0; i < nwin; i++) {
    auto s = std::make_shared<PlaneRef::Spectra>();
    s->re.assign(wre[i], wre[i] + total);
    s->im.assign(wim[i], wim[i] + total);
    win.push_back(s);
  }
  std::vector<float> o;
  rp->ref->filterAndSynthesize(win, o);
  std::memcpy(out, o.data(), o.size() * 4);
}

void ref_fft1d(int n, const float* in, float* out, int inverse) {
  FftPlan plan(n);
  const auto* ci = reinterpret_cast<const std::complex<float>*>(in);
  auto* co = reinterpret_cast<std::complex<float>*>(out);
  if (inverse) plan.inverse(ci, co); else plan.forward(ci, co);
}

void ref_fft2d_real_forward(const float* block, int bw, int bh, float* full) {
  FftPlan rp(bw), cp(bh);
  fft2dRealForward(block, bw, bh, rp, cp, reinterpret_cast<std::complex<float>*>(full));
}

void ref_fft2d_real_inverse_half(const float* half, int bw, int bh, float* block) {
  FftPlan rp(bw), cp(bh);
  fft2dRealInverseHalf(reinterpret_cast<const std::complex<float>*>(half), bw, bh, bw / 2 + 1,
                       rp, cp, block);
}

// Clip driver over the reference's PlaneRef, with the window, degrid and cache
// logic of its CPU getFrame path (src/fft3dcl.cpp:200-258, 319-331, 375-386,
// 471-521). Same signature as orc_run_clip (float_out is float here).
int ref_run_clip(const orc_script* s, int colorFamily, int width, int height, int ssw, int ssh,
                 int bits, int sampleType, int nframes, const void* const* in,
                 void* const* out, float* const* floatOut, const int* order, int norder,
                 const char** err) {
  const int nplanes = colorFamily == 1 ? 1 : 3;
  const int bps = sampleType == 1 ? 4 : (bits <= 8 ? 1 : 2);
  ScriptParams sp = toScript(s);
  int plane = sp.plane;
  if (colorFamily == 2) plane = 4;
  if (colorFamily == 1) {
    if (plane == 1) { *err = "FFT3DCL: cannot set chroma processing for a greyscale clip"; return -1; }
    if (plane == 4) plane = 0;
  }
  if (plane == 2 || plane == 3) plane = 1;
  if (plane != 0 && plane != 1 && plane != 4) {
    *err = "FFT3DCL: plane must be 0 (luma), 1 (chroma) or 4 (all)";
    return -1;
  }
  std::vector<int> planes;
  if (colorFamily == 1 || plane == 0) planes = {0};
  else if (plane == 1) planes = {1, 2};
  else planes = {0, 1, 2};
  std::vector<std::unique_ptr<PlaneRef>> refs;
  std::vector<InternalParams> ips;
  std::vector<PlaneGeom> geoms;
  InternalParams ip;
  for (int q : planes) {
    int pw = width >> (q ? ssw : 0), ph = height >> (q ? ssh : 0);
    if (!validateParams(sp, pw, ph, bits, ip, err)) return -1;
    ips.push_back(ip);
    PlaneGeom g{pw, ph, bits, sampleType, colorFamily == 3 && q > 0};
    geoms.push_back(g);
    refs.push_back(std::make_unique<PlaneRef>(ip, g));
  }
  std::vector<std::map<int, std::shared_ptr<PlaneRef::Spectra>>> cache(planes.size());
  std::vector<bool> degridDone(planes.size(), false);
  std::vector<float> planeBuf, outBuf;
  for (int oi = 0; oi < norder; oi++) {
    const int n = order[oi];
    for (int q = 0; q < nplanes; q++) {
      bool processed = false;
      for (int p : planes) processed |= p == q;
      if (processed) continue;
      int w = width >> (q ? ssw : 0), h = height >> (q ? ssh : 0);
      std::memcpy(out[(size_t)n * nplanes + q], in[(size_t)n * nplanes + q], (size_t)w * h * bps);
    }
    for (size_t k = 0; k < planes.size(); k++) {
      const int q = planes[k];
      PlaneRef& ref = *refs[k];
      const PlaneGeom& g = geoms[k];
      if (ip.degridOn && !degridDone[k]) {
        ref.buildDegridSample();
        degridDone[k] = true;
      }
      auto getSpectra = [&](int fn) {
        fn = clampFrame(fn, nframes);
        auto it = cache[k].find(fn);
        if (it != cache[k].end()) return it->second;
        planeBuf.clear();
        ref.upload(static_cast<const uint8_t*>(in[(size_t)fn * nplanes + q]),
                   (ptrdiff_t)g.width * bps, planeBuf);
        auto sp2 = ref.computeSpectra(planeBuf);
        cache[k][fn] = sp2;
        return sp2;
      };
      std::vector<std::shared_ptr<PlaneRef::Spectra>> window;
      if (ip.bm >= 2) {
        for (int i = 0; i < ip.bt; i++) window.push_back(getSpectra(n - ip.current + i));
      } else {
        window.push_back(getSpectra(n));
      }
      outBuf.clear();
      ref.filterAndSynthesize(window, outBuf);
      if (floatOut && floatOut[(size_t)n * nplanes + q])
        std::memcpy(floatOut[(size_t)n * nplanes + q], outBuf.data(), outBuf.size() * 4);
      ref.download(outBuf, static_cast<uint8_t*>(out[(size_t)n * nplanes + q]),
                   (ptrdiff_t)g.width * bps);
      if (ip.bm >= 2) {
        int lo = n - ip.current - 1, hi = n - ip.current + ip.bt;
        for (auto it = cache[k].begin(); it != cache[k].end();)
          it = (it->first < lo || it->first > hi) ? cache[k].erase(it) : std::next(it);
      } else if (ip.bm != 0 && !cache[k].empty()) {
        cache[k].clear();
      }
    }
  }
  return 0;
}

} // extern "C"
