// Host <-> device transfers of planes.
//
// A pageable buffer cannot be read or written by the device directly, so the
// plane goes through a few pinned chunks owned by the plane object. Host
// threads copy chunk i between the caller's buffer and a pinned chunk while
// the device moves its neighbours:
//   upload   : copy chunk i into pinned slot i % nbuf, queue its DMA, go on
//              (a slot is refilled once the DMA that read it has finished);
//   download : queue the device-side copies of the first nbuf chunks, then
//              for each chunk wait for its copy, copy it out, and queue the
//              chunk nbuf ahead into the slot just emptied.
// Page-locked and device buffers are handed over as they are. Device ->
// page-locked host copies run as a kernel (k_copy_out) on the caller's
// stream rather than on the DMA engine: on this platform a host-to-device DMA
// and a device-to-host DMA do not overlap, but a copy kernel writing host
// memory overlaps a host-to-device DMA (fft3d_plane_prefetch relies on it).
//
// The reference moves each plane with one blocking clEnqueueWriteBuffer /
// clEnqueueReadBuffer (src/cl/gpu.cpp:360-365, src/fft3dcl.cpp:396-448).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace fx {

// dst[i] = src[i] for i < n16 (16-byte units): device memory to page-locked
// host memory, written by vector stores over the bus.
__global__ __launch_bounds__(256) void k_copy_out(const uint4* __restrict__ src,
                                                  uint4* __restrict__ dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Device -> page-locked host on stream s (asynchronous). The kernel writes
// through the device address the runtime reports for h; host memory without
// one, and sizes or addresses that are not multiples of 16, go to
// hipMemcpyAsync.
inline hipError_t copy_to_host(void* h, const void* d, size_t n, hipStream_t s) {
  hipPointerAttribute_t a{};
  void* dh = nullptr;
  if (hipPointerGetAttributes(&a, h) == hipSuccess && a.type == hipMemoryTypeHost)
    dh = a.devicePointer;
  else
    (void)hipGetLastError();
  if (!dh || n % 16 || reinterpret_cast<uintptr_t>(dh) % 16 ||
      reinterpret_cast<uintptr_t>(d) % 16)
    return hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s);
  const size_t n16 = n / 16;
  const unsigned grid = (unsigned)std::min<size_t>(1024, (n16 + 255) / 256);
  if (grid == 0) return hipSuccess;
  FK_LAUNCH(k_copy_out, dim3(grid), dim3(256), s, static_cast<const uint4*>(d),
            static_cast<uint4*>(dh), n16);
  return hipGetLastError();
}

// Process-wide pool of host threads for large memcpy: the caller and the
// helpers each copy one contiguous piece. One copy at a time; a caller that
// finds the pool busy copies alone.
class CopyPool {
 public:
  CopyPool() {
    const unsigned hc = std::thread::hardware_concurrency();
    const int n = (int)std::min(8u, std::max(1u, hc / 2));
    for (int i = 1; i < n; i++) th_.emplace_back([this, i] { worker(i); });
    nthr_ = n;
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void copy(void* dst, const void* src, size_t n) {
    std::unique_lock<std::mutex> busy(job_, std::try_to_lock);
    if (nthr_ == 1 || n < kMinSplit || !busy.owns_lock()) {
      std::memcpy(dst, src, n);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      d_ = static_cast<char*>(dst);
      s_ = static_cast<const char*>(src);
      n_ = n;
      pending_ = nthr_ - 1;
      gen_++;
    }
    cv_.notify_all();
    piece(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  static constexpr size_t kMinSplit = 256 << 10;
  void piece(int i) {
    // 64-byte aligned piece boundaries
    const size_t per = ((n_ + nthr_ - 1) / nthr_ + 63) & ~(size_t)63;
    const size_t a = std::min(n_, per * i), b = std::min(n_, a + per);
    if (b > a) std::memcpy(d_ + a, s_ + a, b - a);
  }
  void worker(int i) {
    unsigned long long seen = 0;
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      lk.unlock();
      piece(i);
      lk.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  int nthr_ = 1;
  std::mutex job_;  // held by the one caller whose copy the helpers work on
  std::mutex m_;
  std::condition_variable cv_, done_;
  char* d_ = nullptr;
  const char* s_ = nullptr;
  size_t n_ = 0;
  int pending_ = 0;
  unsigned long long gen_ = 0;
  bool stop_ = false;
};

inline CopyPool& copy_pool() {
  static CopyPool pool;
  return pool;
}

// True for host memory the DMA engine cannot address directly (not
// page-locked, not device memory).
inline bool pageable(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return a.type == hipMemoryTypeUnregistered;
}

// Pinned chunks of one plane and the events of the DMAs that use them. With
// nbuf chunks covering a whole plane, an upload returns as soon as the host
// copies are done; with fewer, it also waits for the DMAs of earlier chunks.
struct Staging {
  static constexpr int MAXBUF = 16;
  static constexpr size_t kDefaultChunk = 4 << 20;
  size_t chunk = kDefaultChunk;
  int nbuf = 3;
  void* pin[MAXBUF] = {};
  hipEvent_t ev[MAXBUF] = {};
  bool inFlight[MAXBUF] = {};

  bool ready() const { return pin[0] != nullptr; }
  // n chunks (1..MAXBUF), allocated now
  hipError_t init(int n = 3) {
    nbuf = std::max(1, std::min(n, MAXBUF));
    for (int k = 0; k < nbuf; k++) {
      hipError_t e = hipHostMalloc(&pin[k], chunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
      if (e != hipSuccess) {
        release();
        return e;
      }
    }
    return hipSuccess;
  }
  // the caller makes sure no DMA on these chunks is still queued
  void release() {
    for (int k = 0; k < MAXBUF; k++) {
      if (pin[k]) (void)hipHostFree(pin[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
      pin[k] = nullptr;
      ev[k] = nullptr;
      inFlight[k] = false;
    }
  }
  hipError_t wait(int k) {
    if (!inFlight[k]) return hipSuccess;
    inFlight[k] = false;
    return hipEventSynchronize(ev[k]);
  }
  hipError_t queue(int k, void* dst, const void* src, size_t n, hipMemcpyKind kind,
                   hipStream_t s) {
    hipError_t e = kind == hipMemcpyDeviceToHost ? copy_to_host(dst, src, n, s)
                                                 : hipMemcpyAsync(dst, src, n, kind, s);
    if (e == hipSuccess) e = hipEventRecord(ev[k], s);
    if (e == hipSuccess) inFlight[k] = true;
    return e;
  }

  // Host -> device. Returns once the host buffer has been read; the DMAs run
  // on s ahead of whatever the caller queues on s next.
  hipError_t upload(void* dDst, const void* hSrc, size_t n, hipStream_t s) {
    char* d = static_cast<char*>(dDst);
    const char* h = static_cast<const char*>(hSrc);
    for (size_t o = 0, i = 0; o < n; o += chunk, i++) {
      const int k = (int)(i % nbuf);
      const size_t len = std::min(chunk, n - o);
      hipError_t e = wait(k);
      if (e != hipSuccess) return e;
      copy_pool().copy(pin[k], h + o, len);
      e = queue(k, d + o, pin[k], len, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }

  // Device -> host, after the work queued on s so far. Returns once the
  // whole plane is in hDst.
  hipError_t download(void* hDst, const void* dSrc, size_t n, hipStream_t s) {
    char* h = static_cast<char*>(hDst);
    const char* d = static_cast<const char*>(dSrc);
    const size_t nch = (n + chunk - 1) / chunk;
    auto issue = [&](size_t i) {
      const int k = (int)(i % nbuf);
      hipError_t e = wait(k);  // an upload's DMA may still read the slot
      if (e != hipSuccess) return e;
      const size_t o = i * chunk;
      return queue(k, pin[k], d + o, std::min(chunk, n - o), hipMemcpyDeviceToHost, s);
    };
    for (size_t i = 0; i < nch && i < (size_t)nbuf; i++) {
      const hipError_t e = issue(i);
      if (e != hipSuccess) return e;
    }
    for (size_t i = 0; i < nch; i++) {
      const int k = (int)(i % nbuf);
      hipError_t e = wait(k);
      if (e != hipSuccess) return e;
      const size_t o = i * chunk;
      copy_pool().copy(h + o, pin[k], std::min(chunk, n - o));
      if (i + nbuf < nch && (e = issue(i + nbuf)) != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

}  // namespace fx
