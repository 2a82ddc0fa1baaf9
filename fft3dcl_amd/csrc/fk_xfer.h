// Host <-> device transfers of planes held in pageable host memory.
//
// A pageable buffer cannot be read or written by the DMA engine directly, so
// the plane goes through a few pinned chunks owned by the plane object. Host
// threads copy chunk i between the caller's buffer and a pinned chunk while
// the DMA engine moves its neighbours:
//   upload   : copy chunk i into pinned slot i % NBUF, queue its DMA, go on
//              (a slot is refilled once the DMA that read it has finished);
//   download : queue the DMAs of the first NBUF chunks, then for each chunk
//              wait for its DMA, copy it out, and queue the DMA of the chunk
//              NBUF ahead into the slot just emptied.
// Pinned (page-locked) and device buffers go to the DMA engine as they are.
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

// Pinned chunks of one plane and the events of the DMAs that use them.
struct Staging {
  static constexpr int NBUF = 3;
  static constexpr size_t kDefaultChunk = 4 << 20;
  size_t chunk = kDefaultChunk;
  void* pin[NBUF] = {};
  hipEvent_t ev[NBUF] = {};
  bool inFlight[NBUF] = {};

  bool ready() const { return pin[0] != nullptr; }
  hipError_t init() {
    for (int k = 0; k < NBUF; k++) {
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
    for (int k = 0; k < NBUF; k++) {
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
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, s);
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
      const int k = (int)(i % NBUF);
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
      const int k = (int)(i % NBUF);
      hipError_t e = wait(k);  // an upload's DMA may still read the slot
      if (e != hipSuccess) return e;
      const size_t o = i * chunk;
      return queue(k, pin[k], d + o, std::min(chunk, n - o), hipMemcpyDeviceToHost, s);
    };
    for (size_t i = 0; i < nch && i < (size_t)NBUF; i++) {
      const hipError_t e = issue(i);
      if (e != hipSuccess) return e;
    }
    for (size_t i = 0; i < nch; i++) {
      const int k = (int)(i % NBUF);
      hipError_t e = wait(k);
      if (e != hipSuccess) return e;
      const size_t o = i * chunk;
      copy_pool().copy(h + o, pin[k], std::min(chunk, n - o));
      if (i + NBUF < nch && (e = issue(i + NBUF)) != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

}  // namespace fx
