// Thread-per-lane CPU emulation of the HIP subset used by fft3dcl_amd/csrc.
//
// Test infrastructure: lets the kernels and the C-ABI host run on a machine
// with no GPU, so index arithmetic, LDS layouts and barriers are checked
// (under AddressSanitizer in tests/test_emu_asan.py) before a kernel ever
// runs on the device. One OS thread per lane; workgroups run one after the
// other so a kernel's __shared__ arrays (function-local statics here) are
// private to the running workgroup. "Device" memory is host memory.
#pragma once

#include <algorithm>
#include <barrier>
#include <memory>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iterator>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float4 {
  float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
struct float2 {
  float x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
struct uint4 {
  unsigned x, y, z, w;
};
struct uint2 {
  unsigned x, y;
};
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
struct int2 {
  int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }
using std::max;
using std::min;

inline thread_local dim3 threadIdx, blockIdx;
inline dim3 blockDim, gridDim;
inline std::barrier<>* emu_barrier_ = nullptr;
inline thread_local std::barrier<>* emu_wave_barrier_ = nullptr;

inline void __syncthreads() { emu_barrier_->arrive_and_wait(); }
// Barrier over the 64 lanes of the calling lane's wavefront only, so that a
// kernel relying on it for data shared across waves races here as on a GPU.
inline void emu_wave_sync() { emu_wave_barrier_->arrive_and_wait(); }
inline float emu_sqrtf(float x) { return std::sqrt(x); }
// IEEE binary16 conversions, round to nearest even (the GPU's conversion).
inline uint16_t emu_f2h(float f) {
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
inline float emu_h2f(uint16_t h) {
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
using std::copysign;
using std::floor;
inline float copysignf(float a, float b) { return std::copysign(a, b); }
inline float floorf(float a) { return std::floor(a); }
inline float fmaxf(float a, float b) { return std::fmax(a, b); }

// Runs fn() once per lane of every workgroup of the grid.
inline void emu_launch(dim3 grid, dim3 block, const std::function<void()>& fn) {
  const unsigned nt = block.x * block.y * block.z;
  blockDim = block;
  gridDim = grid;
  std::barrier<> bar(nt);
  emu_barrier_ = &bar;
  std::vector<std::unique_ptr<std::barrier<>>> waves;
  for (unsigned w = 0; w * 64 < nt; w++)
    waves.emplace_back(new std::barrier<>(std::min<unsigned>(64, nt - w * 64)));
  std::vector<std::thread> th;
  th.reserve(nt);
  for (unsigned t = 0; t < nt; t++) {
    th.emplace_back([&, t] {
      threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
      emu_wave_barrier_ = waves[t / 64].get();
      for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
          for (unsigned bx = 0; bx < grid.x; bx++) {
            blockIdx = dim3(bx, by, bz);
            fn();
            bar.arrive_and_wait();  // the next workgroup reuses the shared arrays
          }
    });
  }
  for (auto& x : th) x.join();
  emu_barrier_ = nullptr;
}

// ---- runtime API subset (synchronous, host memory) ----
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
#define hipStreamNonBlocking 1
struct hipDeviceProp_t {
  char name[256];
  size_t sharedMemPerBlock;
  size_t totalGlobalMem;
  int multiProcessorCount;
  char gcnArchName[256];
};

inline hipError_t emu_fail_alloc_ = 0;  // tests set it to force allocation failures
inline hipError_t hipMalloc(void** p, size_t n) {
  if (emu_fail_alloc_) return hipErrorOutOfMemory;
  *p = std::calloc(1, n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = reinterpret_cast<hipStream_t>(0x1);
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
// A file-scope __device__ array is an ordinary host array here; its "device
// address" is its host address.
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) __attribute__((aligned(n)))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->name, "cpu-emulation");
  std::strcpy(p->gcnArchName, "emu");
  p->sharedMemPerBlock = 160 * 1024;
  p->totalGlobalMem = (size_t)1 << 34;
  // the emulated part: MI355X (FFT3D_EMU_CUS sets fewer, so that small
  // planes take the paths of planes larger than one round of workgroups)
  const char* cus = std::getenv("FFT3D_EMU_CUS");
  p->multiProcessorCount = cus && std::atoi(cus) > 0 ? std::atoi(cus) : 256;
  return hipSuccess;
}
// two resident workgroups per CU, whatever the kernel
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
  *n = 2;
  return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t e) {
  return e == hipSuccess ? "no error" : (e == hipErrorOutOfMemory ? "out of memory" : "error");
}
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = nullptr;
  return hipSuccess;
}
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// host memory: every buffer of the emulation is plain host memory. The
// pointer query reports the blocks of hipHostMalloc as page-locked (device
// address = host address) and everything else as pageable, so the staged
// transfer path and its copy kernel both run.
#define hipHostMallocDefault 0
inline std::mutex emu_pinned_mtx_;
inline std::map<uintptr_t, size_t> emu_pinned_;  // base -> size
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
  if (emu_fail_alloc_) return hipErrorOutOfMemory;
  *p = std::calloc(1, n ? n : 1);
  if (!*p) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> lk(emu_pinned_mtx_);
  emu_pinned_[reinterpret_cast<uintptr_t>(*p)] = n ? n : 1;
  return hipSuccess;
}
inline hipError_t hipHostFree(void* p) {
  {
    std::lock_guard<std::mutex> lk(emu_pinned_mtx_);
    emu_pinned_.erase(reinterpret_cast<uintptr_t>(p));
  }
  std::free(p);
  return hipSuccess;
}
enum hipMemoryType { hipMemoryTypeUnregistered = 0, hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 };
struct hipPointerAttribute_t {
  hipMemoryType type;
  int device;
  void* devicePointer;
};
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t* a, const void* p) {
  a->type = hipMemoryTypeUnregistered;
  a->device = 0;
  a->devicePointer = nullptr;
  const uintptr_t u = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lk(emu_pinned_mtx_);
  auto it = emu_pinned_.upper_bound(u);
  if (it != emu_pinned_.begin() && u < std::prev(it)->first + std::prev(it)->second) {
    a->type = hipMemoryTypeHost;
    a->devicePointer = const_cast<void*>(p);
  }
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.0f;
  return hipSuccess;
}

// hipGraph API subset: present so that the library compiles; the emulation
// runs every launch at once, so the library never captures under it.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return 1; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return 1; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) {
  return 1;
}
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 1; }
struct hipFuncAttributes {
  int maxThreadsPerBlock;
};
inline hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*) {
  a->maxThreadsPerBlock = 1024;
  return hipSuccess;
}

#define FK_LAUNCH(fn, grid, block, stream, ...) \
  emu_launch((grid), (block), [&]() { fn(__VA_ARGS__); })
