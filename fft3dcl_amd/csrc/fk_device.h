// Device-side basics shared by the FFT3D kernels.
//
// The kernels are plain HIP for gfx950. The same source also compiles as host
// C++ against a thread-per-lane emulation of the few HIP facilities they use
// (FFT3D_CPU_EMU, tests/emu/hip_emu.h), which is how the index arithmetic,
// LDS layouts and barriers are checked on a machine without a GPU.
#pragma once

#if defined(FFT3D_CPU_EMU)
#include "hip_emu.h"
#else
#include <hip/hip_runtime.h>
// Launch through a kernel function pointer (template instantiations are
// selected at run time from a dispatch table).
#define FK_LAUNCH(fn, grid, block, stream, ...) \
  fn<<<(grid), (block), 0, (stream)>>>(__VA_ARGS__)
#endif

#include <cstdint>

namespace fk {

struct c32 {
  float x, y;
};

__device__ __forceinline__ c32 cmake(float x, float y) { return c32{x, y}; }
__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return c32{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return c32{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c32 cconj(c32 a) { return c32{a.x, -a.y}; }
__device__ __forceinline__ c32 cscale(c32 a, float s) { return c32{a.x * s, a.y * s}; }

// Two floats as FP16 (round to nearest even) packed low-high in 32 bits, and
// back: the storage of a half-precision spectra ring.
__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
#if defined(FFT3D_CPU_EMU)
  return (uint32_t)emu_f2h(a) | ((uint32_t)emu_f2h(b) << 16);
#else
  const _Float16 ha = (_Float16)a, hb = (_Float16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, ha) |
         ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
#endif
}
__device__ __forceinline__ float2 unpack_h2(uint32_t v) {
#if defined(FFT3D_CPU_EMU)
  return float2{emu_h2f((uint16_t)(v & 0xffffu)), emu_h2f((uint16_t)(v >> 16))};
#else
  const _Float16 a = __builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
  const _Float16 b = __builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
  return make_float2((float)a, (float)b);
#endif
}

// Reciprocal and square root used by the filter. Tolerances of the path are
// far above one ulp, so the hardware approximations (1 ulp) are used.
__device__ __forceinline__ float frcp(float x) {
#if defined(FFT3D_CPU_EMU)
  return 1.0f / x;
#else
  return __builtin_amdgcn_rcpf(x);
#endif
}
__device__ __forceinline__ float fsqrt(float x) {
#if defined(FFT3D_CPU_EMU)
  return emu_sqrtf(x);
#else
  return __builtin_amdgcn_sqrtf(x);
#endif
}

// Orders LDS accesses among the lanes of one wavefront. The LDS instructions
// of a wave execute in issue order, so a compiler fence at wavefront scope
// around the wave barrier suffices for data that stays inside an LDS area
// owned by the calling wave; areas shared between waves need __syncthreads.
__device__ __forceinline__ void wave_sync() {
#if defined(FFT3D_CPU_EMU)
  emu_wave_sync();
#else
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}

// wave_sync that also orders global-memory accesses of the wave's lanes: a
// workgroup-scope release (the wave's stores complete) before the barrier and
// the matching acquire after it.
__device__ __forceinline__ void wave_sync_global() {
#if defined(FFT3D_CPU_EMU)
  emu_wave_sync();
#else
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#endif
}

} // namespace fk
