// Kernel entry points of one block shape, as launched by the C ABI
// (fft3d_capi.hip). The templates are instantiated per block shape in
// fk_set_bw.hip, one translation unit per shape, so that they compile in
// parallel.
#pragma once

#include "fk_batch.h"

namespace fk {

using AnalysisFn = void (*)(PlaneArgs, const void*, c32*);
using FilterFn = void (*)(PlaneArgs, WindowArgs, float4*, int, float*);
using SynthFn = void (*)(PlaneArgs, const float*, void*);
using SweepFn = void (*)(PlaneArgs, WindowArgs, float4*, int, SweepArgs);
using CombineFn = void (*)(PlaneArgs, SweepArgs);
using BatchFn = void (*)(const BatchItem*, int);

struct KernelSet {
  AnalysisFn analysis = nullptr;
  FilterFn filter_inverse = nullptr;
  SweepFn sweep = nullptr;
  SweepFn sweepFwd = nullptr;  // with the new frame's analysis inside
  // the same over a ring of FP16 spectra (square blocks of 16, 32 and 64)
  AnalysisFn analysisH = nullptr;
  FilterFn filterH = nullptr;
  SweepFn sweepH = nullptr;
  // the MFMA DFT form (32x32 blocks): [precision - 1][FP16 ring]
  AnalysisFn analysisM[2][2] = {};
  SweepFn sweepM[2][2] = {};
  // batched launches (fft3d_step): [DFT form 0..2][FP16 ring]
  BatchFn analysisB[3][2] = {};
  BatchFn sweepB[3][2] = {};
  // the register-FFT sweep sized for more resident workgroups per CU:
  // [workgroups per CU - 2][FP16 ring], where the LDS holds them (square 8
  // and 16 blocks); [0] is the default sweep
  SweepFn sweepOcc[3][2] = {};
  BatchFn sweepBOcc[3][2] = {};
  int maxOcc = 2;  // largest sweep workgroups per CU with a kernel
  int G = 0;
  int fftThreads = 64;   // threads per analysis / filter-inverse workgroup
  int sweepThreads = 0;  // 64 * waves per sweep workgroup
  int nbw = 0;           // block columns per sweep strip
  int sweepPerCU = 0;    // sweep workgroups one CU holds (LDS, 2 waves per SIMD)
};

// The kernels of block shape (BW, BH), for BW, BH in {4, 8, 16, 32, 64};
// each shape is defined by its own compile of fk_set_bw.hip.
template <int BW, int BH>
KernelSet kernel_set();

// Shapes with a side above 64 (fk_set_large.hip); empty for others.
KernelSet kernel_set_large(int bw, int bh);

}  // namespace fk
