// The kernel set of one block shape FK_BW x FK_BH (compiled once per shape:
// -DFK_BW=... -DFK_BH=..., each in 4, 8, 16, 32, 64).
#include "fk_set.h"

#if !defined(FK_BW) || !defined(FK_BH)
#error "FK_BW and FK_BH (block width and height) must be defined"
#endif

namespace fk {
namespace {

// Sweep workgroups hold WPW waves; 4 unless the per-wave LDS of large blocks
// would leave room for only one workgroup per CU.
template <int BW, int BH>
KernelSet make_set() {
  constexpr int WPW = (BW * BH >= 4096) ? 2 : 4;
  constexpr int LDS_BYTES = 4 * SweepShape<BW, BH, WPW>::LDS_FLOATS;
  constexpr int BY_LDS = 160 * 1024 / LDS_BYTES, BY_WAVES = 8 / WPW;
  KernelSet k;
  k.analysis = &k_analysis<BW, BH, false>;
  k.filter_inverse = &k_filter_inverse<BW, BH, false>;
  k.sweep = &k_filter_synth<BW, BH, WPW, false, false>;
  k.sweepFwd = &k_filter_synth<BW, BH, WPW, true, false>;
  k.analysisB[0][0] = &k_analysis_batch<BW, BH, false>;
  k.sweepB[0][0] = &k_sweep_batch<BW, BH, WPW, false, 0>;
  if constexpr (BW == BH && BW >= 16) {
    k.analysisH = &k_analysis<BW, BH, true>;
    k.filterH = &k_filter_inverse<BW, BH, true>;
    k.sweepH = &k_filter_synth<BW, BH, WPW, false, true>;
    k.analysisB[0][1] = &k_analysis_batch<BW, BH, true>;
    k.sweepB[0][1] = &k_sweep_batch<BW, BH, WPW, true, 0>;
  }
  if constexpr (BW == 32 && BH == 32) {
    k.analysisM[0][0] = &k_analysis_mfma<mm::F16, false>;
    k.analysisM[0][1] = &k_analysis_mfma<mm::F16, true>;
    k.analysisM[1][0] = &k_analysis_mfma<mm::BF16, false>;
    k.analysisM[1][1] = &k_analysis_mfma<mm::BF16, true>;
    k.sweepM[0][0] = &k_filter_synth<BW, BH, WPW, false, false, mm::F16>;
    k.sweepM[0][1] = &k_filter_synth<BW, BH, WPW, false, true, mm::F16>;
    k.sweepM[1][0] = &k_filter_synth<BW, BH, WPW, false, false, mm::BF16>;
    k.sweepM[1][1] = &k_filter_synth<BW, BH, WPW, false, true, mm::BF16>;
    k.analysisB[1][0] = &k_analysis_mfma_batch<mm::F16, false>;
    k.analysisB[1][1] = &k_analysis_mfma_batch<mm::F16, true>;
    k.analysisB[2][0] = &k_analysis_mfma_batch<mm::BF16, false>;
    k.analysisB[2][1] = &k_analysis_mfma_batch<mm::BF16, true>;
    k.sweepB[1][0] = &k_sweep_batch<BW, BH, WPW, false, mm::F16>;
    k.sweepB[1][1] = &k_sweep_batch<BW, BH, WPW, true, mm::F16>;
    k.sweepB[2][0] = &k_sweep_batch<BW, BH, WPW, false, mm::BF16>;
    k.sweepB[2][1] = &k_sweep_batch<BW, BH, WPW, true, mm::BF16>;
  }
  k.sweepOcc[0][0] = k.sweep;
  k.sweepOcc[0][1] = k.sweepH;
  k.sweepBOcc[0][0] = k.sweepB[0][0];
  k.sweepBOcc[0][1] = k.sweepB[0][1];
  if constexpr (BW == BH && (BW == 8 || BW == 16) && BY_LDS >= 3) {
    k.sweepOcc[1][0] = &k_filter_synth<BW, BH, WPW, false, false, 0, 3>;
    k.sweepBOcc[1][0] = &k_sweep_batch<BW, BH, WPW, false, 0, 3>;
    if constexpr (BW >= 16) {
      k.sweepOcc[1][1] = &k_filter_synth<BW, BH, WPW, false, true, 0, 3>;
      k.sweepBOcc[1][1] = &k_sweep_batch<BW, BH, WPW, true, 0, 3>;
    }
    k.maxOcc = 3;
  }
  if constexpr (BW == BH && BW == 8 && BY_LDS >= 4) {
    k.sweepOcc[2][0] = &k_filter_synth<BW, BH, WPW, false, false, 0, 4>;
    k.sweepBOcc[2][0] = &k_sweep_batch<BW, BH, WPW, false, 0, 4>;
    k.maxOcc = 4;
  }
  k.G = Shape<BW, BH>::G;
  k.sweepThreads = 64 * WPW;
  k.nbw = SweepShape<BW, BH, WPW>::NBW;
  k.sweepPerCU = BY_LDS < BY_WAVES ? BY_LDS : BY_WAVES;
  return k;
}

}  // namespace

template <int BW, int BH>
KernelSet kernel_set() {
  return make_set<BW, BH>();
}

template KernelSet kernel_set<FK_BW, FK_BH>();

}  // namespace fk
