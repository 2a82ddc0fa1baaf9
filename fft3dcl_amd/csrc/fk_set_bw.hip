// The kernel set of every block height for one block width FK_BW (compiled
// once per width: -DFK_BW=4, 8, 16, 32 or 64).
#include "fk_set.h"

#ifndef FK_BW
#error "FK_BW (block width) must be defined"
#endif

namespace fk {
namespace {

// Sweep workgroups hold WPW waves; 4 unless the per-wave LDS of large blocks
// would leave room for only one workgroup per CU.
template <int BW, int BH>
KernelSet make_set() {
#ifdef FK_SWEEP_WAVES
  constexpr int WPW = FK_SWEEP_WAVES;
#else
  constexpr int WPW = (BW * BH >= 4096) ? 2 : 4;
#endif
  constexpr int LDS_BYTES = 4 * SweepShape<BW, BH, WPW>::LDS_FLOATS;
  constexpr int BY_LDS = 160 * 1024 / LDS_BYTES, BY_WAVES = 8 / WPW;
  KernelSet k;
  k.analysis = &k_analysis<BW, BH, false>;
  k.filter_inverse = &k_filter_inverse<BW, BH, false>;
  k.sweep = &k_filter_synth<BW, BH, WPW, false, false>;
  k.sweepFwd = &k_filter_synth<BW, BH, WPW, true, false>;
  if constexpr (BW == BH && BW >= 16) {
    k.analysisH = &k_analysis<BW, BH, true>;
    k.filterH = &k_filter_inverse<BW, BH, true>;
    k.sweepH = &k_filter_synth<BW, BH, WPW, false, true>;
  }
  k.G = Shape<BW, BH>::G;
  k.sweepThreads = 64 * WPW;
  k.nbw = SweepShape<BW, BH, WPW>::NBW;
  k.sweepPerCU = BY_LDS < BY_WAVES ? BY_LDS : BY_WAVES;
  return k;
}

}  // namespace

template <int BW>
KernelSet kernel_set(int bh) {
  switch (bh) {
  case 4: return make_set<BW, 4>();
  case 8: return make_set<BW, 8>();
  case 16: return make_set<BW, 16>();
  case 32: return make_set<BW, 32>();
  case 64: return make_set<BW, 64>();
  }
  return KernelSet{};
}

template KernelSet kernel_set<FK_BW>(int bh);

}  // namespace fk
