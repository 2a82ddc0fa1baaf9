// The kernel sets of the block shapes with a side above 64 (fk_large.h):
// bw, bh powers of two in [4, 512] with bw * bh <= 4096 and max(bw, bh) >= 128.
#include "fk_set.h"
#include "fk_large.h"

namespace fk {
namespace {

template <int BW, int BH>
KernelSet make_large() {
  KernelSet k;
  k.analysis = &k_analysis_lds<BW, BH>;
  k.filter_inverse = &k_filter_inverse_lds<BW, BH>;
  k.G = 1;
  k.fftThreads = lg::NT;
  return k;
}

}  // namespace

KernelSet kernel_set_large(int bw, int bh) {
#define FK_LARGE(W, H) \
  if (bw == W && bh == H) return make_large<W, H>();
  FK_LARGE(128, 4) FK_LARGE(128, 8) FK_LARGE(128, 16) FK_LARGE(128, 32)
  FK_LARGE(4, 128) FK_LARGE(8, 128) FK_LARGE(16, 128) FK_LARGE(32, 128)
  FK_LARGE(256, 4) FK_LARGE(256, 8) FK_LARGE(256, 16)
  FK_LARGE(4, 256) FK_LARGE(8, 256) FK_LARGE(16, 256)
  FK_LARGE(512, 4) FK_LARGE(512, 8) FK_LARGE(4, 512) FK_LARGE(8, 512)
#undef FK_LARGE
  return KernelSet{};
}

}  // namespace fk
