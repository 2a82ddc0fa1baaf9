// The C ABI unit, compiled as host C++ against the HIP emulation.
#include "fft3d_capi.hip"
