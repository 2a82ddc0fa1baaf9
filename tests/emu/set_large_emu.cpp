// The kernel sets of the shapes with a side above 64, compiled as host C++
// against the HIP emulation.
#include "fk_set_large.hip"
