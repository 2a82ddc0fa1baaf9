// One block width's kernel unit (-DFK_BW=...), compiled as host C++ against
// the HIP emulation.
#include "fk_set_bw.hip"
