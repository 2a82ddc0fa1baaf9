// One block shape's kernel unit (-DFK_BW=... -DFK_BH=...), compiled as host
// C++ against the HIP emulation.
#include "fk_set_bw.hip"
