// 2D instantiations of the directional sweep kernels, compiled apart from
// bx_explicit.hip so that the dimensions build in parallel.
#define BX_SWEEP_PART 2
#include "bx_explicit.hip"
