// pt_kernels_any.hip -- the PT-ICM sweep kernel for launches that hold a group whose spin planes
// stay in global memory (too large for LDS), and its launch wrapper.  The kernel is the same source
// as pt_sweep_kernel, instantiated with generic spin pointers; it is a translation unit of its own
// so that the two instantiations compile in parallel.
#define TAMC_SWEEP_KERNEL_ANY
#include "pt_kernels.hip"
