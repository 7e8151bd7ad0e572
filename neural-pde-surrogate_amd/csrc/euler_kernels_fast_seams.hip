// The seam-layout stage kernels of the fast-arithmetic build (see euler_kernels_fast.hip), in a
// translation unit of their own so that the Makefile compiles them with the ILP-oriented machine
// scheduler (stage_kernel.hpp, pick_stage_seams).
#define NPDE_FAST_ARITH 1
#define NPDE_SEAMS_UNIT 1
#include "stage_kernel.hpp"
