// The seam-layout stage kernels of the exact build, in a translation unit of their own so that the
// Makefile compiles them with the ILP-oriented machine scheduler (stage_kernel.hpp, pick_stage_seams);
// the kernels themselves are those of stage_kernel.hpp, with the exact arithmetic of the other layouts.
#define NPDE_SEAMS_UNIT 1
#include "stage_kernel.hpp"
