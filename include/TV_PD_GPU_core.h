/* TV_PD_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/headers/TV_PD_GPU_core.h:7); declares TV_PD_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_TV_PD_GPU_CORE_H
#define CILREG_TV_PD_GPU_CORE_H
#include "cilregcuda.h"
#endif
