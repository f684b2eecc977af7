/* NonlDiff_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/NonlDiff_GPU_core.h:6); declares NonlDiff_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_NONLDIFF_GPU_CORE_H
#define CILREG_NONLDIFF_GPU_CORE_H
#include "cilregcuda.h"
#endif
