/* TGV_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/TGV_GPU_core.h:8); declares TGV_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_TGV_GPU_CORE_H
#define CILREG_TGV_GPU_CORE_H
#include "cilregcuda.h"
#endif
