/* LLT_ROF_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/LLT_ROF_GPU_core.h:6); declares LLT_ROF_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_LLT_ROF_GPU_CORE_H
#define CILREG_LLT_ROF_GPU_CORE_H
#include "cilregcuda.h"
#endif
