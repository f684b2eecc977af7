/* Diffus_4thO_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/Diffus_4thO_GPU_core.h:6); declares Diffus4th_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_DIFFUS_4THO_GPU_CORE_H
#define CILREG_DIFFUS_4THO_GPU_CORE_H
#include "cilregcuda.h"
#endif
