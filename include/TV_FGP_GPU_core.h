/* TV_FGP_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/TV_FGP_GPU_core.h:7); declares TV_FGP_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_TV_FGP_GPU_CORE_H
#define CILREG_TV_FGP_GPU_CORE_H
#include "cilregcuda.h"
#endif
