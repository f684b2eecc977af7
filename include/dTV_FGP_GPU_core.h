/* dTV_FGP_GPU_core.h -- drop-in for the reference header of the same name
 * (src/Core/regularisers_GPU/dTV_FGP_GPU_core.h:7); declares dTV_FGP_GPU_main, defined in libcilregcuda.so.
 * All declarations live in cilregcuda.h. */
#ifndef CILREG_DTV_FGP_GPU_CORE_H
#define CILREG_DTV_FGP_GPU_CORE_H
#include "cilregcuda.h"
#endif
