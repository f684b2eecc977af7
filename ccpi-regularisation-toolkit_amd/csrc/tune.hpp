// tune.hpp -- the autotuner's interface (tune.hip).
#pragma once
#include "runtime.hpp"

namespace cilreg {

struct TuneChoice {
    int tile_y = 8;          // 8 or 4
    int zchunk = 0;          // planes per workgroup; 0 = occupancy-planned
    int depth = 1;           // iterations per pass
    double ms_per_iter = 0;  // measured kernel time per iteration of the choice
};

// Fastest (tile_y, zchunk, depth) for `method` (CILREG_ROF / _FGP / _PD /
// _TGV) on an nx x ny x nz volume with base's dual storage; iso = isotropic
// TV (FGP, PD).  fixed_ty: tile height 8 and a concrete chunk depth only (the
// grouped kernels' constraints), timed with the device full (the probe
// stacks copies of the extent along z).  Cached per process.
int autotune(int method, int nx, int ny, int nz, int iso, const Opts& base, bool fixed_ty, TuneChoice* out);

// o with the tuned tiling and depth when o.autotune is set and none of them
// was given explicitly; o unchanged otherwise (or if tuning fails).
Opts tuned_opts(int method, int nx, int ny, int nz, int iso, const Opts& o);

}  // namespace cilreg
