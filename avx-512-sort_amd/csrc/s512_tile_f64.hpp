// s512_tile_f64.hpp -- fp64 leaf sorts for inputs without NaN and -0.0.
#pragma once
#include <hip/hip_runtime.h>
#include "s512_kernels.hpp"

namespace s512 {
// Same classes as the 64-bit integer tile kernels (caps 256 .. 8192).
void launch_tile_f64_fast(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count,
                          double* a, const double* b, hipStream_t st);
// The same classes; each tile picks the min/max network when its own keys hold
// no NaN and no -0.0, the encoded integer network otherwise (classes <= 8192).
void launch_tile_f64_auto(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count,
                          double* a, const double* b, hipStream_t st);
}  // namespace s512
