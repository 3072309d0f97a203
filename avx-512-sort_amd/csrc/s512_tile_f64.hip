// s512_tile_f64.hip -- fp64 leaf sorts on raw doubles with v_min_f64 /
// v_max_f64 compare-exchanges (two instructions per pair instead of a 64-bit
// integer compare and four selects).
//
// Built with -fno-honor-nans (see the Makefile), which lets fmin/fmax lower
// to bare v_min_f64/v_max_f64.  That is only exact for inputs with no NaN and
// no -0.0: the sort engine checks every key for both in its first count pass
// and uses these kernels only when neither occurs; otherwise the encoded
// integer kernels of s512_sort.hip run.  Pads are +inf (a real +inf key and a
// pad are the same value, so the first `size` outputs are unchanged).
#include "s512_kernels.hpp"
#include "s512_tile_f64.hpp"

namespace s512 {

void launch_tile_f64_fast(int cls, const TileSeg* segs, const uint32_t* dcount, uint32_t count, double* a,
                          const double* b, hipStream_t st) {
    if (count == 0) return;
    const uint32_t gw = (count + 3) / 4;
    switch (cls) {
        case 0: k_wave_sort<double, 4, F64Order><<<gw, 256, 0, st>>>(segs, dcount, a, b); break;
        case 1: k_tile_sort<double, 128, 4, F64Order><<<count, 128, 0, st>>>(segs, dcount, a, b); break;
        case 2: k_tile_sort<double, 128, 8, F64Order><<<count, 128, 0, st>>>(segs, dcount, a, b); break;
        case 3: k_tile_sort<double, 256, 8, F64Order><<<count, 256, 0, st>>>(segs, dcount, a, b); break;
        case 4: k_tile_sort<double, 256, 16, F64Order><<<count, 256, 0, st>>>(segs, dcount, a, b); break;
        default: k_tile_sort<double, 512, 16, F64Order><<<count, 512, 0, st>>>(segs, dcount, a, b); break;
    }
}

}  // namespace s512
