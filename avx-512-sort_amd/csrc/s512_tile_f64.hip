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

void launch_tile_f64_fast(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count,
                          double* a, const double* b, hipStream_t st) {
    if (count == 0) return;
    const uint32_t gw = (count + 3) / 4;
#define S512_TF(PP, EE) k_tile_sort<double, PP, EE, F64Order><<<count, PP, 0, st>>>(segs, dcount, a, b)
    if (cls == 0) { k_wave_sort<double, 4, F64Order><<<gw, 256, 0, st>>>(segs, dcount, a, b); return; }
    switch (plan * 8 + cls) {
        case 1: S512_TF(128, 4); break;
        case 2: S512_TF(128, 8); break;
        case 3: S512_TF(256, 8); break;
        case 4: S512_TF(256, 16); break;
        case 8 + 1: S512_TF(64, 8); break;
        case 8 + 2: S512_TF(64, 16); break;
        case 8 + 3: S512_TF(128, 16); break;
        case 8 + 4: S512_TF(256, 16); break;
        case 16 + 1: S512_TF(64, 8); break;
        case 16 + 2: S512_TF(64, 16); break;
        case 16 + 3: S512_TF(64, 32); break;
        case 16 + 4: S512_TF(128, 32); break;
        case 16 + 5: S512_TF(256, 32); break;
        case 6:
        case 8 + 6: S512_TF(1024, 16); break;
        case 16 + 6: S512_TF(512, 32); break;
        default: S512_TF(512, 16); break;
    }
#undef S512_TF
}

}  // namespace s512
