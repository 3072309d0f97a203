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
#define S512_TFL(PP, EE, LP) k_tile_sort<double, PP, EE, F64Order, LP><<<count, PP, 0, st>>>(segs, dcount, a, b)
#define S512_TFW(PP, EE, LP, W) \
    k_tile_sort_wpe<double, PP, EE, F64Order, LP, W><<<count, PP, 0, st>>>(segs, dcount, a, b)
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
        // plan 2 keeps 32 keys per thread up to the 2048-key class only: its
        // 4096- and 8192-key tiles measured faster with 16 (2^28 keys, uniform and
        // normal: 115 against 140 us, 24 against 32 us per launch).  Lane phases
        // per shape (per launch): three for 512-2048 keys (2048 keys at 2^28:
        // 897 against 924 us with two), two for 4096 and 8192 keys; 4096 keys: 160
        // registers leave room for three waves per SIMD (2^28: 94 against 115 us
        // for the tile with every phase through LDS, built for three waves with
        // spills; 2^30: 1.86 against 2.17 ms)
        case 16 + 1: S512_TFL(64, 8, 3); break;
        case 16 + 2: S512_TFL(64, 16, 3); break;
        case 16 + 3: S512_TFL(64, 32, 3); break;
        case 16 + 4: S512_TFL(256, 16, 2); break;  // capped at four waves: 1.94 against 1.88 ms (2^30)
        case 16 + 5: S512_TFW(512, 16, 2, 4); break;  // two workgroups per CU (k_tile_sort_wpe)
        case 6:
        case 8 + 6: S512_TF(1024, 16); break;
        case 16 + 6: S512_TF(512, 32); break;
        default: S512_TF(512, 16); break;
    }
#undef S512_TFW
#undef S512_TFL
#undef S512_TF
}

// ---------------------------------------------------------------- per-tile choice
// For callers that have not looked at every key beforehand (the segmented
// sort: its small segments go straight to the tiles).  Each tile checks its
// own keys while loading them: with no NaN and no -0.0 it runs the min/max
// network on raw doubles, otherwise the encoded integer network; both give
// the same bit patterns.  The test is on the bits: under -fno-honor-nans
// `x != x` is no NaN test.
__device__ __forceinline__ bool f64_special(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    return (u & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull || u == 0x8000000000000000ull;
}

template <int P, int E>
__global__ __launch_bounds__(P) void k_tile_sort_f64_auto(const TileSeg* segs, const uint32_t* count, double* a,
                                                          const double* b) {
    using BF = BlockBitonic<F64Order, P, E>;
    using BU = BlockBitonic<UintOrder<uint64_t>, P, E>;
    using Tr = KeyTraits<double>;
    __shared__ uint64_t lds[BF::LDS_N];
    double* ldf = reinterpret_cast<double*>(lds);
    const uint32_t cnt = *count;
    const uint32_t t = threadIdx.x;
    const uint64_t inf_bits = 0x7FF0000000000000ull;
    for (uint32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const TileSeg sg = segs[s];
        const double* src = (sg.buf ? b : a) + sg.begin;
        double* dst = a + sg.begin;
        bool bad = false;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            const double x = src[i < sg.size ? i : sg.size - 1];
            bad |= i < sg.size && f64_special(x);
            lds[BF::phys(i)] = i < sg.size ? __builtin_bit_cast(uint64_t, x) : inf_bits;
        }
        if (__syncthreads_or(bad)) {
            // each thread re-encodes the slots it wrote; pads sort after every key
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const uint32_t i = m * P + t;
                lds[BU::phys(i)] = i < sg.size ? Tr::enc_bits(lds[BU::phys(i)]) : Tr::KMAX;
            }
            uint64_t v[E];
            BU::from_striped(v, lds, t);
            BU::sort(v, lds, t);
            BU::to_striped(v, lds, t);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const uint32_t i = m * P + t;
                if (i < sg.size) dst[i] = Tr::dec(lds[BU::phys(i)]);
            }
        } else {
            double v[E];
            BF::from_striped(v, ldf, t);
            BF::sort(v, ldf, t);
            BF::to_striped(v, ldf, t);
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const uint32_t i = m * P + t;
                if (i < sg.size) dst[i] = ldf[BF::phys(i)];
            }
        }
        __syncthreads();
    }
}

// One wave per segment (<= 256 keys), four waves per workgroup.
__global__ __launch_bounds__(256) void k_wave_sort_f64_auto(const TileSeg* segs, const uint32_t* count, double* a,
                                                            const double* b) {
    constexpr int E = 4;
    using Tr = KeyTraits<double>;
    const uint32_t cnt = *count;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6); s < cnt; s += gridDim.x * 4) {
        const TileSeg sg = segs[s];
        const double* src = (sg.buf ? b : a) + sg.begin;
        double* dst = a + sg.begin;
        double x[E];
        bool bad = false;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            x[r] = src[i < sg.size ? i : sg.size - 1];
            bad |= i < sg.size && f64_special(x[r]);
            if (i >= sg.size) x[r] = __builtin_huge_val();
        }
        if (__any(bad)) {
            uint64_t v[E];
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = lane * E + r < sg.size ? Tr::enc(x[r]) : Tr::KMAX;
            WaveBitonic<UintOrder<uint64_t>, E>::sort(v, lane);
#pragma unroll
            for (int r = 0; r < E; ++r)
                if (lane * E + r < sg.size) dst[lane * E + r] = Tr::dec(v[r]);
        } else {
            WaveBitonic<F64Order, E>::sort(x, lane);
#pragma unroll
            for (int r = 0; r < E; ++r)
                if (lane * E + r < sg.size) dst[lane * E + r] = x[r];
        }
    }
}

void launch_tile_f64_auto(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count,
                          double* a, const double* b, hipStream_t st) {
    if (count == 0) return;
    const uint32_t gw = (count + 3) / 4;
#define S512_TA(PP, EE) k_tile_sort_f64_auto<PP, EE><<<count, PP, 0, st>>>(segs, dcount, a, b)
    if (cls == 0) { k_wave_sort_f64_auto<<<gw, 256, 0, st>>>(segs, dcount, a, b); return; }
    switch (plan * 8 + cls) {
        case 1: S512_TA(128, 4); break;
        case 2: S512_TA(128, 8); break;
        case 3: S512_TA(256, 8); break;
        case 4: S512_TA(256, 16); break;
        case 8 + 1: S512_TA(64, 8); break;
        case 8 + 2: S512_TA(64, 16); break;
        case 8 + 3: S512_TA(128, 16); break;
        case 8 + 4: S512_TA(256, 16); break;
        case 16 + 1: S512_TA(64, 8); break;
        case 16 + 2: S512_TA(64, 16); break;
        case 16 + 3: S512_TA(64, 32); break;
        case 16 + 4: S512_TA(128, 32); break;
        case 16 + 5: S512_TA(256, 32); break;
        default: S512_TA(512, 16); break;  // class 5 of plans 0, 1
    }
#undef S512_TA
}

}  // namespace s512
