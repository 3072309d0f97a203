// s512_synth.hip -- synthetic inputs generated on the device (benchmark and
// test utility; no counterpart in the reference).  Same definitions as the
// oracle's generators (SplitMix64 counter hash of (seed, index)); Box-Muller
// uses the device's log/sqrt/cos, so normal keys may differ from the host
// restatement in the last ulp -- parity checks always copy the exact device
// input to the host.
#include "s512_common.hpp"
#include "s512_runtime.hpp"
#include "../../include/s512_synth.h"

namespace s512 {

__device__ __forceinline__ uint64_t sm64(uint64_t seed, uint64_t i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t i) { return (double)(sm64(seed, i) >> 11) * 0x1.0p-53; }

__global__ void k_fill_f64(double* out, uint64_t n, int kind, uint64_t seed, uint64_t param) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        double v;
        switch (kind) {
            case S512X_UNIFORM: v = u01(seed, i); break;
            case S512X_NORMAL: {
                double u1 = u01(seed, 2 * i), u2 = u01(seed, 2 * i + 1);
                if (u1 < 0x1.0p-53) u1 = 0x1.0p-53;
                v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
                break;
            }
            case S512X_DISTINCT: v = (double)(sm64(seed, i) % (param ? param : 1)); break;
            case S512X_SORTED: v = (double)i / (double)n; break;
            case S512X_REVERSED: v = (double)(n - 1 - i) / (double)n; break;
            case S512X_MIXTURE: {
                // 90%: one of 8 Gaussian components (equal weights), means in
                // [-20, 20) and sd in [0.02, 2.02) drawn from the seed; 10%:
                // lognormal(mu = 0, sigma = 2)
                const double u0 = u01(seed, 3 * i);
                double u1 = u01(seed, 3 * i + 1), u2 = u01(seed, 3 * i + 2);
                if (u1 < 0x1.0p-53) u1 = 0x1.0p-53;
                const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
                if (u0 < 0.1) {
                    v = exp(2.0 * z);
                } else {
                    const uint64_t k = (uint64_t)((u0 - 0.1) / 0.9 * 8.0) & 7u;
                    const double mean = -20.0 + 40.0 * u01(seed ^ 0xA5A5ull, 1000 + k);
                    const double su = u01(seed ^ 0xA5A5ull, 2000 + k);
                    v = mean + (0.02 + 2.0 * su * su) * z;
                }
                break;
            }
            case S512X_PIECEWISE: {
                // param pieces (default 64) of equal mass, widths alternating
                // 1 and 1e-9: piece p starts after ceil(p/2) wide and floor(p/2)
                // narrow pieces; uniform inside the piece
                const uint64_t K = param ? param : 64;
                const uint64_t pc = sm64(seed, 2 * i) % K;
                const double w = (pc & 1) ? 1e-9 : 1.0;
                const double x0 = (double)((pc + 1) / 2) + 1e-9 * (double)(pc / 2);
                v = x0 + w * u01(seed, 2 * i + 1);
                break;
            }
            default: v = 0.0; break;
        }
        out[i] = v;
    }
}

// Zipf(s) on ranks 1..N by rejection-inversion (W. Hormann, G. Derflinger,
// "Rejection-inversion to generate variates from monotone discrete
// distributions", ACM TOMACS 6(3), 1996): invert the integral H of the hat
// h(x) = x^-s and accept k = round(x) when x falls under the histogram bar of k.
struct ZipfGen {
    double s, hx1, hn, sq;
    __device__ double H(double x) const { return (exp((1.0 - s) * log(x)) - 1.0) / (1.0 - s); }
    __device__ double Hinv(double y) const { return exp(log1p(y * (1.0 - s)) / (1.0 - s)); }
    __device__ double h(double x) const { return exp(-s * log(x)); }
    __device__ ZipfGen(double s_, double n) : s(s_) {
        hx1 = H(1.5) - 1.0;
        hn = H(n + 0.5);
        sq = 2.0 - Hinv(H(2.5) - h(2.0));
    }
    __device__ uint64_t draw(uint64_t seed, uint64_t i, uint64_t n) const {
        double x = 1.0;
        for (uint64_t t = 0; t < 64; ++t) {  // acceptance is > 90% per try
            const double u = hn + u01(seed, i * 64 + t) * (hx1 - hn);
            x = Hinv(u);
            double k = floor(x + 0.5);
            k = k < 1.0 ? 1.0 : (k > (double)n ? (double)n : k);
            if (k - x <= sq || u >= H(k + 0.5) - h(k)) return (uint64_t)k;
        }
        return (uint64_t)fmin(fmax(floor(x + 0.5), 1.0), (double)n);
    }
};

// A seeded bijection of [0, n): xor-shift / odd-multiply rounds on the
// smallest power-of-two domain holding n, cycle-walking back into [0, n).
__device__ __forceinline__ uint64_t perm_mod(uint64_t x, uint64_t n, uint64_t seed) {
    int b = 1;
    while ((1ull << b) < n) ++b;
    const uint64_t mask = (1ull << b) - 1;
    const uint64_t m1 = (sm64(seed, 1) | 1ull) & mask, m2 = (sm64(seed, 2) | 1ull) & mask;
    const uint64_t k1 = sm64(seed, 3) & mask, k2 = sm64(seed, 4) & mask;
    const int sh = b / 2 > 0 ? b / 2 : 1;
    do {
        x = (x ^ k1) & mask;
        x = (x * m1) & mask;
        x ^= x >> sh;
        x = (x * m2) & mask;
        x ^= x >> sh;
        x = (x + k2) & mask;
    } while (x >= n);
    return x;
}

__global__ void k_fill_zipf(int32_t* out, uint64_t n, uint64_t seed, uint64_t nranks) {
    const ZipfGen z(1.1, (double)nranks);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)perm_mod(z.draw(seed, i, nranks) - 1, nranks, seed ^ 0x5EEDull);
}

__global__ void k_fill_i32(int32_t* out, uint64_t n, int kind, uint64_t seed, uint64_t param) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        int32_t v;
        switch (kind) {
            case S512X_UNIFORM: v = (int32_t)(uint32_t)(sm64(seed, i) >> 32); break;
            case S512X_DISTINCT: v = (int32_t)(sm64(seed, i) % (param ? param : 1)); break;
            case S512X_SORTED: v = (int32_t)(i - n / 2); break;
            case S512X_REVERSED: v = (int32_t)(n / 2 - i); break;
            case S512X_INDEX: v = (int32_t)i; break;
            default: v = 0; break;
        }
        out[i] = v;
    }
}

}  // namespace s512

using namespace s512;

static unsigned grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

extern "C" s512_status s512x_fill_f64(double* d, size_t n, int kind, uint64_t seed, uint64_t param, void* stream) {
    if (!d && n) { set_error("null array"); return S512_ERR_INVALID; }
    if (kind < S512X_UNIFORM || kind > S512X_PIECEWISE || kind == S512X_INDEX) {  /* zipf: i32 only */
        set_error("unknown f64 distribution %d", kind);
        return S512_ERR_INVALID;
    }
    if (!n) return S512_OK;
    k_fill_f64<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(d, n, kind, seed, param);
    S512_CHECK_LAUNCH();
    return S512_OK;
}

extern "C" s512_status s512x_fill_i32(int32_t* d, size_t n, int kind, uint64_t seed, uint64_t param, void* stream) {
    if (!d && n) { set_error("null array"); return S512_ERR_INVALID; }
    if (kind == S512X_NORMAL || kind < S512X_UNIFORM || kind > S512X_ZIPF || kind == S512X_MIXTURE ||
        kind == S512X_PIECEWISE) {
        set_error("unknown i32 distribution %d", kind);
        return S512_ERR_INVALID;
    }
    if (kind == S512X_ZIPF && param > (1ull << 31)) {
        set_error("zipf: %llu ranks exceed the int32 key range", (unsigned long long)param);
        return S512_ERR_INVALID;
    }
    if (!n) return S512_OK;
    if (kind == S512X_ZIPF) {
        k_fill_zipf<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(d, n, seed, param ? param : (1ull << 24));
        S512_CHECK_LAUNCH();
        return S512_OK;
    }
    k_fill_i32<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(d, n, kind, seed, param);
    S512_CHECK_LAUNCH();
    return S512_OK;
}
