/*
 * s512_synth.h -- device-side synthetic inputs for benchmarks and tests.
 * No counterpart in the reference (its benchmark fills host arrays with
 * drand48, sort512perf.cpp); these fill device memory at HBM speed with the
 * distributions of the benchmark workloads.  Not part of the sort API.
 */
#ifndef S512_SYNTH_H
#define S512_SYNTH_H
#include "s512.h"
#ifdef __cplusplus
extern "C" {
#endif

enum {
    S512X_UNIFORM = 0,   /* f64: [0,1) from SplitMix64(seed,i); i32: full int32 range */
    S512X_NORMAL = 1,    /* f64 only: Box-Muller from uniforms (2i, 2i+1)             */
    S512X_DISTINCT = 2,  /* SplitMix64(seed,i) % param                                 */
    S512X_SORTED = 3,    /* f64: i/n;  i32: i - n/2                                    */
    S512X_REVERSED = 4,  /* f64: (n-1-i)/n;  i32: n/2 - i                              */
    S512X_INDEX = 5,     /* i32 only: i                                                */
    S512X_MIXTURE = 6,   /* f64 only: 8-component Gaussian mixture + 10% lognormal(0,2) */
    S512X_PIECEWISE = 7, /* f64 only: piecewise-constant density, param pieces of equal
                            mass with widths alternating 1 and 1e-9 (default 64)     */
    S512X_ZIPF = 8       /* i32 only: Zipf rank r in 1..N (P(r) ~ r^-1.1, N = param,
                            default 2^24; rejection-inversion sampling), mapped to
                            the key perm(r - 1) by a seeded bijection of [0, N)    */
};

s512_status s512x_fill_f64(double* d_out, size_t n, int kind, uint64_t seed, uint64_t param, void* stream);
s512_status s512x_fill_i32(int32_t* d_out, size_t n, int kind, uint64_t seed, uint64_t param, void* stream);

#ifdef __cplusplus
}
#endif
#endif
