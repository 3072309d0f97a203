// s512_learned.hpp -- learned partitioning: a monotone 1-32-32-1 fp32 MLP that
// models the CDF of the keys, trained on the GPU on a sorted ~1% sample, then
// used to place the splitters of the first two partition levels.  No
// counterpart in the reference.
//
// Model (same definition as the oracle's fp64 restatement):
//   u  = feature of the key: (float)(enc(x) - kmin) / (kmax - kmin), clamped
//        to [0,1]; enc is the order-preserving integer encoding, kmin/kmax the
//        sample's smallest/largest encoded key.  Every step is monotone.
//   h1 = r(p1^2 * u + b1), h2 = r(P2^2 h1 + b2), y = p3^2 . h2 + b3,
//        r(z) = min(max(z,0),1).  Squared weights make y non-decreasing in u.
// Knots: the sorted sample is cut into LEARN_KNOTS pieces of equal mass; knot
// k is the sample key of rank r_k = k (ns-1) / LEARN_KNOTS and its CDF value
// is the network's prediction at that key, kept within a quarter of a knot's
// mass around the empirical value (the network smooths within the sample's
// resolution, it cannot move a knot further; neighbouring bands do not
// overlap), then made non-decreasing by a running maximum.  Equal-mass knots put the model's
// resolution where the keys are, however skewed they are.
// Splitters: the model's quantiles, i.e. the inverse of the piecewise-linear
// CDF through the knots (k_model_splitters).  A level partitioned by the model
// therefore uses the same classifier as sampled splitters; the model decides
// where the bucket boundaries fall, and since its quantiles are non-decreasing
// the buckets are ordered key ranges whatever its accuracy: the model only
// changes bucket sizes, never the sorted result.
#pragma once
#include "s512_common.hpp"

namespace s512 {

constexpr int MLP_H = 32;
constexpr int MLP_NP = 3 * MLP_H + MLP_H * MLP_H + MLP_H + 1;  // 1153
constexpr int MLP_OFF_B1 = MLP_H;
constexpr int MLP_OFF_P2 = 2 * MLP_H;
constexpr int MLP_OFF_B2 = 2 * MLP_H + MLP_H * MLP_H;
constexpr int MLP_OFF_P3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_P3 + MLP_H;
constexpr int TRAIN_P = 256;        // training workgroup (one point per thread)

__device__ __forceinline__ float ramp01(float z) { return fminf(fmaxf(z, 0.0f), 1.0f); }

// The model as the partition levels use it: LEARN_KNOTS + 1 knots (encoded
// key, CDF value), both non-decreasing, in device memory.
constexpr int LEARN_KNOTS = 8192;
struct LearnedModelDev {
    const uint64_t* key;
    const float* cdf;
    const void* smp = nullptr;  // the sorted sample (encoded keys) the knots are taken from
    uint32_t ns = 0;            // its length
    uint32_t kbytes = 8;        // bytes per sample key (4 or 8)
};

template <class K>
__device__ __forceinline__ float learned_feature(K x, uint64_t kmin, float inv_range) {
    if ((uint64_t)x <= kmin) return 0.0f;
    return fminf((float)((uint64_t)x - kmin) * inv_range, 1.0f);
}

// Position (knot interval lo-1..lo and fraction f) where the piecewise-linear
// CDF through the knots reaches q; the key there is key[lo-1] + f (key[lo] -
// key[lo-1]).  lo = 0: below the first knot (f = 0); lo = LEARN_KNOTS + 1:
// above the last.
__device__ __forceinline__ void model_position(const float* cdf, double q, uint32_t* lo_out, double* f_out) {
    if (q <= (double)cdf[0]) { *lo_out = 0; *f_out = 0.0; return; }
    if (q >= (double)cdf[LEARN_KNOTS]) { *lo_out = LEARN_KNOTS + 1; *f_out = 0.0; return; }
    uint32_t lo = 1, hi = LEARN_KNOTS;  // first knot with cdf >= q, in [1, LEARN_KNOTS]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((double)cdf[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    const double c0 = cdf[lo - 1], c1 = cdf[lo];
    const double f = c1 > c0 ? (q - c0) / (c1 - c0) : 1.0;
    *lo_out = lo;
    *f_out = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
}

// The CDF value another knot table over the same keys takes at that position.
__device__ __forceinline__ double model_cdf_at(const float* cdf, uint32_t lo, double f) {
    if (lo == 0) return cdf[0];
    if (lo > LEARN_KNOTS) return cdf[LEARN_KNOTS];
    return (double)cdf[lo - 1] + f * ((double)cdf[lo] - (double)cdf[lo - 1]);
}

// Encoded key at CDF value q: the sample key at the rank the model's
// position maps to.  Knot k sits at sample rank k (ns - 1) / LEARN_KNOTS, so
// a position (interval, fraction) interpolates the rank, not the key: the
// quantiles are actual sample keys, spaced as the model says inside a knot's
// interval, whatever the key encoding's spacing (the encoded space of doubles
// is close to logarithmic).
__device__ __forceinline__ uint64_t model_quantile_rank(const LearnedModelDev& m, double q) {
    uint32_t lo;
    double f;
    model_position(m.cdf, q, &lo, &f);
    const uint64_t last = (uint64_t)m.ns - 1;
    uint64_t r;
    if (lo == 0) r = 0;
    else if (lo > LEARN_KNOTS) r = last;
    else {
        const uint64_t r0 = ((uint64_t)(lo - 1) * last) / LEARN_KNOTS, r1 = ((uint64_t)lo * last) / LEARN_KNOTS;
        r = r0 + (uint64_t)(f * (double)(r1 - r0));
        if (r > r1) r = r1;
    }
    return m.kbytes == 8 ? ((const uint64_t*)m.smp)[r] : (uint64_t)((const uint32_t*)m.smp)[r];
}

// Encoded key at CDF value q of the piecewise-linear model through the knots.
__device__ __forceinline__ uint64_t model_quantile(const uint64_t* key, const float* cdf, double q) {
    if (q <= (double)cdf[0]) return key[0];
    if (q >= (double)cdf[LEARN_KNOTS]) return key[LEARN_KNOTS];
    uint32_t lo = 1, hi = LEARN_KNOTS;  // first knot with cdf >= q, in [1, LEARN_KNOTS]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((double)cdf[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    const double c0 = cdf[lo - 1], c1 = cdf[lo];
    const double f = c1 > c0 ? (q - c0) / (c1 - c0) : 1.0;
    const uint64_t k0 = key[lo - 1], k1 = key[lo];
    const uint64_t d = (uint64_t)((double)(k1 - k0) * (f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f)));
    return k0 + (d > k1 - k0 ? k1 - k0 : d);
}

}  // namespace s512
