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
// is the network's prediction at that key, kept within half the sample mass
// of the two adjacent pieces around the empirical value (the network smooths
// within the sample's resolution, it cannot move a knot further), then made
// non-decreasing by a running maximum.  Equal-mass knots put the model's
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
};

template <class K>
__device__ __forceinline__ float learned_feature(K x, uint64_t kmin, float inv_range) {
    if ((uint64_t)x <= kmin) return 0.0f;
    return fminf((float)((uint64_t)x - kmin) * inv_range, 1.0f);
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
