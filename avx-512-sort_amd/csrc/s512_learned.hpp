// s512_learned.hpp -- learned partitioning: a monotone 1-32-32-1 fp32 MLP that
// models the CDF of the keys, trained on the GPU on a sorted ~1% sample, then
// tabulated into a monotone table that predicts bucket indices.  No
// counterpart in the reference.
//
// Model (same definition as the oracle's fp64 restatement):
//   u  = feature of the key: (float)(enc(x) - kmin) / (kmax - kmin), clamped
//        to [0,1]; enc is the order-preserving integer encoding, kmin/kmax the
//        sample's smallest/largest encoded key.  Every step is monotone.
//   h1 = r(p1^2 * u + b1), h2 = r(P2^2 h1 + b2), y = p3^2 . h2 + b3,
//        r(z) = min(max(z,0),1).  Squared weights make y non-decreasing in u.
// Prediction: the model is evaluated on a grid of G+1 points and each grid
// value is calibrated against the sorted sample (second stage of a recursive
// model index): c_g = clamp(m_g, e_g - d_g, e_g + d_g), e_g the sample's
// empirical CDF at the grid point and d_g half the sample mass of the two
// adjacent cells.  The network thus only smooths within the sample's local
// resolution; where the sample has no key nearby the empirical value stands
// (an unconstrained network value there, carried by the running maximum
// below, would merge whole ranges into one bucket).  The table is
// made non-decreasing by a running maximum, and a key's CDF value is the table
// interpolated at u (clamped to the next knot, so the prediction is
// non-decreasing in the key despite rounding).  The fine bucket is
// fb = min(NB-1, floor(c * NB)); a partition level uses a bit field of fb.
// Because fb is non-decreasing in the key, buckets are ordered ranges of keys
// whatever the model's accuracy: the model only changes bucket sizes, never
// the sorted result.
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
constexpr int LEARN_G = 8192;       // table cells
constexpr int TRAIN_P = 256;        // training workgroup (one point per thread)

__device__ __forceinline__ float ramp01(float z) { return fminf(fmaxf(z, 0.0f), 1.0f); }

// Tabulated model used by the count / scatter kernels.
struct LearnedTable {
    const float* tab;   // [LEARN_G + 1], non-decreasing, in [0,1]
    uint64_t kmin;      // feature origin (encoded key)
    float inv_range;    // 1 / (kmax - kmin)
    uint32_t nb_log;    // log2 of the number of fine buckets
};

template <class K>
__device__ __forceinline__ float learned_feature(K x, uint64_t kmin, float inv_range) {
    if ((uint64_t)x <= kmin) return 0.0f;
    return fminf((float)((uint64_t)x - kmin) * inv_range, 1.0f);
}

// Fine bucket of key x from the tabulated model held in LDS.
template <class K>
__device__ __forceinline__ uint32_t learned_fine_bucket(const float* tab, uint64_t kmin, float inv_range,
                                                        uint32_t nb_log, K x) {
    const float u = learned_feature(x, kmin, inv_range);
    const float pos = u * (float)LEARN_G;
    uint32_t gi = (uint32_t)pos;
    gi = gi < LEARN_G - 1 ? gi : LEARN_G - 1;
    const float fr = pos - (float)gi;
    const float t0 = tab[gi], t1 = tab[gi + 1];
    const float c = fminf(fmaf(t1 - t0, fr, t0), t1);
    const float nb = (float)(1u << nb_log);
    const uint32_t fb = (uint32_t)(c * nb);
    return fb < (1u << nb_log) - 1 ? fb : (1u << nb_log) - 1;
}

}  // namespace s512
