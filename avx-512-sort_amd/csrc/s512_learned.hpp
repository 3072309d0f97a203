// s512_learned.hpp -- learned partitioning: a monotone 1-32-32-1 fp32 MLP that
// models the CDF of the keys, trained on the GPU on a ~1% sample, and a CDF
// table that puts every key in its bucket in O(1) (no search over splitters)
// in the count and scatter passes of the first two partition levels.  No
// counterpart in the reference.
//
// Model (same definition as the oracle's fp64 restatement):
//   u  = feature of the key: (float)(enc(x) - kmin) / (kmax - kmin), clamped
//        to [0,1]; enc is the order-preserving integer encoding, kmin/kmax the
//        sample's smallest/largest encoded key.  Every step is monotone.
//   h1 = r(p1^2 * u + b1), h2 = r(P2^2 h1 + b2), y = p3^2 . h2 + b3,
//        r(z) = min(max(z,0),1).  Squared weights make y non-decreasing in u.
// CDF table (CT_WORDS u32 words), a shallow tree of uniform cells in encoded
// key space: CT_TOP top cells of width 2^shift from an origin at a multiple of
// 2^shift, covering the sampled range [kmin, kmax]; a heavy cell is split into
// CT_FAN equal sub-cells (a node), and a heavy sub-cell again, down to
// CT_DEPTH levels below the top, at most CT_NODES nodes in all.  Every entry
// holds the CDF (27-bit fixed point) at its cell's lower boundary and the
// node that splits the cell (0: none).  A key's CDF value is the linear
// interpolation inside the deepest cell that holds it: one table read per
// level, one multiply -- no search.  The first level's table holds the
// sample's own CDF; the second level's holds the trained network's CDF at the
// same boundaries, kept within a quarter of the adjacent cells' sampled mass
// of the sample's (the network smooths inside the sample's resolution).
// Buckets: at the root level of fan-out F, y = CDF(x) gives bucket y F / 2^32.
// A root bucket's keys are split again by a local table of the bucket
// (LT_CELLS uniform cells over its key range, origin on a multiple of the
// cell width): the network's CDF at the cell boundaries, renormalised so the
// bucket's bounds (where the sample's table crosses them) map to 0 and 1;
// bucket y_local F2 / 2^32, one table read per key.  Both are non-decreasing
// functions of the key, so the buckets are ordered key ranges whatever the
// model's accuracy -- the model only changes bucket sizes, never the sorted
// result.
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

// The model as the partition levels use it, in device memory: the root's CDF
// table (first level) and the root buckets' local tables (second level).
struct LearnedModelDev {
    const uint32_t* tab1;
    const uint32_t* tab2;
};

template <class K>
__device__ __forceinline__ float learned_feature(K x, uint64_t kmin, float inv_range) {
    if ((uint64_t)x <= kmin) return 0.0f;
    return fminf((float)((uint64_t)x - kmin) * inv_range, 1.0f);
}

// ---------------------------------------------------------------- CDF table
constexpr uint32_t CT_TOP = 1024;   // top cells
constexpr uint32_t CT_TOP_LOG = 10;
constexpr uint32_t CT_FAN = 256;    // sub-cells of a node
constexpr uint32_t CT_FAN_LOG = 8;
constexpr uint32_t CT_NODES = 21;   // nodes at most
constexpr uint32_t CT_DEPTH = 3;    // node levels below the top at most
constexpr uint32_t CT_CDF_BITS = 27;
constexpr uint32_t CT_CDF_MASK = (1u << CT_CDF_BITS) - 1u;
// words: origin lo, origin hi, shift, nodes, last node of each of the
// CT_DEPTH node levels, 1 spare | top[CT_TOP + 1] | node k (1..CT_NODES):
// [CT_FAN + 1]
constexpr uint32_t CT_OFF_TOP = 8;
constexpr uint32_t CT_OFF_NODE = CT_OFF_TOP + CT_TOP + 1;
constexpr uint32_t CT_WORDS = CT_OFF_NODE + CT_NODES * (CT_FAN + 1);
__host__ __device__ constexpr uint32_t ct_node(uint32_t k) { return CT_OFF_NODE + (k - 1) * (CT_FAN + 1); }

// CDF value (2^32 = 1) of encoded key x.
template <class K>
__device__ __forceinline__ uint32_t table_cdf(const uint32_t* tb, K xk) {
    const uint64_t x = (uint64_t)xk;
    const uint64_t origin = (uint64_t)tb[0] | ((uint64_t)tb[1] << 32);
    const uint32_t shift = tb[2];
    const uint64_t d = x > origin ? x - origin : 0ull;
    const uint64_t c64 = d >> shift;
    // keys above the range: the last cell's last position
    const uint32_t c = c64 < CT_TOP ? (uint32_t)c64 : CT_TOP - 1;
    uint64_t rem = c64 < CT_TOP ? d - (c64 << shift) : (1ull << shift) - 1ull;
    uint32_t w = shift;  // bits of rem
    uint32_t at = CT_OFF_TOP + c;
    uint32_t e0 = tb[at];
#pragma unroll
    for (uint32_t lv = 0; lv < CT_DEPTH; ++lv) {
        const uint32_t child = e0 >> CT_CDF_BITS;
        if (!child) break;
        const uint32_t i = (uint32_t)(rem >> (w - CT_FAN_LOG));
        w -= CT_FAN_LOG;
        rem &= (1ull << w) - 1ull;
        at = ct_node(child) + i;
        e0 = tb[at];
    }
    const uint32_t e1 = tb[at + 1];
    const uint32_t f16 = w >= 16 ? (uint32_t)(rem >> (w - 16)) : (uint32_t)(rem << (16 - w));
    const uint32_t lo = e0 & CT_CDF_MASK, hi = e1 & CT_CDF_MASK;
    return (lo + (uint32_t)(((uint64_t)(hi - lo) * f16) >> 16)) << (32 - CT_CDF_BITS);
}

// Node levels of a table (0..CT_DEPTH): header words 4 + l - 1 hold the last
// node of level l.
__device__ __forceinline__ uint32_t table_depth(const uint32_t* tb) {
    return tb[6] > tb[5] ? 3u : (tb[5] > tb[4] ? 2u : (tb[4] ? 1u : 0u));
}

// Local table of a root bucket: origin (2 words), shift, 1 spare, then the
// normalised CDF (2^32 = 1, capped at 2^32 - 1) at LT_CELLS + 1 boundaries.
constexpr uint32_t LT_CELLS = 512;
constexpr uint32_t LT_WORDS = 4 + LT_CELLS + 1;

// Bucket from a CDF value y (2^32 = 1) of the root table: y F / 2^32 (ordinary
// buckets only: id 2 b).
__device__ __forceinline__ uint32_t table_bucket_of(uint32_t y, uint32_t logf) { return 2u * (y >> (32 - logf)); }

// Bucket (id 2 b) of encoded key x inside a root bucket, from its local table.
template <class K>
__device__ __forceinline__ uint32_t local_bucket(const uint32_t* lt, uint32_t logf, K xk) {
    const uint64_t x = (uint64_t)xk;
    const uint64_t origin = (uint64_t)lt[0] | ((uint64_t)lt[1] << 32);
    const uint32_t shift = lt[2];
    const uint64_t d = x > origin ? x - origin : 0ull;
    const uint64_t c64 = d >> shift;
    const uint32_t c = c64 < LT_CELLS ? (uint32_t)c64 : LT_CELLS - 1;
    const uint64_t rem = c64 < LT_CELLS ? d - (c64 << shift) : (1ull << shift) - 1ull;
    const uint32_t f16 = shift >= 16 ? (uint32_t)(rem >> (shift - 16)) : (uint32_t)(rem << (16 - shift));
    const uint32_t lo = lt[4 + c], hi = lt[4 + c + 1];
    const uint32_t y = lo + (uint32_t)(((uint64_t)(hi - lo) * f16) >> 16);
    return 2u * (y >> (32 - logf));
}

template <class K>
__device__ __forceinline__ uint32_t table_bucket(const uint32_t* tb, uint32_t logf, K x) {
    return table_bucket_of(table_cdf(tb, x), logf);
}

// Buckets of NB keys at once, without branches per key: every key takes the
// same steps down the table (a key in an unsplit cell stays on its entry), so
// the table reads of the NB keys overlap level by level; the steps stop when
// no key of the wave is in a split cell.
template <class K, int NB>
__device__ __forceinline__ void table_buckets(const uint32_t* tb, uint32_t logf, const K (&xk)[NB],
                                              uint32_t (&bk)[NB]) {
    const uint64_t origin = (uint64_t)tb[0] | ((uint64_t)tb[1] << 32);
    const uint32_t shift = tb[2];
    const uint32_t depth = table_depth(tb);
    uint32_t at[NB], e0[NB], w[NB];
    uint64_t rem[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const uint64_t x = (uint64_t)xk[u];
        const uint64_t d = x > origin ? x - origin : 0ull;
        const uint64_t c64 = d >> shift;
        const uint32_t c = c64 < CT_TOP ? (uint32_t)c64 : CT_TOP - 1;
        rem[u] = c64 < CT_TOP ? d - (c64 << shift) : (1ull << shift) - 1ull;
        w[u] = shift;
        at[u] = CT_OFF_TOP + c;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) e0[u] = tb[at[u]];
    for (uint32_t lv = 0; lv < depth; ++lv) {
        bool deeper = false;
#pragma unroll
        for (int u = 0; u < NB; ++u) deeper |= (e0[u] >> CT_CDF_BITS) != 0;
        if (!__any(deeper)) break;  // the whole wave's keys are in unsplit cells
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const uint32_t child = e0[u] >> CT_CDF_BITS;
            const uint32_t sh = child ? w[u] - CT_FAN_LOG : 0u;
            const uint32_t i = (uint32_t)(rem[u] >> sh) & (CT_FAN - 1u);
            rem[u] = child ? rem[u] & ((1ull << sh) - 1ull) : rem[u];
            w[u] = child ? sh : w[u];
            at[u] = child ? ct_node(child) + i : at[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) e0[u] = tb[at[u]];
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const uint32_t e1 = tb[at[u] + 1];
        const uint32_t f16 = w[u] >= 16 ? (uint32_t)(rem[u] >> (w[u] - 16)) : (uint32_t)(rem[u] << (16 - w[u]));
        const uint32_t lo = e0[u] & CT_CDF_MASK, hi = e1 & CT_CDF_MASK;
        const uint32_t y = (lo + (uint32_t)(((uint64_t)(hi - lo) * f16) >> 16)) << (32 - CT_CDF_BITS);
        bk[u] = table_bucket_of(y, logf);
    }
}

}  // namespace s512
