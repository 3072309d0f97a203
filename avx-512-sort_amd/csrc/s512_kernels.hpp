// s512_kernels.hpp -- device kernels of the multiway sample-sort pipeline.
//
// Reference mapping.  The reference's Sort512::Sort is a recursive quicksort:
// pick a pivot (sort512.hpp:5916-5926), 2-way Partition512 (5729-5910),
// recurse, and sort leaves of <= 16 vectors with SmallSort16V (2255, 5312).
// Here one level of recursion partitions every open segment at once into
// F = 2^logf ordinary buckets plus one equality bucket per splitter (so runs of
// equal keys finish in O(n), where the reference's `<= pivot` partition goes
// quadratic), and leaves of up to one workgroup tile are sorted by the bitonic
// tile kernels of s512_tile.hpp.
//
// Pipeline of one level (all segments of the level in each launch):
//   k_sample      draw nsample keys per segment at hashed positions
//   (tile sort)   sort each segment's sample
//   k_build_tree  evenly spaced splitters, de-duplicated, Eytzinger tree
//   k_count       per chunk: classify, LDS histogram -> hist[chunk][bucket]
//   k_colscan     per segment: prefix over chunks per bucket, bucket starts
//   k_scatter     per chunk: classify, rank in LDS, stage by bucket in LDS,
//                 write coalesced runs to the other array
//   k_plan        route each bucket: done / tile sort class / next level
#pragma once
#include "s512_common.hpp"
#include "s512_tile.hpp"
#include "s512_learned.hpp"
#include <type_traits>

namespace s512 {

constexpr int FMAX_LOG = 9;                 // at most 512-way fan-out per level
constexpr int FMAX = 1 << FMAX_LOG;
constexpr int BMAX = 2 * FMAX;              // ordinary + equality buckets
constexpr int SCAT_ROUND = 8192;            // chunks are multiples of this many keys
constexpr int SCATTER_SHAPES = 4;           // 512x16, 256x16, 1024x8, 512x16 write-combining
constexpr int COUNT_SHAPES = 5;            // (threads, keys per batch, classified together), see launch_count
constexpr uint32_t COPY_PIECE = 1u << 16;   // keys per copy work item
constexpr int CLS_NB = 8;                   // keys classified together (overlapped LDS reads)

struct TileSeg {
    uint64_t begin;
    uint32_t size;
    uint32_t buf;
};

__device__ __forceinline__ uint64_t mix64(uint64_t seed, uint64_t i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------- block scan
// Exclusive scan of one value per thread; returns the prefix, sets *total.
template <int P>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wtot, uint32_t* total) {
    constexpr int NW = P / 64;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) wtot[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t s = wtot[w];
        off += (w < (int)wid) ? s : 0u;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return off + x - v;
}

// ---------------------------------------------------------------- classify
// Splitters of a segment: srt[0 .. F-1) sorted and de-duplicated (nuniq real
// ones, then KMAX pads), srt[F-1] = KMAX.  Bucket of key x: j = #{srt < x}
// (at most nuniq), ordinary bucket 2j, or the equality bucket 2j+1 when
// x == srt[j].
//
// j is found with a two-level guide over the encoded key range [kmin, kmax]
// of the segment's sample.  The guide only narrows the search: every key,
// inside that range or not, gets its exact j.
//   top[c], c = (x - kmin) >> shift clamped to [0, TOP-1]:
//     j0 | j1 << 16    j lies in [j0, j1] for every key of the cell, or
//     SUBF | h         more than GUIDE_DENSE splitters fall in the cell:
//   hdr[h] = {origin lo, origin hi, shift2, base | k << 16}: a sub-table of
//     2^k cells starting at origin = the cell's first splitter; its cell
//     q = (x - origin) >> shift2 (clamped) holds j0 | j1 << 16 as above.
// Encoded doubles are log-like, so a normal or lognormal sample crowds most
// splitters into a few top cells; their sub-tables spread them out again and
// the binary search in [j0, j1] stays at zero to two steps.  A segment whose
// top cells are all direct (uniform-like keys) skips the sub-table stage.
//
// Equality buckets are used only when the sample repeats a splitter value
// (nuniq < F-1, i.e. some key is frequent enough to span a bucket).  Without
// them a key equal to srt[j] goes to ordinary bucket 2j, which holds the keys
// in (srt[j-1], srt[j]]: buckets stay ordered either way.
constexpr uint32_t GUIDE_TOP_MAX = 2048;
constexpr uint32_t GUIDE_SUBTABLES = 1u;  // PartSeg::flags: some top cell has a sub-table
constexpr uint32_t GUIDE_EQBUCKETS = 2u;  // PartSeg::flags: keys equal to a splitter use 2j+1
constexpr uint32_t GUIDE_DENSE = 4;
constexpr uint32_t GUIDE_SUBF = 0x80000000u;
__host__ __device__ constexpr uint32_t guide_top(uint32_t F) { return 8 * F < GUIDE_TOP_MAX ? 8 * F : GUIDE_TOP_MAX; }
__host__ __device__ constexpr uint32_t guide_hdrs(uint32_t F) { return F / (GUIDE_DENSE + 1) + 1; }
__host__ __device__ constexpr uint32_t guide_words(uint32_t F) { return guide_top(F) + 4 * guide_hdrs(F) + 8 * F; }
constexpr uint32_t GMAX = guide_words(FMAX);

struct Guide {
    uint64_t kmin;
    uint32_t shift;
    uint32_t topmask;  // top cells - 1
    uint32_t hdr;      // word offset of the headers
    uint32_t sub;      // word offset of the sub-tables
    uint32_t flags;    // GUIDE_SUBTABLES | GUIDE_EQBUCKETS
};

__host__ __device__ inline Guide make_guide(uint64_t kmin, uint32_t shift, uint32_t F, uint32_t flags) {
    const uint32_t top = guide_top(F);
    return Guide{kmin, shift, top - 1, top, top + 4 * guide_hdrs(F), flags};
}

// Range [j0, j1] (packed) that holds j for key x.
template <class K>
__device__ __forceinline__ uint32_t guide_range(const uint32_t* gd, const Guide& g, K x) {
    const uint64_t xv = (uint64_t)x;
    const uint64_t c = xv < g.kmin ? 0 : (xv - g.kmin) >> g.shift;
    uint32_t e = gd[c > g.topmask ? g.topmask : (uint32_t)c];
    if (e & GUIDE_SUBF) {
        const uint4 h = *reinterpret_cast<const uint4*>(gd + g.hdr + 4 * (e & 0xFFFFu));
        const uint64_t org = (uint64_t)h.x | ((uint64_t)h.y << 32);
        const uint32_t m = (1u << (h.w >> 16)) - 1u;
        const uint64_t q = xv < org ? 0 : (xv - org) >> h.z;
        e = gd[g.sub + (h.w & 0xFFFFu) + (q > m ? m : (uint32_t)q)];
    }
    return e;
}

template <class K>
__device__ __forceinline__ uint32_t classify(const K* srt, const uint32_t* guide, const Guide& g, K x) {
    const uint32_t e = guide_range(guide, g, x);
    uint32_t lo = e & 0xFFFFu, hi = (e >> 16) & 0x7FFFu;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (srt[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return 2 * lo + ((g.flags & GUIDE_EQBUCKETS) && srt[lo] == x ? 1u : 0u);
}

template <class K>
__device__ __forceinline__ uint32_t lower_bound_lds(const K* srt, uint32_t n, K x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (srt[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Load a segment's classifier into LDS: splitters + guide table.
constexpr int LDS_TABLE_WORDS = GMAX;
static_assert(CT_WORDS <= LDS_TABLE_WORDS, "the learned CDF table fits the guide's LDS");
// LRN: a build for levels whose segments are all learned (mode 2; the host
// launches it for those levels only).  Each build carries one classifier, so
// neither pays registers (and occupancy) for the other's code.
template <class K, int NT, bool LRN>
__device__ __forceinline__ Guide load_classifier(const PartSeg& s, const K* trees, const uint32_t* guides, K* lsrt,
                                                 uint32_t* lguide) {
    if constexpr (LRN) {
        // the learned model's CDF table (s512_learned.hpp), no splitters: the
        // root's tree table, or below the root the bucket's local table
        const uint32_t words = s.lshift ? LT_WORDS : CT_WORDS;
        for (uint32_t j = threadIdx.x; j < words; j += NT) lguide[j] = guides[s.guide_base + j];
        return Guide{0, 0, 0, 0, 0, 0};
    }
    const uint32_t F = 1u << s.logf, G = guide_words(F);
    for (uint32_t j = threadIdx.x; j < F; j += NT) lsrt[j] = trees[s.tree_base + j];
    for (uint32_t j = threadIdx.x; j < G; j += NT) lguide[j] = guides[s.guide_base + j];
    return make_guide(s.kmin, s.shift, F, s.flags);
}


// Buckets of NB keys of one segment at once.  Each stage issues its LDS reads
// for all NB keys before any result is used and no stage branches per key, so
// the reads of different keys overlap (one classification is a chain of three
// to five dependent LDS reads).  Keys of direct top cells also run the
// sub-table stage (on header 0, which the tree build always writes) and keep
// their top entry.
template <class K, int NB>
__device__ __forceinline__ void classify_batch(const K* srt, const uint32_t* gd, const Guide& g, const K (&x)[NB],
                                               uint32_t (&bk)[NB]) {
    uint32_t e[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const uint64_t xv = (uint64_t)x[u];
        const uint64_t c = xv < g.kmin ? 0 : (xv - g.kmin) >> g.shift;
        e[u] = gd[c > g.topmask ? g.topmask : (uint32_t)c];
    }
    uint32_t lo[NB], len[NB];
    if (g.flags & GUIDE_SUBTABLES) {
        uint4 h[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const uint32_t hi = (e[u] & GUIDE_SUBF) ? (e[u] & 0xFFFFu) : 0u;
            h[u] = *reinterpret_cast<const uint4*>(gd + g.hdr + 4 * hi);
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const uint64_t xv = (uint64_t)x[u];
            const uint64_t org = (uint64_t)h[u].x | ((uint64_t)h[u].y << 32);
            const uint32_t m = (1u << (h[u].w >> 16)) - 1u;
            const uint64_t q = xv < org ? 0 : (xv - org) >> h[u].z;
            const uint32_t e2 = gd[g.sub + (h[u].w & 0xFFFFu) + (q > m ? m : (uint32_t)q)];
            e[u] = (e[u] & GUIDE_SUBF) ? e2 : e[u];
        }
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        lo[u] = e[u] & 0xFFFFu;
        len[u] = ((e[u] >> 16) & 0x7FFFu) - lo[u] + 1;
    }
    // j is in [lo, lo + len) and srt[lo + len - 1] >= x: halve until one is left
    for (;;) {
        uint32_t mx = 0;
#pragma unroll
        for (int u = 0; u < NB; ++u) mx = len[u] > mx ? len[u] : mx;
        if (mx <= 1) break;
        K sv[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const uint32_t half = len[u] >> 1;
            sv[u] = srt[lo[u] + (half ? half - 1 : 0)];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const uint32_t half = len[u] >> 1;
            lo[u] += (half && sv[u] < x[u]) ? half : 0u;
            len[u] -= half;
        }
    }
    if (g.flags & GUIDE_EQBUCKETS) {
        K ev[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) ev[u] = srt[lo[u]];
#pragma unroll
        for (int u = 0; u < NB; ++u) bk[u] = 2 * lo[u] + (ev[u] == x[u] ? 1u : 0u);
    } else {
#pragma unroll
        for (int u = 0; u < NB; ++u) bk[u] = 2 * lo[u];
    }
}


// Buckets of keys of segment s: splitters and guide (modes 0, 1), or in the
// learned build the model's CDF table, O(1) per key (mode 2: ordinary buckets
// only).
template <class K, int NB, bool LRN>
__device__ __forceinline__ void classify_keys(const PartSeg& s, const K* srt, const uint32_t* gd, const Guide& g,
                                              const K (&x)[NB], uint32_t (&bk)[NB]) {
    if constexpr (LRN) {
        if (s.lshift) {
#pragma unroll
            for (int u = 0; u < NB; ++u) bk[u] = local_bucket(gd, s.logf, x[u]);
        } else {
            table_buckets<K, NB>(gd, s.logf, x, bk);
        }
    } else {
        classify_batch<K, NB>(srt, gd, g, x, bk);
    }
}
template <bool LRN, class K>
__device__ __forceinline__ uint32_t classify_key(const PartSeg& s, const K* srt, const uint32_t* gd, const Guide& g,
                                                 K x) {
    if constexpr (LRN) return s.lshift ? local_bucket(gd, s.logf, x) : table_bucket(gd, s.logf, x);
    else return classify(srt, gd, g, x);
}

// ---------------------------------------------------------------- tile kernels
// How keys of type T enter and leave a network ordered by policy Pol.
template <class T, class Pol>
struct KeyIO {  // encoded unsigned keys
    using K = typename Pol::key_t;
    __device__ static __forceinline__ K load(T x) { return KeyTraits<T>::enc(x); }
    __device__ static __forceinline__ T store(K k) { return KeyTraits<T>::dec(k); }
    __device__ static __forceinline__ K pad() { return KeyTraits<T>::KMAX; }
};
template <>
struct KeyIO<double, F64Order> {  // raw doubles (no NaN, no -0.0)
    __device__ static __forceinline__ double load(double x) { return x; }
    __device__ static __forceinline__ double store(double k) { return k; }
    __device__ static __forceinline__ double pad() { return __builtin_huge_val(); }
};
template <class T>
using DefaultOrder = UintOrder<typename KeyTraits<T>::K>;

// Sort every listed segment (size <= P*E) in one workgroup; result goes to the
// caller's array `a` at the same range.
template <class T, int P, int E, class Pol, int LP = LANE_PHASES>
__device__ __forceinline__ void tile_sort_body(const TileSeg* segs, const uint32_t* count, T* a, const T* b) {
    using IO = KeyIO<T, Pol>;
    using K = typename Pol::key_t;
    using BB = BlockBitonic<Pol, P, E, LP>;
    __shared__ K lds[BB::LDS_N];
    const uint32_t cnt = *count;
    const uint32_t t = threadIdx.x;
    for (uint32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const TileSeg sg = segs[s];
        const T* src = (sg.buf ? b : a) + sg.begin;
        T* dst = a + sg.begin;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            const K x = IO::load(src[i < sg.size ? i : sg.size - 1]);
            lds[BB::phys(i)] = i < sg.size ? x : IO::pad();
        }
        K v[E];
        BB::from_striped(v, lds, t);
        BB::sort(v, lds, t);
        BB::to_striped(v, lds, t);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) dst[i] = IO::store(lds[BB::phys(i)]);
        }
        __syncthreads();
    }
}
template <class T, int P, int E, class Pol = DefaultOrder<T>, int LP = LANE_PHASES>
__global__ __launch_bounds__(P) void k_tile_sort(const TileSeg* segs, const uint32_t* count, T* a, const T* b) {
    tile_sort_body<T, P, E, Pol, LP>(segs, count, a, b);
}

// The same with a floor of W waves per SIMD (registers capped at 512 / W, some
// spill to scratch): for tiles whose LDS holds more workgroups per CU than
// their registers would.  The 8192-key raw-double tile (512 x 16, 64 KiB of
// LDS) runs two workgroups per CU this way instead of one: 5.41 against
// 6.13 ms per 2^30-key sort.
template <class T, int P, int E, class Pol, int LP, int W>
__global__ __launch_bounds__(P) __attribute__((amdgpu_waves_per_eu(W))) void k_tile_sort_wpe(const TileSeg* segs,
                                                                                           const uint32_t* count,
                                                                                           T* a, const T* b) {
    tile_sort_body<T, P, E, Pol, LP>(segs, count, a, b);
}

// Sample sorts of many segments whose sample sizes differ (the segmented
// sort's large segments: 16 F samples for fan-outs F from 2 to 512).  One
// launch per network size P*E sorts the listed segments with lo < size <=
// P*E and skips the others, so a sample runs the smallest network that holds
// it instead of being padded to the largest.  (One kernel choosing among the
// four networks per segment needs their registers at once, and spills.)
template <class T, int P, int E, class Pol = DefaultOrder<T>>
__global__ __launch_bounds__(P) void k_tile_sort_range(const TileSeg* segs, const uint32_t* count, T* a, const T* b,
                                                       uint32_t lo) {
    using IO = KeyIO<T, Pol>;
    using K = typename Pol::key_t;
    using BB = BlockBitonic<Pol, P, E>;
    __shared__ K lds[BB::LDS_N];
    const uint32_t cnt = *count;
    const uint32_t t = threadIdx.x;
    for (uint32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const TileSeg sg = segs[s];
        if (sg.size <= lo || sg.size > (uint32_t)BB::N) continue;  // uniform: the whole workgroup skips
        const T* src = (sg.buf ? b : a) + sg.begin;
        T* dst = a + sg.begin;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            const K x = IO::load(src[i < sg.size ? i : sg.size - 1]);
            lds[BB::phys(i)] = i < sg.size ? x : IO::pad();
        }
        K v[E];
        BB::from_striped(v, lds, t);
        BB::sort(v, lds, t);
        BB::to_striped(v, lds, t);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) dst[i] = IO::store(lds[BB::phys(i)]);
        }
        __syncthreads();
    }
}

// One wave per segment (size <= 64*E), four waves per workgroup.
template <class T, int E, class Pol = DefaultOrder<T>>
__global__ __launch_bounds__(256) void k_wave_sort(const TileSeg* segs, const uint32_t* count, T* a, const T* b) {
    using IO = KeyIO<T, Pol>;
    using K = typename Pol::key_t;
    const uint32_t cnt = *count;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6); s < cnt; s += gridDim.x * 4) {
        const TileSeg sg = segs[s];
        const T* src = (sg.buf ? b : a) + sg.begin;
        T* dst = a + sg.begin;
        K v[E];
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            const K x = IO::load(src[i < sg.size ? i : sg.size - 1]);
            v[r] = i < sg.size ? x : IO::pad();
        }
        WaveBitonic<Pol, E>::sort(v, lane);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            if (i < sg.size) dst[i] = IO::store(v[r]);
        }
    }
}

// Key/value tiles: sort (encoded key << 32 | position) as one 64-bit key, then
// gather the values by position.  Pads (all ones) sort after every real pair,
// including real pairs whose key is INT_MAX.
__device__ __forceinline__ uint64_t kv_pack(int32_t k, uint32_t pos) {
    return ((uint64_t)KeyTraits<int32_t>::enc(k) << 32) | pos;
}

template <int P, int E>
__global__ __launch_bounds__(P) void k_tile_sort_kv(const TileSeg* segs, const uint32_t* count, int32_t* ka,
                                                    int32_t* va, const int32_t* kb, const int32_t* vb) {
    using BB = BlockBitonic<UintOrder<uint64_t>, P, E>;
    __shared__ uint64_t lds[BB::LDS_N];
    const uint32_t cnt = *count;
    const uint32_t t = threadIdx.x;
    for (uint32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const TileSeg sg = segs[s];
        const int32_t* sk = (sg.buf ? kb : ka) + sg.begin;
        const int32_t* sv = (sg.buf ? vb : va) + sg.begin;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            const uint64_t x = kv_pack(sk[i < sg.size ? i : sg.size - 1], i);
            lds[BB::phys(i)] = i < sg.size ? x : ~0ull;
        }
        uint64_t v[E];
        BB::from_striped(v, lds, t);
        BB::sort(v, lds, t);
        BB::to_striped(v, lds, t);
        int32_t ok[E], ov[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) {
                const uint64_t c = lds[BB::phys(i)];
                ok[m] = KeyTraits<int32_t>::dec((uint32_t)(c >> 32));
                ov[m] = sv[(uint32_t)c];
            }
        }
        __syncthreads();  // every value gathered before any is overwritten
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) {
                ka[sg.begin + i] = ok[m];
                va[sg.begin + i] = ov[m];
            }
        }
        __syncthreads();
    }
}

// Key/value tiles with 64-bit keys: (key, position) does not fit one 64-bit
// network key, so the keys are sorted alone and each value is then placed at
// its key's rank -- the first sorted slot holding an equal key (binary search
// in LDS) plus a per-slot counter for duplicates.  Equal keys keep their
// values in no particular order, as in the reference's key/value sort.
template <class T, int P, int E>
__global__ __launch_bounds__(P) void k_tile_sort_kv64(const TileSeg* segs, const uint32_t* count, T* ka, int32_t* va,
                                                      const T* kb, const int32_t* vb) {
    using Tr = KeyTraits<T>;
    using BB = BlockBitonic<UintOrder<uint64_t>, P, E>;
    constexpr uint32_t N = BB::N;
    __shared__ uint64_t lds[BB::LDS_N];
    __shared__ uint32_t dup[N];
    __shared__ int32_t lval[N];
    const uint32_t cnt = *count;
    const uint32_t t = threadIdx.x;
    for (uint32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const TileSeg sg = segs[s];
        const T* sk = (sg.buf ? kb : ka) + sg.begin;
        const int32_t* sv = (sg.buf ? vb : va) + sg.begin;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            const uint64_t x = Tr::enc(sk[i < sg.size ? i : sg.size - 1]);
            lds[BB::phys(i)] = i < sg.size ? x : ~0ull;
            dup[i] = 0;
        }
        uint64_t v[E];
        BB::from_striped(v, lds, t);
        BB::sort(v, lds, t);
        BB::to_striped(v, lds, t);
        // pads (all ones) sort after every key below all ones, so a key's
        // lower bound plus its duplicate count stays below sg.size
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) {
                const uint64_t x = Tr::enc(sk[i]);
                uint32_t lo = 0, hi = N;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (lds[BB::phys(mid)] < x) lo = mid + 1;
                    else hi = mid;
                }
                lval[lo + atomicAdd(&dup[lo], 1u)] = sv[i];
            }
        }
        __syncthreads();  // every source key and value read before any is overwritten
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const uint32_t i = m * P + t;
            if (i < sg.size) {
                ka[sg.begin + i] = Tr::dec(lds[BB::phys(i)]);
                va[sg.begin + i] = lval[i];
            }
        }
        __syncthreads();
    }
}

template <int E>
__global__ __launch_bounds__(256) void k_wave_sort_kv(const TileSeg* segs, const uint32_t* count, int32_t* ka,
                                                      int32_t* va, const int32_t* kb, const int32_t* vb) {
    const uint32_t cnt = *count;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6); s < cnt; s += gridDim.x * 4) {
        const TileSeg sg = segs[s];
        const int32_t* sk = (sg.buf ? kb : ka) + sg.begin;
        const int32_t* sv = (sg.buf ? vb : va) + sg.begin;
        uint64_t v[E];
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            const uint64_t x = kv_pack(sk[i < sg.size ? i : sg.size - 1], i);
            v[r] = i < sg.size ? x : ~0ull;
        }
        WaveBitonic<UintOrder<uint64_t>, E>::sort(v, lane);
        int32_t ov[E];
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            if (i < sg.size) ov[r] = sv[(uint32_t)v[r]];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const uint32_t i = lane * E + r;
            if (i < sg.size) {
                ka[sg.begin + i] = KeyTraits<int32_t>::dec((uint32_t)(v[r] >> 32));
                va[sg.begin + i] = ov[r];
            }
        }
    }
}

// ---------------------------------------------------------------- sampling
// Kernels over segments or chunks take the number of work items from device
// memory and loop over them, so a fixed grid (hipGraph capture, no host
// round trip) and an exact grid (one item per workgroup) both work.
template <class T>
__global__ __launch_bounds__(256) void k_sample(const PartSeg* ps, const uint32_t* nseg, const T* a, const T* b,
                                                typename KeyTraits<T>::K* samples, uint32_t smax, uint64_t seed) {
    using Tr = KeyTraits<T>;
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        const PartSeg s = ps[w];
        if (s.mode) continue;  // learned segment: classified by the model's CDF table
        const T* src = (s.buf ? b : a) + s.begin;
        typename Tr::K* out = samples + s.sample_base;
        for (uint32_t i = threadIdx.x; i < s.nsample; i += blockDim.x) {
            const uint64_t h = mix64(seed ^ (s.begin * 0x632BE59BD9B4E019ull), i);
            const uint64_t pos = __umul64hi(h, s.size);  // uniform in [0, size)
            out[i] = Tr::enc(src[pos]);
        }
    }
}

// Splitters at evenly spaced ranks of the sorted sample, duplicates removed,
// padded with KMAX to F entries; then the two-level guide over [kmin, kmax]
// of the sample.  One workgroup of FMAX threads per segment.  Writes
// kmin/shift into the segment descriptor for the count and scatter kernels.
struct TreeLds {
    uint32_t wtot[FMAX / 64];
    uint32_t hbase[guide_hdrs(FMAX)];  // sub-table base | log2(cells) << 16
    uint32_t hj[guide_hdrs(FMAX)];     // j0 | j1 << 16 of the dense top cell
    uint32_t hsh[guide_hdrs(FMAX)];    // shift2
    uint64_t horg[guide_hdrs(FMAX)];   // origin (encoded key)
    uint32_t nuniq, nhdr;
};

__device__ __forceinline__ uint64_t sat_add_shl(uint64_t org, uint64_t q, uint32_t sh, uint64_t kmax) {
    // min(kmax, org + (q << sh)) without overflow
    if (q > ((kmax - org) >> sh)) return kmax;
    const uint64_t off = q << sh;
    return off > kmax - org ? kmax : org + off;
}

__device__ __forceinline__ uint32_t pow2ceil32(uint32_t v) { return v <= 1 ? 1u : 1u << (32 - __clz(v - 1)); }

template <class K>
__device__ __forceinline__ void build_tree_body(PartSeg& s, const K* smp, K* trees, uint32_t* guides, K* srt,
                                                TreeLds& L) {
    const uint32_t F = 1u << s.logf;
    const uint32_t ns = s.nsample;
    const uint32_t t = threadIdx.x;
    const uint64_t KM = (uint64_t)(K)~(K)0;
    K c = (K)~(K)0;
    bool keep = false;
    if (t + 1 < F) {
        const uint32_t r = (uint32_t)(((uint64_t)(t + 1) * ns) / F);
        c = smp[r > 0 ? r - 1 : 0];
        keep = true;
        if (t > 0) {
            const uint32_t rp = (uint32_t)(((uint64_t)t * ns) / F);
            keep = smp[rp > 0 ? rp - 1 : 0] != c;
        }
    }
    uint32_t nuniq;
    const uint32_t pos = block_excl_scan<FMAX>(keep ? 1u : 0u, L.wtot, &nuniq);
    if (t < F) srt[t] = (K)~(K)0;
    __syncthreads();
    if (keep) srt[pos] = c;
    __syncthreads();
    K* out = trees + s.tree_base;
    if (t < F) out[t] = srt[t];
    // top cells over the sample's key range
    const uint64_t kmin = (uint64_t)smp[0], kmax = (uint64_t)smp[ns - 1];
    const uint32_t TOP = guide_top(F);
    uint32_t shift = 0;
    while (((kmax - kmin) >> shift) >= TOP) ++shift;
    uint32_t* gd = guides + s.guide_base;
    const Guide g = make_guide(kmin, shift, F, 0);
    // each thread computes CPT consecutive top cells (TOP <= 4 * FMAX)
    constexpr uint32_t CPT_MAX = GUIDE_TOP_MAX / FMAX;
    static_assert(CPT_MAX * FMAX == GUIDE_TOP_MAX, "top cells per thread");
    const uint32_t CPT = TOP > FMAX ? TOP / FMAX : 1u;
    uint32_t cj0[CPT_MAX], cj1[CPT_MAX], csc[CPT_MAX];
    uint32_t my_dense = 0, my_sc = 0;
#pragma unroll
    for (uint32_t k = 0; k < CPT_MAX; ++k) {
        const uint32_t c = t * CPT + k;
        cj0[k] = cj1[k] = csc[k] = 0;
        if (k < CPT && c < TOP) {
            const uint64_t lo_key = sat_add_shl(kmin, c, shift, KM);
            const uint64_t hi_key = sat_add_shl(kmin, c + 1, shift, KM);
            cj0[k] = c == 0 ? 0u : lower_bound_lds(srt, nuniq, (K)lo_key);
            cj1[k] = c == TOP - 1 ? nuniq : lower_bound_lds(srt, nuniq, (K)hi_key);
            if (cj1[k] - cj0[k] > GUIDE_DENSE) {
                csc[k] = pow2ceil32(4 * (cj1[k] - cj0[k]));
                my_dense += 1;
                my_sc += csc[k];
            }
        }
    }
    uint32_t nhdr, nsub;
    uint32_t h = block_excl_scan<FMAX>(my_dense, L.wtot, &nhdr);
    uint32_t sbase = block_excl_scan<FMAX>(my_sc, L.wtot, &nsub);
#pragma unroll
    for (uint32_t k = 0; k < CPT_MAX; ++k) {
        const uint32_t c = t * CPT + k;
        if (!(k < CPT && c < TOP)) continue;
        const uint32_t j0 = cj0[k], j1 = cj1[k], sc = csc[k];
        if (sc) {
            const uint64_t org = (uint64_t)srt[j0], last = (uint64_t)srt[j1 - 1];
            uint32_t sh2 = 0;
            while (((last - org) >> sh2) >= sc) ++sh2;
            const uint32_t lg = 31 - __clz(sc);
            L.hbase[h] = sbase | (lg << 16);
            L.hj[h] = j0 | (j1 << 16);
            L.hsh[h] = sh2;
            L.horg[h] = org;
            gd[c] = GUIDE_SUBF | h;
            gd[g.hdr + 4 * h + 0] = (uint32_t)org;
            gd[g.hdr + 4 * h + 1] = (uint32_t)(org >> 32);
            gd[g.hdr + 4 * h + 2] = sh2;
            gd[g.hdr + 4 * h + 3] = sbase | (lg << 16);
            h += 1;
            sbase += sc;
        } else {
            gd[c] = j0 | (j1 << 16);
        }
    }
    // unused headers are zero: classify_batch reads header 0 for every key
    for (uint32_t q = nhdr + t; q < guide_hdrs(F); q += FMAX) {
#pragma unroll
        for (int k = 0; k < 4; ++k) gd[g.hdr + 4 * q + k] = 0u;
    }
    __syncthreads();
    // sub-table cells: q -> its header by binary search over the bases
    for (uint32_t q = t; q < nsub; q += FMAX) {
        uint32_t lo = 0, hi = nhdr;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((L.hbase[mid] & 0xFFFFu) <= q) lo = mid;
            else hi = mid;
        }
        const uint32_t hb = L.hbase[lo], hjj = L.hj[lo], sh2 = L.hsh[lo];
        const uint64_t org = L.horg[lo];
        const uint32_t loc = q - (hb & 0xFFFFu), cells = 1u << (hb >> 16);
        const uint32_t a0 = hjj & 0xFFFFu, a1 = hjj >> 16;
        const uint32_t jl = loc == 0 ? a0 : a0 + lower_bound_lds(srt + a0, a1 - a0, (K)sat_add_shl(org, loc, sh2, KM));
        const uint32_t jh =
            loc == cells - 1 ? a1 : a0 + lower_bound_lds(srt + a0, a1 - a0, (K)sat_add_shl(org, loc + 1, sh2, KM));
        gd[g.sub + q] = jl | (jh << 16);
    }
    if (t == 0) {
        s.kmin = kmin;
        s.shift = shift;
        s.flags = (nhdr ? GUIDE_SUBTABLES : 0u) | (nuniq + 1 < F ? GUIDE_EQBUCKETS : 0u);
    }
}

template <class K>
__global__ __launch_bounds__(FMAX) void k_build_tree(PartSeg* ps, const uint32_t* nseg, const K* samples,
                                                     uint32_t smax, K* trees, uint32_t* guides) {
    __shared__ K srt[FMAX];
    __shared__ TreeLds L;
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        PartSeg& s = ps[w];
        if (s.nsample == 0) continue;
        build_tree_body<K>(s, samples + s.sample_base, trees, guides, srt, L);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- count
// Encoded fp64 keys that are NaN or -0.0 (they rule out the min/max leaf sort).
__device__ __forceinline__ bool nan_or_negzero(uint64_t k) {
    return k == 0x7FFFFFFFFFFFFFFFull || k > 0xFFF0000000000000ull || k < 0x000FFFFFFFFFFFFFull;
}
__device__ __forceinline__ bool nan_or_negzero(uint32_t) { return false; }
// The same test on the raw bits of a key of type T: for doubles one
// v_cmp_class_f64 with the classes signalling NaN, quiet NaN and negative zero
// (three 64-bit compares on the encoded key otherwise); other types never.
template <class T, class K>
__device__ __forceinline__ bool nan_or_negzero_raw(K raw) {
    if constexpr (std::is_same<T, double>::value) return __builtin_amdgcn_class(__builtin_bit_cast(double, raw), 0x23);
    else return false;
}

// Histogram and rank helpers over the E keys a thread holds (all lanes of the
// wave active).  Every lane adds one to its key's bucket.  At a key position
// where the whole wave holds one bucket (sorted or run-heavy input) the first
// lane adds the wave's count and the other lanes add zero to a private dummy
// slot, so 64 lanes do not serialise on one LDS word.  No branch depends on
// the keys, and all atomics are issued before any return value is used.
constexpr int DUMMY_SLOTS = 64;

template <int E>
__device__ __forceinline__ void wave_hist_add_many(uint32_t* cnt, uint32_t* dummy, const uint32_t (&bk)[E]) {
    const uint32_t lane = __lane_id();
    const bool first = lane == 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(bk[e]);
        const bool same = __ballot(bk[e] == b0) == ~0ull;
        uint32_t* addr = same ? (first ? &cnt[b0] : &dummy[lane]) : &cnt[bk[e]];
        atomicAdd(addr, same ? (first ? 64u : 0u) : 1u);
    }
}

template <int E>
__device__ __forceinline__ void wave_rank_many(uint32_t* cnt, uint32_t* dummy, const uint32_t (&bk)[E],
                                               uint32_t (&lr)[E]) {
    const uint32_t lane = __lane_id();
    const bool first = lane == 0;
    bool same[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(bk[e]);
        same[e] = __ballot(bk[e] == b0) == ~0ull;
        uint32_t* addr = same[e] ? (first ? &cnt[b0] : &dummy[lane]) : &cnt[bk[e]];
        lr[e] = atomicAdd(addr, same[e] ? (first ? 64u : 0u) : 1u);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) lr[e] = same[e] ? __builtin_amdgcn_readfirstlane(lr[e]) + lane : lr[e];
}

// Per chunk: bucket histogram -> hist[chunk][bucket].  With ids != nullptr
// the bucket of every key is also stored (ids[position] = bucket), so the
// scatter does not classify again.  Batches of CNT_P*CNT_U keys alternate
// between two register buffers: the next batch is in flight while this one is
// classified, with no register copies that would wait for it.
template <class T, int CNT_P, int CNT_U, int NB, bool LRN>
__global__ __launch_bounds__(CNT_P) void k_count(const PartSeg* ps, const ChunkRef* cr, const uint32_t* nchunk,
                                                 const T* a, const T* b, const typename KeyTraits<T>::K* trees,
                                                 const uint32_t* guides, uint32_t* hist, uint32_t* special,
                                                 uint16_t* ids) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    constexpr int STEP = CNT_P * CNT_U;
    __shared__ K lsrt[FMAX];
    __shared__ __align__(16) uint32_t lguide[LDS_TABLE_WORDS];
    __shared__ uint32_t lh[BMAX];
    __shared__ uint32_t dummy[DUMMY_SLOTS];
    for (uint32_t w = blockIdx.x; w < *nchunk; w += gridDim.x) {
    const ChunkRef c = cr[w];
    const PartSeg s = ps[c.seg];
    const uint32_t F = 1u << s.logf, B = 2 * F;
    const uint32_t t = threadIdx.x;
    const Guide g = load_classifier<K, CNT_P, LRN>(s, trees, guides, lsrt, lguide);
    for (uint32_t j = t; j < B; j += CNT_P) lh[j] = 0;
    __syncthreads();
    const T* src = (s.buf ? b : a) + s.begin;
    uint16_t* oid = ids ? ids + s.begin : nullptr;
    const uint64_t lo = (uint64_t)c.idx * s.chunk;
    const uint64_t hi = lo + s.chunk < s.size ? lo + s.chunk : s.size;
    bool odd = false;
    // clamped indices keep every load in bounds and unconditional; the buffers
    // hold the raw bits, encoded only where a batch is classified, so that no
    // instruction right after the prefetch has to wait for it.  (Non-temporal
    // loads measured 5% faster here but slowed the scatter that follows by
    // more: ordinary loads.)
    auto load = [&](uint64_t base, K (&d)[CNT_U]) {
#pragma unroll
        for (int u = 0; u < CNT_U; ++u) {
            const uint64_t i = base + (uint64_t)u * CNT_P + t;
            d[u] = Tr::bits(src[i < hi ? i : hi - 1]);
        }
    };
    auto batch = [&](uint64_t base, K (&xr)[CNT_U], K (&xn)[CNT_U]) {
        // unconditional (clamped, in bounds past the end): a conditional
        // prefetch makes the compiler wait for it before this batch's keys
        load(base + STEP, xn);
        K x[CNT_U];
#pragma unroll
        for (int u = 0; u < CNT_U; ++u) x[u] = Tr::enc_bits(xr[u]);
        if (base + STEP <= hi) {
            uint32_t bk[CNT_U];
#pragma unroll
            for (int u0 = 0; u0 < CNT_U; u0 += NB) {
                K xb[NB];
                uint32_t kb[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) xb[u] = x[u0 + u];
                classify_keys<K, NB, LRN>(s, lsrt, lguide, g, xb, kb);
#pragma unroll
                for (int u = 0; u < NB; ++u) bk[u0 + u] = kb[u];
            }
            wave_hist_add_many<CNT_U>(lh, dummy, bk);
            if (special) {
#pragma unroll
                for (int u = 0; u < CNT_U; ++u) odd |= nan_or_negzero_raw<T>(xr[u]);
            }
            if (oid) {
#pragma unroll
                for (int u = 0; u < CNT_U; ++u) oid[base + (uint64_t)u * CNT_P + t] = (uint16_t)bk[u];
            }
        } else {
#pragma unroll
            for (int u = 0; u < CNT_U; ++u) {
                const uint64_t i = base + (uint64_t)u * CNT_P + t;
                if (i < hi) {
                    const uint32_t bk = classify_key<LRN>(s, lsrt, lguide, g, x[u]);
                    if (special) odd |= nan_or_negzero_raw<T>(xr[u]);
                    atomicAdd(&lh[bk], 1u);
                    if (oid) oid[i] = (uint16_t)bk;
                }
            }
        }
    };
    K xa[CNT_U], xb[CNT_U];
    if (lo < hi) load(lo, xa);
    for (uint64_t base = lo; base < hi;) {
        batch(base, xa, xb);
        base += STEP;
        if (base >= hi) break;
        batch(base, xb, xa);
        base += STEP;
    }
    if (special && __any(odd) && (t & 63u) == 0) atomicOr(special, 1u);
    __syncthreads();
    uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
    for (uint32_t j = t; j < B; j += CNT_P) h[j] = lh[j];
    __syncthreads();
    }
}

// Key types the learned model covers: only their kernels get a learned build.
template <class T>
constexpr bool learnable_key() { return std::is_same<T, double>::value || std::is_same<T, int32_t>::value; }

// Count variants: workgroup threads, keys per thread per batch, and keys
// classified together; grid = chunks (the kernel loops over the device count).
// lrn: every segment of the level is learned.
template <class T, bool LRN>
inline void launch_count_as(int shape, uint32_t grid, hipStream_t st, const PartSeg* ps, const ChunkRef* cr,
                            const uint32_t* nchunk, const T* a, const T* b, const typename KeyTraits<T>::K* trees,
                            const uint32_t* guides, uint32_t* hist, uint32_t* special, uint16_t* ids) {
#define S512_COUNT(PP, UU, NN) \
    k_count<T, PP, UU, NN, LRN><<<grid, PP, 0, st>>>(ps, cr, nchunk, a, b, trees, guides, hist, special, ids)
    switch (shape) {
        case 1: S512_COUNT(256, 16, 4); break;
        case 2: S512_COUNT(256, 8, 8); break;
        case 3: S512_COUNT(512, 8, 8); break;
        case 4: S512_COUNT(256, 8, 4); break;
        default: S512_COUNT(256, 16, 8); break;
    }
#undef S512_COUNT
}
template <class T>
inline void launch_count(int shape, uint32_t grid, hipStream_t st, const PartSeg* ps, const ChunkRef* cr,
                         const uint32_t* nchunk, const T* a, const T* b, const typename KeyTraits<T>::K* trees,
                         const uint32_t* guides, uint32_t* hist, uint32_t* special, uint16_t* ids, bool lrn = false) {
    if constexpr (learnable_key<T>()) {
        if (lrn) {
            launch_count_as<T, true>(shape, grid, st, ps, cr, nchunk, a, b, trees, guides, hist, special, ids);
            return;
        }
    }
    launch_count_as<T, false>(shape, grid, st, ps, cr, nchunk, a, b, trees, guides, hist, special, ids);
}

// ---------------------------------------------------------------- column scan
// hist[chunk][bucket] of every segment -> prefix over earlier chunks of the
// same bucket (in place); bsize[b] = bucket size.  One workgroup per
// (segment, slice of 64 buckets): lane = bucket, wave = 1/NG of the chunks;
// loads are batched so a column of thousands of chunks is not a chain of
// dependent round trips.  The host launches NG = 16 when a segment has many
// chunks (the root level: thousands of chunks in one segment, only 16
// workgroups in all) and NG = 4 otherwise.
constexpr int CS_SLICE = 64;
constexpr int CS_NSLICE = BMAX / CS_SLICE;
// runs (may be null): a sample of the runs the scatter will write, for the
// full-line scatter's gate: the first slice of buckets of each of the first
// RUN_SAMPLE_SEGS segments adds its nonzero (chunk, bucket) counts to runs[0]
// and the (chunk, bucket) pairs it looked at to runs[1] (one pair of atomics
// per workgroup).
constexpr uint32_t RUN_SAMPLE_SEGS = 64;
template <int NG>
__global__ __launch_bounds__(NG * 64) void k_colscan(const PartSeg* ps, const uint32_t* nseg, uint32_t* hist,
                                                     uint64_t* bsize, uint32_t* runs_out = nullptr) {
    constexpr int CS_BATCH = NG >= 16 ? 16 : 8;  // loads in flight per lane
    __shared__ uint32_t part[NG][CS_SLICE];
    __shared__ uint32_t rsum[NG][1];
    const uint32_t lane = threadIdx.x & 63u, grp = threadIdx.x >> 6;
    const uint32_t nwork = *nseg * CS_NSLICE;
    for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const PartSeg s = ps[w / CS_NSLICE];
        const uint32_t B = 2u << s.logf;
        const uint32_t b0 = (w % CS_NSLICE) * CS_SLICE;
        if (b0 >= B) continue;  // whole workgroup: no barrier is skipped by part of it
        const uint32_t bkt = b0 + lane;
        const bool on = bkt < B;
        const uint32_t nc = s.nchunks;
        const uint32_t c0 = (uint32_t)((uint64_t)nc * grp / NG), c1 = (uint32_t)((uint64_t)nc * (grp + 1) / NG);
        uint32_t* h = hist + s.hist_base + bkt;
        uint32_t sum = 0, runs = 0;
        for (uint32_t cb = c0; cb < c1; cb += CS_BATCH) {
            uint32_t v[CS_BATCH];
#pragma unroll
            for (int k = 0; k < CS_BATCH; ++k) v[k] = (on && cb + k < c1) ? h[(uint64_t)(cb + k) * B] : 0u;
#pragma unroll
            for (int k = 0; k < CS_BATCH; ++k) {
                sum += v[k];
                runs += v[k] ? 1u : 0u;
            }
        }
        const bool sample = runs_out && w % CS_NSLICE == 0 && w / CS_NSLICE < RUN_SAMPLE_SEGS;
        if (sample) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) runs += __shfl_xor(runs, d);
            if (lane == 0) rsum[grp][0] = runs;
        }
        part[grp][lane] = sum;
        __syncthreads();
        if (sample && threadIdx.x == 0) {
            uint32_t r = 0;
#pragma unroll
            for (int q = 0; q < NG; ++q) r += rsum[q][0];
            atomicAdd(&runs_out[0], r);
            atomicAdd(&runs_out[1], min(B - b0, (uint32_t)CS_SLICE) * nc);
        }
        uint32_t run = 0, total = 0;
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const uint32_t p = part[q][lane];
            run += q < (int)grp ? p : 0u;
            total += p;
        }
        for (uint32_t cb = c0; cb < c1; cb += CS_BATCH) {
            uint32_t v[CS_BATCH];
#pragma unroll
            for (int k = 0; k < CS_BATCH; ++k) v[k] = (on && cb + k < c1) ? h[(uint64_t)(cb + k) * B] : 0u;
#pragma unroll
            for (int k = 0; k < CS_BATCH; ++k) {
                if (on && cb + k < c1) h[(uint64_t)(cb + k) * B] = run;
                run += v[k];
            }
        }
        if (on && grp == 0) bsize[s.bkt_base + bkt] = total;
        __syncthreads();
    }
}

// bstart[b] = first position of bucket b inside its segment.
[[maybe_unused]] static __global__ __launch_bounds__(BMAX) void k_bucket_scan(const PartSeg* ps, const uint32_t* nseg,
                                                                             const uint64_t* bsize, uint64_t* bstart) {
    __shared__ uint32_t wtot[BMAX / 64];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        const PartSeg s = ps[w];
        const uint32_t B = 2u << s.logf;
        const uint32_t t = threadIdx.x;
        const uint32_t v = t < B ? (uint32_t)bsize[s.bkt_base + t] : 0u;
        uint32_t total;
        const uint32_t start = block_excl_scan<BMAX>(v, wtot, &total);
        if (t < B) bstart[s.bkt_base + t] = start;
    }
}

// ---------------------------------------------------------------- scatter
// Keys (and values for key/value sorts) move from array buf to array 1-buf.
// A workgroup of P threads walks its chunk in rounds of P*E keys:
//   bucket of each key (classified here, or read from the ids the count
//   pass stored), rank inside the round with LDS atomics     | barrier 1
//   wave-level scan of the round histogram (BMAX/P buckets per thread)
//                                                            | barrier 2
//   bucket owners: round start rs, destination base, advance the chunk's
//   cursor, clear the counter for the next round             | barrier 3
//   stage keys grouped by bucket in LDS                      | barrier 4
//   write the staged keys: consecutive threads, consecutive addresses
// The next round's keys are loaded into registers while this round runs.
// Write combining (WC): a bucket's run usually ends inside a 128-byte line
// whose rest the next round writes; the keys of that unfinished line wait in
// an LDS list and are written during the next round's staging phase, a few
// microseconds before the rest of their line, instead of a whole round
// earlier (the L2 then merges the two halves rather than evicting half a line).
constexpr int WC_CAP = 4096;  // deferred tail-line keys per round (write combining)

template <class T, bool KV, bool IDS, int P, int E, bool WC>
struct ScatterLds {
    using K = typename KeyTraits<T>::K;
    static constexpr int ROUND = P * E;
    static constexpr bool CLS = !IDS;  // classifier tables needed
    K lsrt[CLS ? FMAX : 1];
    __align__(16) uint32_t lguide[CLS ? LDS_TABLE_WORDS : 1];
    uint32_t cursor[BMAX];  // next destination of each bucket (segment-relative)
    uint32_t rh[BMAX];      // round histogram
    uint16_t rs[BMAX];      // round start of each bucket in the staging area
    uint32_t dbase[BMAX];   // destination of staged slot i is dbase[bucket] + i
    uint32_t wtot[P / 64];
    uint32_t dummy[DUMMY_SLOTS];
    T stg[ROUND];
    uint16_t stb[ROUND];
    int32_t stv[KV ? ROUND : 1];
    T dkey[WC ? WC_CAP : 1];              // deferred keys (write combining)
    int32_t dval[WC && KV ? WC_CAP : 1];
    uint32_t dpos[WC ? WC_CAP : 1];
    uint32_t dcount;
};

// Choice between the full-line scatter and the plain one, made on the device
// from how many (chunk, bucket) pairs of the level hold keys (sampled by
// k_colscan).  When few do (sorted or reversed input: a chunk's keys fall in a
// few neighbouring buckets, in long runs), the plain scatter writes whole lines
// too and measured faster (2^28 fp64 keys, sorted: 1.21 + 0.96 ms for both
// levels against 1.10 + 1.69); when a quarter or more do, the full-line
// scatter (uniform, where half the pairs, the ordinary buckets, hold keys:
// 1.25 + 1.31 against 1.66 + 1.59; 2^30 mixture keys, runs of 512 keys: 30.1
// against 32.8 ms for the sort).  The host launches both kernels of a pair;
// one of them returns at once.
struct ScatterGate {
    const uint32_t* runs = nullptr;  // sampled nonzero pairs, pairs looked at (k_colscan)
    uint32_t side = 0;               // 0: no gate; 1: runs on dense pairs; 2: runs on sparse pairs
    template <class T>
    __device__ __forceinline__ bool skip() const {
        if (!side) return false;
        const bool sparse = (uint64_t)runs[0] * 4 < (uint64_t)runs[1];
        return side == 1 ? sparse : !sparse;
    }
};

template <class T, bool KV, bool IDS, int P, int E, bool WC, bool LRN = false>
__global__ __launch_bounds__(P) void k_scatter(const PartSeg* ps, const ChunkRef* cr, T* a, T* b, int32_t* va,
                                               int32_t* vb, const typename KeyTraits<T>::K* trees,
                                               const uint32_t* guides, const uint32_t* hist, const uint64_t* bstart,
                                               const uint32_t* nchunk, const uint16_t* ids,
                                               const ScatterGate gate = ScatterGate{}) {
    if (gate.template skip<T>()) return;  // the whole workgroup, before any barrier
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    using L = ScatterLds<T, KV, IDS, P, E, WC>;
    constexpr uint64_t LINE = 128 / sizeof(T);  // keys per 128-byte line
    constexpr int NW = P / 64;
    constexpr int ROUND = L::ROUND;
    constexpr int BPT = BMAX / P;  // buckets per thread in the scan
    static_assert(BPT * P == BMAX && ROUND <= 32768, "scatter shape");
    __shared__ L sm;

    for (uint32_t wi = blockIdx.x; wi < *nchunk; wi += gridDim.x) {
    uint32_t wc = wi;
    if (gridDim.x >= *nchunk) {
        // workgroups reach the 8 XCDs round-robin: give XCD x a contiguous run of
        // chunks, so the line two neighbouring chunks share in a bucket is
        // written through one L2
        const uint32_t N = *nchunk, q = N / 8, r = N % 8, x = wi % 8;
        wc = x * q + (x < r ? x : r) + wi / 8;
    }
    const ChunkRef c = cr[wc];
    const PartSeg s = ps[c.seg];
    const uint32_t B = 2u << s.logf;
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    Guide g{0, 0, 0, 0, 0, 0};
    if constexpr (!IDS) g = load_classifier<K, P, LRN>(s, trees, guides, sm.lsrt, sm.lguide);
    const uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
    for (uint32_t j = t; j < BMAX; j += P) {
        sm.cursor[j] = j < B ? (uint32_t)bstart[s.bkt_base + j] + h[j] : 0u;
        sm.rh[j] = 0;
    }
    if (WC && t == 0) sm.dcount = 0;
    const T* src = (s.buf ? b : a) + s.begin;
    T* dst = (s.buf ? a : b) + s.begin;
    const int32_t* vsrc = KV ? ((s.buf ? vb : va) + s.begin) : nullptr;
    int32_t* vdst = KV ? ((s.buf ? va : vb) + s.begin) : nullptr;
    const uint16_t* isrc = IDS ? ids + s.begin : nullptr;
    auto flush_deferred = [&](uint32_t nd) {
        for (uint32_t k = t; k < nd; k += P) {
            dst[sm.dpos[k]] = sm.dkey[k];
            if constexpr (KV) vdst[sm.dpos[k]] = sm.dval[k];
        }
    };
    const uint64_t lo = (uint64_t)c.idx * s.chunk;
    const uint64_t hi = lo + s.chunk < s.size ? lo + s.chunk : s.size;

    // rounds alternate between two register buffers: the next round's keys
    // are in flight while this one runs, with no copies that would wait for them
    struct Buf {
        T x[E];
        int32_t v[E];
        uint32_t b[E];
    };
    // a round at or past the end (the prefetch after the last round) reads
    // key hi - 1 everywhere: in bounds, unconditional
    auto load_round = [&](uint64_t base, Buf& R) {
        const uint64_t bb = base < hi ? base : hi - 1;
        const uint32_t cnt = (uint32_t)((hi - bb) < (uint64_t)ROUND ? (hi - bb) : (uint64_t)ROUND);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;
            const uint64_t ic = bb + (i < cnt ? i : cnt - 1);
            R.x[e] = src[ic];
            if constexpr (KV) R.v[e] = vsrc[ic];
            if constexpr (IDS) R.b[e] = isrc[ic];
        }
    };
    auto do_round = [&](uint64_t base, Buf& cur, Buf& nxt) {
        const uint32_t cnt = (uint32_t)((hi - base) < (uint64_t)ROUND ? (hi - base) : (uint64_t)ROUND);
        load_round(base + ROUND, nxt);  // prefetch, unconditional so that no wait covers it
        uint32_t bk[E], lr[E];
        if constexpr (IDS) {
#pragma unroll
            for (int e = 0; e < E; ++e) bk[e] = cur.b[e];
        } else if (cnt == ROUND) {
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += CLS_NB) {
                K kx[CLS_NB];
                uint32_t kb[CLS_NB];
#pragma unroll
                for (int u = 0; u < CLS_NB; ++u) kx[u] = Tr::enc(cur.x[e0 + u]);
                classify_keys<K, CLS_NB, LRN>(s, sm.lsrt, sm.lguide, g, kx, kb);
#pragma unroll
                for (int u = 0; u < CLS_NB; ++u) bk[e0 + u] = kb[u];
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) bk[e] = classify_key<LRN>(s, sm.lsrt, sm.lguide, g, Tr::enc(cur.x[e]));
        }
        if (cnt == ROUND) {
            wave_rank_many<E>(sm.rh, sm.dummy, bk, lr);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (e * P + t < cnt) lr[e] = atomicAdd(&sm.rh[bk[e]], 1u);
        }
        __syncthreads();  // 1
        // exclusive scan of rh: thread t owns buckets t*BPT .. t*BPT+BPT-1
        uint32_t hv[BPT], v = 0;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            hv[k] = sm.rh[t * BPT + k];
            v += hv[k];
        }
        uint32_t incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += y;
        }
        if (lane == 63) sm.wtot[wid] = incl;
        __syncthreads();  // 2
        uint32_t woff = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) woff += (w < (int)wid) ? sm.wtot[w] : 0u;
        const uint32_t nd = WC ? sm.dcount : 0u;  // deferred by the previous round
        uint32_t p = woff + incl - v;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const uint32_t j = t * BPT + k;
            sm.rs[j] = (uint16_t)p;
            sm.dbase[j] = sm.cursor[j] - p;
            sm.cursor[j] += hv[k];
            sm.rh[j] = 0;
            p += hv[k];
        }
        __syncthreads();  // 3
        if constexpr (WC) {
            flush_deferred(nd < WC_CAP ? nd : WC_CAP);
            if (t == 0) sm.dcount = 0;  // every thread read it before barrier 3
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;
            if (i < cnt) {
                const uint32_t q = sm.rs[bk[e]] + lr[e];
                sm.stg[q] = cur.x[e];
                sm.stb[q] = (uint16_t)bk[e];
                if constexpr (KV) sm.stv[q] = cur.v[e];
            }
        }
        __syncthreads();  // 4
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;  // staged keys are grouped by bucket
            const uint32_t bkt = sm.stb[i] & (BMAX - 1u);  // slots >= cnt: stale, never stored
            const uint64_t d = (uint64_t)sm.dbase[bkt] + i;
            const T y = sm.stg[i];
            int32_t yv = 0;
            if constexpr (KV) yv = sm.stv[i];
            if (i < cnt) {
                bool direct = true;
                if constexpr (WC) {
                    // keys in the unfinished last line of this bucket's run wait
                    const uint64_t end = s.begin + sm.cursor[bkt];  // absolute end after this round
                    if ((end & (LINE - 1)) && s.begin + d >= (end & ~(LINE - 1))) {
                        const uint32_t k = atomicAdd(&sm.dcount, 1u);
                        if (k < WC_CAP) {
                            sm.dkey[k] = y;
                            sm.dpos[k] = (uint32_t)d;
                            if constexpr (KV) sm.dval[k] = yv;
                            direct = false;
                        }
                    }
                }
                if (direct) {
                    dst[d] = y;
                    if constexpr (KV) vdst[d] = yv;
                }
            }
        }
        // the next round's atomics touch rh (cleared before barrier 3) and the
        // staging area is rewritten only after its barrier 3, which every
        // thread reaches after finishing the stores above
    };
    Buf A, Bb;
    if (lo < hi) load_round(lo, A);
    __syncthreads();
    for (uint64_t base = lo; base < hi;) {
        do_round(base, A, Bb);
        base += ROUND;
        if (base >= hi) break;
        do_round(base, Bb, A);
        base += ROUND;
    }
    if constexpr (WC) {
        __syncthreads();
        const uint32_t nd = sm.dcount;
        flush_deferred(nd < WC_CAP ? nd : WC_CAP);
    }
    __syncthreads();  // LDS is reloaded for the next chunk
    }
}

// ---------------------------------------------------------------- full-line scatter
// Keys only, reading the bucket ids that the count pass stored.  Measured on
// MI355X (2^28 fp64 keys written as 512 streams per workgroup): a 128-byte
// line stored whole by one store instruction streams at 5.7 TB/s; the same
// bytes as runs of 16 keys off line alignment -- every line stored in two
// parts, one per round, the pattern of k_scatter -- at 1.9 TB/s, at 3.5 TB/s
// when the parts are 64-byte halves, and neither cache hints nor storing the
// two parts close together in time changes that.
//
// So every store here covers whole lines.  A bucket's keys that do not
// complete their line in a round wait in LDS (`carry`, at most LINE - 1 keys
// per ordinary bucket) and go out at the front of the bucket's next run.  The
// store phase works on line jobs: LINE consecutive lanes write one line,
// reading each key from the carry or from the staging area.
//
// A workgroup takes a contiguous range of chunks (a persistent grid of about
// one workgroup per CU).  The chunks of one segment inside that range are one
// stream: bucket j's destination continues from chunk to chunk, so only the
// stream's first line of each bucket (shared with the previous workgroup's
// range) and its last line are stored partially.  Equality buckets and the
// stream's last round store everything.
//
// Per round (the next round's keys and ids are loaded meanwhile):
//   rank every key inside the round with LDS atomics          | barrier 1
//   bucket owners: the last round's waiting keys into the carry; this round's
//   writes (carried keys first, then staged ones, up to the last line
//   boundary), staging start and line jobs; a packed block scan of
//   (keys, jobs)                                              | barrier 2, 3
//   stage the keys grouped by bucket                          | barrier 4
//   line jobs: LINE lanes per line, one store instruction per 64 / LINE lines
constexpr int SCATTER_FL = 5;  // scatter_shape of the full-line kernel

template <class T, int P, int E>
struct ScatterFlLds {
    static constexpr int ROUND = P * E;
    static constexpr int LINE = 128 / (int)sizeof(T);  // keys per 128-byte line
    static constexpr int CW = LINE - 1;                // waiting keys per ordinary bucket, at most
    // line jobs of a round: every staged and carried key in full lines plus a
    // partial first and last line per bucket
    static constexpr int JMAX = (ROUND + FMAX * CW) / LINE + 2 * BMAX;
    struct alignas(16) Rec {
        uint32_t d;   // first place written this round (segment-relative)
        uint16_t w;   // keys written: the k carried ones, then staged ones
        uint16_t rs;  // the bucket's staged keys: stg[rs ...]
        uint16_t js;  // its first line job
        uint16_t ts;  // waiting keys after this round: stg[ts .. ts + tn) ...
        uint8_t k;    // carried keys written first (ordinary buckets)
        uint8_t a;    // line offset of d (nonzero only on a stream's first line)
        uint8_t tn;
        uint8_t cd;   // ... go to carry slots cd .. cd + tn
    };
    static_assert(sizeof(Rec) == 16, "record layout");
    Rec rec[BMAX];
    uint32_t cursor[BMAX];  // next place of each bucket (segment-relative, carried keys included)
    uint32_t rh[BMAX];      // round histogram
    uint16_t jb[JMAX];      // bucket of each line job
    uint32_t wtot[P / 64];
    uint32_t dummy[DUMMY_SLOTS];
    T stg[ROUND];
    T carry[FMAX * CW];
};

template <class T, int P, int E>
__global__ __launch_bounds__(P) void k_scatter_fl(const PartSeg* ps, const ChunkRef* cr, T* a, T* b,
                                                  const uint32_t* hist, const uint64_t* bstart,
                                                  const uint32_t* nchunk, const uint16_t* ids, uint32_t per_wg,
                                                  const ScatterGate gate = ScatterGate{}) {
    if (gate.template skip<T>()) return;  // the whole workgroup, before any barrier
    using L = ScatterFlLds<T, P, E>;
    using Rec = typename L::Rec;
    constexpr int ROUND = L::ROUND, LINE = L::LINE, CW = L::CW;
    constexpr int NW = P / 64;
    constexpr int JPI = 64 / LINE;  // line jobs per wave store instruction
    static_assert(BMAX == 2 * P, "one ordinary and one equality bucket per thread");
    static_assert(ROUND < 65536 && (LINE & (LINE - 1)) == 0 && LINE <= 64, "full-line scatter shape");
    __shared__ L sm;
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    const uint32_t N = *nchunk;
    const uint32_t cend = min(N, (blockIdx.x + 1) * per_wg);
    for (uint32_t ci = blockIdx.x * per_wg; ci < cend;) {
        const ChunkRef c = cr[ci];
        const PartSeg s = ps[c.seg];
        // this segment's chunks c.idx, c.idx + 1, .. inside the range: one stream
        const uint32_t nst = min(cend - ci, s.nchunks - c.idx);
        ci += nst;
        const uint32_t B = 2u << s.logf;
        const uint64_t lo = (uint64_t)c.idx * s.chunk;
        const uint64_t hi = min(lo + (uint64_t)nst * s.chunk, s.size);
        const T* src = (s.buf ? b : a) + s.begin;
        T* dst = (s.buf ? a : b) + s.begin;
        const uint16_t* isrc = ids + s.begin;
        // stores through an explicit global pointer: a flat store would also
        // count against the LDS wait counter and stall the next LDS read
        __attribute__((address_space(1))) T* gdst = (__attribute__((address_space(1))) T*)dst;
        const uint32_t dph = (uint32_t)((uintptr_t)dst / sizeof(T));  // line arithmetic on the address
        const uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
        for (uint32_t j = t; j < BMAX; j += P) {
            sm.cursor[j] = j < B ? (uint32_t)bstart[s.bkt_base + j] + h[j] : 0u;
            sm.rh[j] = 0;
            sm.rec[j] = Rec{};  // nothing carried or waiting
        }
        struct Buf {
            T x[E];
            uint32_t b[E];
        };
        // the prefetch after the last round reads key hi - 1 everywhere: in bounds
        auto load_round = [&](uint64_t base, Buf& R) {
            const uint64_t bb = base < hi ? base : hi - 1;
            const uint32_t cnt = (uint32_t)((hi - bb) < (uint64_t)ROUND ? (hi - bb) : (uint64_t)ROUND);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t i = e * P + t;
                const uint64_t ic = bb + (i < cnt ? i : cnt - 1);
                R.x[e] = src[ic];
                R.b[e] = isrc[ic];
            }
        };
        auto do_round = [&](uint64_t base, Buf& cur, Buf& nxt) {
            const uint32_t cnt = (uint32_t)((hi - base) < (uint64_t)ROUND ? (hi - base) : (uint64_t)ROUND);
            const bool last = base + ROUND >= hi;
            load_round(base + ROUND, nxt);
            uint32_t bk[E], lr[E];
#pragma unroll
            for (int e = 0; e < E; ++e) bk[e] = cur.b[e] & (BMAX - 1u);
            if (cnt == ROUND) {
                wave_rank_many<E>(sm.rh, sm.dummy, bk, lr);
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e)
                    if (e * P + t < cnt) lr[e] = atomicAdd(&sm.rh[bk[e]], 1u);
            }
            __syncthreads();  // 1
            // ---- owner of ordinary bucket jo = 2t and equality bucket je = 2t + 1
            const uint32_t jo = 2 * t, je = jo + 1;
            uint32_t kk;  // keys carried into this round
            {
                // the previous round's waiting keys of bucket jo: staged -> carry
                const uint4 rv = *reinterpret_cast<const uint4*>(&sm.rec[jo]);
                const uint32_t ts = rv.z >> 16, tn = (rv.w >> 16) & 0xFFu, cd = rv.w >> 24;
                T tmp[CW];
#pragma unroll
                for (int q = 0; q < CW; ++q)
                    if (q < (int)tn) tmp[q] = sm.stg[ts + q];
#pragma unroll
                for (int q = 0; q < CW; ++q)
                    if (q < (int)tn) sm.carry[t * CW + cd + q] = tmp[q];
                kk = cd + tn;
            }
            const uint32_t ho = sm.rh[jo], he = sm.rh[je];
            const uint32_t curo = sm.cursor[jo], cure = sm.cursor[je];
            const uint32_t d_o = curo - kk, eo = curo + ho;
            // written: up to the last line boundary at or before the end (then
            // wo >= kk, the carried keys lie in the line that starts at d_o)
            const uint32_t nk = kk + ho, past = (dph + eo) & (LINE - 1u);
            const uint32_t wo = last ? nk : (nk > past ? nk - past : 0u);
            const uint32_t ao = (dph + d_o) & (LINE - 1u), ae = (dph + cure) & (LINE - 1u);
            const uint32_t lo_ = wo ? (ao + wo + LINE - 1) / LINE : 0u;
            const uint32_t le_ = he ? (ae + he + LINE - 1) / LINE : 0u;
            const uint32_t v = (ho + he) | ((lo_ + le_) << 16);  // keys | jobs << 16 (each total < 65536)
            uint32_t incl = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(incl, d);
                if (lane >= (uint32_t)d) incl += y;
            }
            if (lane == 63) sm.wtot[wid] = incl;
            __syncthreads();  // 2
            uint32_t woff = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t x = sm.wtot[w];
                woff += (w < (int)wid) ? x : 0u;
                tot += x;
            }
            const uint32_t p = woff + incl - v;
            const uint32_t rso = p & 0xFFFFu, jso = p >> 16;
            const uint32_t rse = rso + ho, jse = jso + lo_;
            {
                Rec r;
                r.d = d_o;
                r.w = (uint16_t)wo;
                r.rs = (uint16_t)rso;
                r.js = (uint16_t)jso;
                r.k = (uint8_t)kk;
                r.a = (uint8_t)ao;
                if (wo) {  // the staged keys past the written ones wait, from carry slot 0
                    r.ts = (uint16_t)(rso + (wo - kk));
                    r.tn = (uint8_t)(ho - (wo - kk));
                    r.cd = 0;
                } else {   // nothing written: all of them wait, after the carried ones
                    r.ts = (uint16_t)rso;
                    r.tn = (uint8_t)ho;
                    r.cd = (uint8_t)kk;
                }
                sm.rec[jo] = r;
                Rec q;
                q.d = cure;
                q.w = (uint16_t)he;
                q.rs = (uint16_t)rse;
                q.js = (uint16_t)jse;
                q.k = 0;
                q.a = (uint8_t)ae;
                q.ts = 0;
                q.tn = 0;
                q.cd = 0;
                sm.rec[je] = q;
            }
            // bucket of every line job: the owner writes those of a bucket with
            // few lines; the whole wave writes those of each bucket with many
            // (sorted input puts a round's keys in one or two buckets)
            constexpr uint32_t FEW = 8;
            if (lo_ <= FEW)
                for (uint32_t i = 0; i < lo_; ++i) sm.jb[jso + i] = (uint16_t)jo;
            if (le_ <= FEW)
                for (uint32_t i = 0; i < le_; ++i) sm.jb[jse + i] = (uint16_t)je;
            for (uint64_t many = __ballot(lo_ > FEW); many; many &= many - 1) {
                const int src_lane = __ffsll((long long)many) - 1;
                const uint32_t js = __shfl(jso, src_lane), n = __shfl(lo_, src_lane);
                for (uint32_t i = lane; i < n; i += 64) sm.jb[js + i] = (uint16_t)(2 * (64 * wid + src_lane));
            }
            for (uint64_t many = __ballot(le_ > FEW); many; many &= many - 1) {
                const int src_lane = __ffsll((long long)many) - 1;
                const uint32_t js = __shfl(jse, src_lane), n = __shfl(le_, src_lane);
                for (uint32_t i = lane; i < n; i += 64) sm.jb[js + i] = (uint16_t)(2 * (64 * wid + src_lane) + 1);
            }
            sm.cursor[jo] = eo;
            sm.cursor[je] = cure + he;
            sm.rh[jo] = 0;
            sm.rh[je] = 0;
            __syncthreads();  // 3
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t i = e * P + t;
                if (i < cnt) sm.stg[sm.rec[bk[e]].rs + lr[e]] = cur.x[e];
            }
            __syncthreads();  // 4
            // line jobs, U per lane group at once: the LDS reads of U jobs are in
            // flight together (each job is a chain of three dependent reads)
            const uint32_t njobs = tot >> 16;
            const uint32_t g = lane / LINE, l = lane % LINE;
            constexpr int U = 4;
            for (uint32_t q0 = wid * JPI + g; q0 < njobs; q0 += NW * JPI * U) {
                uint32_t jq[U];
                uint4 rv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t q = q0 + (uint32_t)u * NW * JPI;
                    jq[u] = q < njobs ? sm.jb[q] : 0u;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) rv[u] = *reinterpret_cast<const uint4*>(&sm.rec[jq[u]]);
                T y[U];
                uint32_t dp[U];
                bool ok[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t q = q0 + (uint32_t)u * NW * JPI;
                    const uint32_t w = rv[u].y & 0xFFFFu, rs = rv[u].y >> 16, js = rv[u].z & 0xFFFFu;
                    const uint32_t k = rv[u].w & 0xFFu, a = (rv[u].w >> 8) & 0xFFu;
                    const int32_t o = (int32_t)((q - js) * LINE + l) - (int32_t)a;
                    ok[u] = q < njobs && o >= 0 && o < (int32_t)w;
                    const T* p = (uint32_t)o < k ? &sm.carry[(jq[u] >> 1) * CW + (uint32_t)o] : &sm.stg[rs + (uint32_t)o - k];
                    y[u] = ok[u] ? *p : T(0);
                    dp[u] = rv[u].x + (uint32_t)o;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (ok[u]) gdst[dp[u]] = y[u];
            }
            // the next round writes rec, jb, stg and carry only after its
            // barrier 1, which every thread reaches after these stores
        };
        Buf A, Bb;
        load_round(lo, A);
        __syncthreads();
        for (uint64_t base = lo; base < hi;) {
            do_round(base, A, Bb);
            base += ROUND;
            if (base >= hi) break;
            do_round(base, Bb, A);
            base += ROUND;
        }
        __syncthreads();  // LDS is reloaded for the next stream
    }
}

// ---------------------------------------------------------------- line-buffered scatter
// Keys only.  The rounds of k_scatter (classify or stored ids, LDS ranks,
// staging grouped by bucket), but a 128-byte line of an ordinary bucket's run
// is stored only once it is complete: the keys of the bucket's unfinished last
// line wait in an LDS line buffer (one per ordinary bucket, indexed by the
// key's place in its line), and in the round whose keys complete the line they
// are stored just before those keys, so the line reaches the L2 as back-to-back
// stores.  Measured on MI355X (2 GiB, up to 4096 sequential streams per CU):
// whole 128-byte lines stream at 4.3-5.7 TB/s; runs of 16 fp64 keys off line
// alignment, the pattern of a plain multiway scatter, at 1.6-1.9 TB/s, and
// storing a line's waiting part at the start of the round that completes it
// (this kernel's pattern) measured no faster than that.  Only the first line of each
// (chunk, bucket) run (shared with the previous chunk) and its last line at
// the chunk's end are stored partially, so the planner gives this kernel
// large chunks.  Equality buckets (odd) are stored directly.
// Measured slower than k_scatter on 2^28 keys, so it is opt-in
// (scatter_variant 9, 10): fp64 2.07 vs 1.52 ms per launch (the 60 KiB of
// line buffers leave room for 4096-key rounds only, which doubles the
// per-round barriers and scans, and most keys pass through the line buffer,
// two more LDS accesses each); int32 0.96 vs 0.85 ms, and 1.44 vs 0.79 ms on
// 1000 distinct values.
constexpr int SCATTER_LB = 4;               // scatter_shape of the line-buffered kernel
constexpr uint64_t LB_MIN_CHUNK = 1u << 18;  // keys per chunk at least (partial lines ~2 per 1000)

template <class T, bool IDS, int P, int E>
struct ScatterLbLds {
    using K = typename KeyTraits<T>::K;
    static constexpr int ROUND = P * E;
    static constexpr int LINE = 128 / (int)sizeof(T);  // keys per 128-byte line
    static constexpr int LB = LINE - 1;                // waiting keys per ordinary bucket, at most
    static constexpr bool CLS = !IDS;
    K lsrt[CLS ? FMAX : 1];
    __align__(16) uint32_t lguide[CLS ? LDS_TABLE_WORDS : 1];
    uint32_t cursor[BMAX];  // next destination of each bucket (segment-relative)
    uint32_t rh[BMAX];      // round histogram
    uint32_t dbase[BMAX];   // destination of staged slot i is dbase[bucket] + i
    uint16_t rs[BMAX];      // round start of each bucket in the staging area
    uint32_t cut[FMAX];     // ordinary bucket 2o: keys placed at or past cut[o] wait this round
    uint32_t head[FMAX];    // first line boundary of its run in this chunk
    uint16_t nw[FMAX];      // keys waiting: the places [cursor - nw, cursor)
    uint16_t nfl[FMAX];     // waiting keys whose line this round completes (stored before barrier 4)
    uint32_t wtot[P / 64];
    uint32_t dummy[DUMMY_SLOTS];
    T stg[ROUND];
    uint16_t stb[ROUND];
    T lbuf[FMAX * LB];
};

template <class T, bool IDS, int P, int E, bool LRN = false>
__global__ __launch_bounds__(P) void k_scatter_lb(const PartSeg* ps, const ChunkRef* cr, T* a, T* b,
                                                  const typename KeyTraits<T>::K* trees, const uint32_t* guides,
                                                  const uint32_t* hist, const uint64_t* bstart,
                                                  const uint32_t* nchunk, const uint16_t* ids) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    using L = ScatterLbLds<T, IDS, P, E>;
    constexpr int ROUND = L::ROUND, LINE = L::LINE, LB = L::LB;
    constexpr int NW = P / 64;
    constexpr int BPT = BMAX / P;  // buckets per thread in the scan (an even number: ordinary + equality)
    static_assert(BPT * P == BMAX && BPT % 2 == 0 && ROUND <= 32768 && LINE >= 2, "line-buffered scatter shape");
    __shared__ L sm;

    for (uint32_t wi = blockIdx.x; wi < *nchunk; wi += gridDim.x) {
    uint32_t wc = wi;
    if (gridDim.x >= *nchunk) {  // XCD-contiguous runs of chunks, as in k_scatter
        const uint32_t N = *nchunk, q = N / 8, r = N % 8, x = wi % 8;
        wc = x * q + (x < r ? x : r) + wi / 8;
    }
    const ChunkRef c = cr[wc];
    const PartSeg s = ps[c.seg];
    const uint32_t B = 2u << s.logf;
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    const T* src = (s.buf ? b : a) + s.begin;
    T* dst = (s.buf ? a : b) + s.begin;
    // line arithmetic on the element's address: dst need not be line aligned
    const uint64_t dph = (uint64_t)(uintptr_t)dst / sizeof(T);
    const uint16_t* isrc = IDS ? ids + s.begin : nullptr;
    Guide g{0, 0, 0, 0, 0, 0};
    if constexpr (!IDS) g = load_classifier<K, P, LRN>(s, trees, guides, sm.lsrt, sm.lguide);
    const uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
    for (uint32_t j = t; j < BMAX; j += P) {
        const uint32_t cur = j < B ? (uint32_t)bstart[s.bkt_base + j] + h[j] : 0u;
        sm.cursor[j] = cur;
        sm.rh[j] = 0;
        if (!(j & 1u)) {
            sm.head[j >> 1] = (uint32_t)(((dph + cur + LINE - 1) & ~(uint64_t)(LINE - 1)) - dph);
            sm.nw[j >> 1] = 0;
        }
    }
    const uint64_t lo = (uint64_t)c.idx * s.chunk;
    const uint64_t hi = lo + s.chunk < s.size ? lo + s.chunk : s.size;
    // Store waiting keys: f(o) of ordinary bucket o, from line-buffer slot 0
    // on, to the places p0(o)...  One item per (bucket, slot) over the
    // segment's B/2 ordinary buckets, consecutive threads on consecutive slots
    // of a bucket, so that a line's waiting part leaves in one store.
    const uint32_t nitems = (B / 2) * LB;
    auto store_waiting = [&](auto f_of, auto p0_of) {
        for (uint32_t k = t; k < nitems; k += P) {
            const uint32_t o = k / LB, q = k - o * LB;
            const uint32_t f = f_of(o);
            if (q < f) dst[p0_of(o, f) + q] = sm.lbuf[k];
        }
    };

    struct Buf {
        T x[E];
        uint32_t b[E];
    };
    // a round at or past the end (the prefetch after the last round) reads
    // key hi - 1 everywhere: in bounds, unconditional
    auto load_round = [&](uint64_t base, Buf& R) {
        const uint64_t bb = base < hi ? base : hi - 1;
        const uint32_t cnt = (uint32_t)((hi - bb) < (uint64_t)ROUND ? (hi - bb) : (uint64_t)ROUND);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;
            const uint64_t ic = bb + (i < cnt ? i : cnt - 1);
            R.x[e] = src[ic];
            if constexpr (IDS) R.b[e] = isrc[ic];
        }
    };
    auto do_round = [&](uint64_t base, Buf& cur, Buf& nxt) {
        const uint32_t cnt = (uint32_t)((hi - base) < (uint64_t)ROUND ? (hi - base) : (uint64_t)ROUND);
        load_round(base + ROUND, nxt);  // prefetch, unconditional so that no wait covers it
        uint32_t bk[E], lr[E];
        if constexpr (IDS) {
#pragma unroll
            for (int e = 0; e < E; ++e) bk[e] = cur.b[e];
        } else if (cnt == ROUND) {
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += CLS_NB) {
                K kx[CLS_NB];
                uint32_t kb[CLS_NB];
#pragma unroll
                for (int u = 0; u < CLS_NB; ++u) kx[u] = Tr::enc(cur.x[e0 + u]);
                classify_keys<K, CLS_NB, LRN>(s, sm.lsrt, sm.lguide, g, kx, kb);
#pragma unroll
                for (int u = 0; u < CLS_NB; ++u) bk[e0 + u] = kb[u];
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) bk[e] = classify_key<LRN>(s, sm.lsrt, sm.lguide, g, Tr::enc(cur.x[e]));
        }
        if (cnt == ROUND) {
            wave_rank_many<E>(sm.rh, sm.dummy, bk, lr);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (e * P + t < cnt) lr[e] = atomicAdd(&sm.rh[bk[e]], 1u);
        }
        __syncthreads();  // 1
        // exclusive scan of rh: thread t owns buckets t*BPT .. t*BPT+BPT-1
        uint32_t hv[BPT], v = 0;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            hv[k] = sm.rh[t * BPT + k];
            v += hv[k];
        }
        uint32_t incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += y;
        }
        if (lane == 63) sm.wtot[wid] = incl;
        __syncthreads();  // 2
        uint32_t woff = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) woff += (w < (int)wid) ? sm.wtot[w] : 0u;
        uint32_t p = woff + incl - v;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const uint32_t j = t * BPT + k;
            const uint32_t c0 = sm.cursor[j], end = c0 + hv[k];
            sm.rs[j] = (uint16_t)p;
            sm.dbase[j] = c0 - p;
            if (!(k & 1)) {
                // places [c0 - w0, c0) wait (line aligned unless w0 == 0); after
                // this round the places >= cut wait: the unfinished last line, or
                // nothing while the run has not passed its chunk's first boundary
                const uint32_t o = j >> 1, hd = sm.head[o], w0 = sm.nw[o];
                const int64_t le = (int64_t)((dph + end) & ~(uint64_t)(LINE - 1)) - (int64_t)dph;
                const uint32_t ct = end <= hd ? end : (le > (int64_t)hd ? (uint32_t)le : hd);
                const bool done = w0 && ct > c0 - w0;  // their line is complete (else ct == c0 - w0)
                sm.nfl[o] = (uint16_t)(done ? w0 : 0u);
                sm.cut[o] = ct;
                sm.nw[o] = (uint16_t)(end - ct);
            }
            sm.cursor[j] = end;
            sm.rh[j] = 0;
            p += hv[k];
        }
        __syncthreads();  // 3
        // completed lines: their waiting keys first, at the places before the
        // bucket's round start (old cursor = dbase + rs)
        store_waiting([&](uint32_t o) { return (uint32_t)sm.nfl[o]; },
                      [&](uint32_t o, uint32_t f) { return sm.dbase[2 * o] + sm.rs[2 * o] - f; });
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;
            if (i < cnt) {
                const uint32_t q = sm.rs[bk[e]] + lr[e];
                sm.stg[q] = cur.x[e];
                sm.stb[q] = (uint16_t)bk[e];
            }
        }
        __syncthreads();  // 4: staged, and the completed lines' waiting keys are read
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * P + t;  // staged keys are grouped by bucket
            if (i < cnt) {
                const uint32_t bkt = sm.stb[i] & (BMAX - 1u);
                const uint32_t d = sm.dbase[bkt] + i;
                const T y = sm.stg[i];
                bool wait = false;
                if (!(bkt & 1u)) {
                    const uint32_t ct = sm.cut[bkt >> 1];
                    if (d >= ct) {  // d - ct < LB: the unfinished line holds fewer than LINE keys
                        sm.lbuf[(bkt >> 1) * LB + (d - ct)] = y;
                        wait = true;
                    }
                }
                if (!wait) dst[d] = y;
            }
        }
        // the next round reads cut/nw/lbuf only after its barriers 1-3, which
        // every thread reaches after finishing the stores above
    };
    Buf A, Bb;
    if (lo < hi) load_round(lo, A);
    __syncthreads();
    for (uint64_t base = lo; base < hi;) {
        do_round(base, A, Bb);
        base += ROUND;
        if (base >= hi) break;
        do_round(base, Bb, A);
        base += ROUND;
    }
    __syncthreads();
    // the chunk's unfinished last lines
    store_waiting([&](uint32_t o) { return (uint32_t)sm.nw[o]; },
                  [&](uint32_t o, uint32_t f) { return sm.cursor[2 * o] - f; });
    __syncthreads();  // LDS is reloaded for the next chunk
    }
}

// Scatter variants: the workgroup shape, and whether the bucket of each key
// is read from the ids the count pass stored or classified again.
// Compute units of the current device (cached per device).
inline uint32_t device_cu_count() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
    if (dev < 64 && cache[dev] > 0) return (uint32_t)cache[dev];
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (dev < 64) cache[dev] = cus;
    return (uint32_t)cus;
}

template <class T, bool KV>
inline void launch_scatter(int shape, bool use_ids, uint32_t grid, hipStream_t st, const PartSeg* ps,
                           const ChunkRef* cr, T* a, T* b, int32_t* va, int32_t* vb,
                           const typename KeyTraits<T>::K* trees, const uint32_t* guides, const uint32_t* hist,
                           const uint64_t* bstart, const uint32_t* nchunk,
                           const uint16_t* ids, bool lrn = false, const uint32_t* runs = nullptr) {
#define S512_SCATTER(I, PP, EE, W)                                                                              \
    k_scatter<T, KV, I, PP, EE, W><<<grid, PP, 0, st>>>(ps, cr, a, b, va, vb, trees, guides, hist, bstart, nchunk, \
                                                        ids)
#define S512_SCATTER_L(PP, EE)                                                                                  \
    k_scatter<T, false, false, PP, EE, false, true><<<grid, PP, 0, st>>>(ps, cr, a, b, va, vb, trees, guides, hist, \
                                                                          bstart, nchunk, ids)
    if constexpr (!KV && learnable_key<T>()) {
        // a level of learned segments that classifies again in the scatter: the
        // learned build of the same shape (with stored ids nothing classifies)
        if (lrn && !use_ids) {
            constexpr int LE = sizeof(T) == 8 ? 8 : 16;
            switch (shape) {
                case SCATTER_LB:
                    k_scatter_lb<T, false, 512, LE, true><<<grid, 512, 0, st>>>(ps, cr, a, b, trees, guides, hist,
                                                                                  bstart, nchunk, ids);
                    break;
                case 1: S512_SCATTER_L(256, 16); break;
                case 2: S512_SCATTER_L(1024, 8); break;
                default: S512_SCATTER_L(512, 16); break;
            }
            return;
        }
    }
    if constexpr (!KV) {
        if (shape == SCATTER_FL && use_ids) {
            // persistent: about one workgroup per CU, each on a contiguous range of chunks;
            // with the level's runs counted, paired with the plain scatter (one of the two runs)
            const uint32_t cus = device_cu_count();
            const uint32_t per = (grid + cus - 1) / cus;
            const ScatterGate g1{runs, runs ? 1u : 0u}, g2{runs, 2u};
            k_scatter_fl<T, 512, 16><<<(grid + per - 1) / per, 512, 0, st>>>(ps, cr, a, b, hist, bstart, nchunk,
                                                                             ids, per, g1);
            if (runs)
                k_scatter<T, false, true, 512, 16, false><<<grid, 512, 0, st>>>(ps, cr, a, b, nullptr, nullptr, trees,
                                                                               guides, hist, bstart, nchunk, ids, g2);
            return;
        }
        if (shape == SCATTER_LB) {
            // 4096-key rounds beside 8-byte keys (the line buffers take 60 KiB), 8192 beside 4-byte keys
            constexpr int LE = sizeof(T) == 8 ? 8 : 16;
            if (use_ids) k_scatter_lb<T, true, 512, LE><<<grid, 512, 0, st>>>(ps, cr, a, b, trees, guides, hist,
                                                                                bstart, nchunk, ids);
            else k_scatter_lb<T, false, 512, LE><<<grid, 512, 0, st>>>(ps, cr, a, b, trees, guides, hist, bstart,
                                                                         nchunk, ids);
            return;
        }
    }
    if (shape == SCATTER_LB || shape == SCATTER_FL) shape = 0;  // key/value sorts; FL without stored ids
    if (use_ids) {
        switch (shape) {
            case 1: S512_SCATTER(true, 256, 16, false); break;
            case 2: S512_SCATTER(true, 1024, 8, false); break;
            case 3:
                // the deferred-line buffers do not fit beside 8-byte keys with values: shape 0
                if constexpr (KV && sizeof(T) == 8) S512_SCATTER(true, 512, 16, false);
                else S512_SCATTER(true, 512, 16, true);
                break;
            default: S512_SCATTER(true, 512, 16, false); break;
        }
    } else {
        // write combining needs the LDS the classifier tables take: shape 3 runs as shape 0
        switch (shape) {
            case 1: S512_SCATTER(false, 256, 16, false); break;
            case 2: S512_SCATTER(false, 1024, 8, false); break;
            default: S512_SCATTER(false, 512, 16, false); break;
        }
    }
#undef S512_SCATTER_L
#undef S512_SCATTER
}

// ---------------------------------------------------------------- plan
struct PlanLists {
    TileSeg* tile[8];         // per tile class
    uint32_t* tile_count;     // [8]
    uint64_t* tile_keys;      // [8] keys per class; [8] keys copied back; [10] largest
                              // bucket of the learned model's second level
    uint32_t tile_cap[8];     // class capacities (keys)
    uint32_t nclass;
    TileSeg* copy;            // pieces to copy scratch -> caller's array
    uint32_t* copy_count;
    Seg* next;                // segments for the next partition level
    uint32_t* next_count;
    uint32_t cap[10];         // entries each list holds (tile classes, copy, next)
    uint32_t* overflow;       // set when an append would pass its list's capacity
};

// Buckets are counted per workgroup in LDS first, so the global list counters
// see one atomic per (workgroup, list) instead of one per bucket.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_plan(const PartSeg* ps, const uint32_t* nseg,
                                                                     const uint64_t* bstart, const uint64_t* bsize,
                                                                     PlanLists pl) {
    constexpr int NL = 10;  // lists: tile classes 0..7, copy = 8, next = 9
    __shared__ uint32_t lcnt[NL], lbase[NL];
    __shared__ unsigned long long lkeys[NL + 2];  // + the learned second level's largest bucket at 11
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
    const PartSeg s = ps[w];
    const uint32_t B = 2u << s.logf;
    const uint32_t b = blockIdx.y * blockDim.x + threadIdx.x;
    if (threadIdx.x < NL + 2) lkeys[threadIdx.x] = 0;
    if (threadIdx.x < NL) lcnt[threadIdx.x] = 0;
    __syncthreads();
    int list = -1;
    uint32_t rank = 0, np = 1;
    uint64_t sz = 0, begin = 0;
    const uint32_t buf = 1u - s.buf;
    if (b < B) {
        sz = bsize[s.bkt_base + b];
        begin = s.begin + bstart[s.bkt_base + b];
        if (sz == 0) {
            list = -1;
        } else if ((b & 1u) || sz == 1) {
            if (buf == 1) {
                list = 8;
                np = (uint32_t)((sz + COPY_PIECE - 1) / COPY_PIECE);
            }
        } else {
            list = 9;
            for (uint32_t k = 0; k < pl.nclass; ++k)
                if (sz <= pl.tile_cap[k]) { list = (int)k; break; }
        }
        if (list >= 0) {
            rank = atomicAdd(&lcnt[list], np);
            atomicAdd(&lkeys[list], (unsigned long long)sz);
        }
        // largest bucket the learned model placed below its root (statistics):
        // the workgroup's largest, then one global atomic
        if (s.mode == 2 && s.lshift) atomicMax(&lkeys[11], (unsigned long long)sz);
    }
    __syncthreads();
    if (threadIdx.x == 0 && lkeys[11]) atomicMax((unsigned long long*)&pl.tile_keys[10], lkeys[11]);
    if (threadIdx.x < NL && lcnt[threadIdx.x]) {
        const uint32_t l = threadIdx.x;
        uint32_t* cnt = l < 8 ? &pl.tile_count[l] : (l == 8 ? pl.copy_count : pl.next_count);
        lbase[l] = atomicAdd(cnt, lcnt[l]);
        if (l <= 8) atomicAdd((unsigned long long*)&pl.tile_keys[l], lkeys[l]);
        if ((uint64_t)lbase[l] + lcnt[l] > pl.cap[l]) atomicOr(pl.overflow, 1u);
    }
    __syncthreads();
    if (list >= 0 && (uint64_t)lbase[list] + rank + np <= pl.cap[list]) {
        const uint32_t o = lbase[list] + rank;
        if (list < 8) {
            pl.tile[list][o] = TileSeg{begin, (uint32_t)sz, buf};
        } else if (list == 8) {
            for (uint32_t p = 0; p < np; ++p) {
                const uint64_t pb = (uint64_t)p * COPY_PIECE;
                pl.copy[o + p] = TileSeg{begin + pb, (uint32_t)((sz - pb) < COPY_PIECE ? (sz - pb) : COPY_PIECE), 1u};
            }
        } else {
            // a bucket of the learned root level is partitioned by the model
            // again: aux = (bucket index << 1) | 1
            const uint32_t aux = (s.mode && s.lshift == 0) ? (((b >> 1) << 1) | 1u) : 0u;
            pl.next[o] = Seg{begin, sz, buf, aux};
        }
    }
    __syncthreads();
    }
}

// ---------------------------------------------------------------- copy
template <class T, bool KV>
__global__ __launch_bounds__(256) void k_copy_back(const TileSeg* pieces, const uint32_t* count, T* a, const T* b,
                                                   int32_t* va, const int32_t* vb) {
    const uint32_t cnt = *count;
    for (uint32_t p = blockIdx.x; p < cnt; p += gridDim.x) {
        const TileSeg sg = pieces[p];
        // a piece is part of an equality bucket or a single key: all its keys
        // have the bits of its first key, so they are written, not copied
        const T key = b[sg.begin];
        for (uint32_t i = threadIdx.x; i < sg.size; i += blockDim.x) {
            a[sg.begin + i] = key;
            if constexpr (KV) va[sg.begin + i] = vb[sg.begin + i];
        }
    }
}

}  // namespace s512
