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

namespace s512 {

constexpr int FMAX_LOG = 9;                 // at most 512-way fan-out per level
constexpr int FMAX = 1 << FMAX_LOG;
constexpr int BMAX = 2 * FMAX;              // ordinary + equality buckets
constexpr int SCAT_P = 512;                 // scatter workgroup
constexpr int SCAT_E = 16;                  // keys per thread per round
constexpr int SCAT_ROUND = SCAT_P * SCAT_E; // keys staged in LDS per round
constexpr int CNT_P = 256;                  // count workgroup
constexpr int CNT_U = 16;                   // loads in flight per thread
constexpr uint32_t COPY_PIECE = 1u << 16;   // keys per copy work item

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
// Splitters of a segment: srt[0 .. F-1) sorted and de-duplicated, padded with
// KMAX, srt[F-1] = KMAX.  Bucket of key x: j = #{srt < x} ordinary bucket 2j,
// or the equality bucket 2j+1 when x == srt[j].  j is found with a guide
// table over the segment's encoded key range: cell c = (x - kmin) >> shift
// (clamped) stores [j0, j1], the values j can take for keys in the cell, and a
// binary search inside that range finishes (usually zero or one step).
constexpr int GUIDE_PER_SPLIT = 8;
constexpr int GMAX = GUIDE_PER_SPLIT * FMAX;

struct Guide {
    uint64_t kmin;
    uint32_t shift;
    uint32_t gmask;  // cells - 1
};

template <class K>
__device__ __forceinline__ uint32_t guide_cell(const Guide& g, K x) {
    if ((uint64_t)x < g.kmin) return 0;
    const uint64_t c = ((uint64_t)x - g.kmin) >> g.shift;
    return c > g.gmask ? g.gmask : (uint32_t)c;
}

template <class K>
__device__ __forceinline__ uint32_t classify(const K* srt, const uint32_t* guide, const Guide& g, K x) {
    const uint32_t e = guide[guide_cell(g, x)];
    uint32_t lo = e & 0xFFFFu, hi = e >> 16;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (srt[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return 2 * lo + (srt[lo] == x ? 1u : 0u);
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

// Load a segment's classifier into LDS: splitters + guide table, or (learned
// segments) the tabulated model in the guide table's place.
constexpr int LDS_TABLE_WORDS = GMAX > LEARN_G + 1 ? GMAX : LEARN_G + 1;
template <class K, int NT>
__device__ __forceinline__ Guide load_classifier(const PartSeg& s, const K* trees, const uint32_t* guides, K* lsrt,
                                                 uint32_t* lguide, const LearnedTable& lt) {
    if (s.mode) {
        for (uint32_t j = threadIdx.x; j <= LEARN_G; j += NT) lguide[j] = __float_as_uint(lt.tab[j]);
        return Guide{0, 0, 0};
    }
    const uint32_t F = 1u << s.logf, G = GUIDE_PER_SPLIT * F;
    for (uint32_t j = threadIdx.x; j < F; j += NT) lsrt[j] = trees[s.tree_base + j];
    for (uint32_t j = threadIdx.x; j < G; j += NT) lguide[j] = guides[s.guide_base + j];
    return Guide{s.kmin, s.shift, G - 1};
}

// Bucket of key x in segment s (learned segments use only even buckets).
template <class K>
__device__ __forceinline__ uint32_t classify_seg(const PartSeg& s, const K* lsrt, const uint32_t* lguide,
                                                 const Guide& g, const LearnedTable& lt, K x) {
    if (s.mode) {
        const uint32_t fb = learned_fine_bucket((const float*)lguide, lt.kmin, lt.inv_range, lt.nb_log, x);
        int b = (int)(fb >> s.lshift) - (int)s.lbase;
        const int fmax = (1 << s.logf) - 1;
        b = b < 0 ? 0 : (b > fmax ? fmax : b);
        return 2u * (uint32_t)b;
    }
    return classify(lsrt, lguide, g, x);
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
template <class T, int P, int E, class Pol = DefaultOrder<T>>
__global__ __launch_bounds__(P) void k_tile_sort(const TileSeg* segs, const uint32_t* count, T* a, const T* b) {
    using IO = KeyIO<T, Pol>;
    using K = typename Pol::key_t;
    using BB = BlockBitonic<Pol, P, E>;
    __shared__ K lds[BB::N];
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
    __shared__ uint64_t lds[BB::N];
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
// padded with KMAX to F entries; then the guide table over [kmin, kmax] of
// the sample.  One workgroup per segment.  Writes kmin/shift into the
// segment descriptor for the count and scatter kernels.
template <class K>
__device__ __forceinline__ void build_tree_body(PartSeg& s, const K* smp, K* trees, uint32_t* guides, K* srt,
                                                uint32_t* wtot) {
    const uint32_t F = 1u << s.logf, G = GUIDE_PER_SPLIT * F;
    const uint32_t ns = s.nsample;
    const uint32_t t = threadIdx.x;
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
    const uint32_t pos = block_excl_scan<FMAX>(keep ? 1u : 0u, wtot, &nuniq);
    if (t < F) srt[t] = (K)~(K)0;
    __syncthreads();
    if (keep) srt[pos] = c;
    __syncthreads();
    K* out = trees + s.tree_base;
    if (t < F) out[t] = srt[t];
    // guide table over the sample's key range
    const uint64_t kmin = (uint64_t)smp[0], kmax = (uint64_t)smp[ns - 1];
    const uint64_t range = kmax - kmin;
    uint32_t shift = 0;
    while ((range >> shift) >= G) ++shift;
    for (uint32_t cc = t; cc < G; cc += FMAX) {
        const uint64_t off0 = (uint64_t)cc << shift, off1 = (uint64_t)(cc + 1) << shift;
        const uint64_t kmaxk = (uint64_t)(K)~(K)0;
        const K lo_key = (K)(off0 > kmaxk - kmin ? kmaxk : kmin + off0);
        const K hi_key = (K)(off1 > kmaxk - kmin ? kmaxk : kmin + off1);
        const uint32_t j0 = cc == 0 ? 0u : lower_bound_lds(srt, F - 1, lo_key);
        const uint32_t j1 = cc == G - 1 ? F - 1 : lower_bound_lds(srt, F - 1, hi_key);
        guides[s.guide_base + cc] = j0 | (j1 << 16);
    }
    if (t == 0) {
        s.kmin = kmin;
        s.shift = shift;
    }
}

template <class K>
__global__ __launch_bounds__(FMAX) void k_build_tree(PartSeg* ps, const uint32_t* nseg, const K* samples,
                                                     uint32_t smax, K* trees, uint32_t* guides) {
    __shared__ K srt[FMAX];
    __shared__ uint32_t wtot[FMAX / 64];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        PartSeg& s = ps[w];
        if (s.mode || s.nsample == 0) continue;  // learned segment: no splitters
        build_tree_body<K>(s, samples + s.sample_base, trees, guides, srt, wtot);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- count
// Encoded fp64 keys that are NaN or -0.0 (they rule out the min/max leaf sort).
__device__ __forceinline__ bool nan_or_negzero(uint64_t k) {
    return k == 0x7FFFFFFFFFFFFFFFull || k > 0xFFF0000000000000ull || k < 0x000FFFFFFFFFFFFFull;
}
__device__ __forceinline__ bool nan_or_negzero(uint32_t) { return false; }

template <class T>
__global__ __launch_bounds__(CNT_P) void k_count(const PartSeg* ps, const ChunkRef* cr, const uint32_t* nchunk,
                                                 const T* a, const T* b, const typename KeyTraits<T>::K* trees,
                                                 const uint32_t* guides, uint32_t* hist, uint32_t* special,
                                                 LearnedTable lt) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    __shared__ K lsrt[FMAX];
    __shared__ uint32_t lguide[LDS_TABLE_WORDS];
    __shared__ uint32_t lh[BMAX];
    for (uint32_t w = blockIdx.x; w < *nchunk; w += gridDim.x) {
    const ChunkRef c = cr[w];
    const PartSeg s = ps[c.seg];
    const uint32_t F = 1u << s.logf, B = 2 * F;
    const uint32_t t = threadIdx.x;
    const Guide g = load_classifier<K, CNT_P>(s, trees, guides, lsrt, lguide, lt);
    for (uint32_t j = t; j < B; j += CNT_P) lh[j] = 0;
    __syncthreads();
    const T* src = (s.buf ? b : a) + s.begin;
    const uint64_t lo = (uint64_t)c.idx * s.chunk;
    const uint64_t hi = lo + s.chunk < s.size ? lo + s.chunk : s.size;
    bool odd = false;
    for (uint64_t base = lo; base < hi; base += (uint64_t)CNT_P * CNT_U) {
        // clamped indices keep every load in bounds and unconditional, so all
        // CNT_U loads of a thread are in flight together
        K x[CNT_U];
#pragma unroll
        for (int u = 0; u < CNT_U; ++u) {
            const uint64_t i = base + (uint64_t)u * CNT_P + t;
            x[u] = Tr::enc(src[i < hi ? i : hi - 1]);
        }
#pragma unroll
        for (int u = 0; u < CNT_U; ++u) {
            const uint64_t i = base + (uint64_t)u * CNT_P + t;
            if (i < hi) {
                odd |= nan_or_negzero(x[u]);
                atomicAdd(&lh[classify_seg(s, lsrt, lguide, g, lt, x[u])], 1u);
            }
        }
    }
    if (__any(odd) && (t & 63u) == 0) atomicOr(special, 1u);
    __syncthreads();
    uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
    for (uint32_t j = t; j < B; j += CNT_P) h[j] = lh[j];
    __syncthreads();
    }
}

// ---------------------------------------------------------------- column scan
// hist[chunk][bucket] -> prefix over earlier chunks of the same bucket;
// bstart[b] = first position of bucket b inside the segment; bsize[b] = size.
[[maybe_unused]] static __global__ __launch_bounds__(BMAX) void k_colscan(const PartSeg* ps, const uint32_t* nseg,
                                                                         uint32_t* hist, uint64_t* bstart,
                                                                         uint64_t* bsize) {
    __shared__ uint32_t wtot[BMAX / 64];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
    const PartSeg s = ps[w];
    const uint32_t B = 2u << s.logf;
    const uint32_t t = threadIdx.x;
    uint32_t run = 0;
    if (t < B) {
        uint32_t* h = hist + s.hist_base + t;
        for (uint32_t c = 0; c < s.nchunks; ++c) {
            const uint32_t v = h[(uint64_t)c * B];
            h[(uint64_t)c * B] = run;
            run += v;
        }
    }
    uint32_t total;
    const uint32_t start = block_excl_scan<BMAX>(run, wtot, &total);
    if (t < B) {
        bstart[s.bkt_base + t] = start;
        bsize[s.bkt_base + t] = run;
    }
    }
}

// ---------------------------------------------------------------- scatter
// Keys (and values for key/value sorts) move from array buf to array 1-buf.
// Per round of SCAT_ROUND keys:
//   classify, rank inside the round with LDS atomics           | barrier 1
//   wave-level scan of the round histogram (2 buckets/thread)  | barrier 2
//   bucket owners: round start rs, destination base, advance the
//   chunk's cursor, clear the counter for the next round       | barrier 3
//   stage keys grouped by bucket in LDS                        | barrier 4
//   write the staged keys: consecutive threads, consecutive addresses
// The next round's keys are loaded into registers while this round runs.
template <class T, bool KV>
__global__ __launch_bounds__(SCAT_P) void k_scatter(const PartSeg* ps, const ChunkRef* cr, T* a, T* b,
                                                    int32_t* va, int32_t* vb,
                                                    const typename KeyTraits<T>::K* trees, const uint32_t* guides,
                                                    const uint32_t* hist, const uint64_t* bstart, LearnedTable lt,
                                                    const uint32_t* nchunk) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    constexpr int NW = SCAT_P / 64;
    static_assert(2 * SCAT_P == BMAX, "scan: two buckets per thread");
    __shared__ K lsrt[FMAX];
    __shared__ uint32_t lguide[LDS_TABLE_WORDS];
    __shared__ uint32_t cursor[BMAX];  // next destination of each bucket (segment-relative)
    __shared__ uint32_t rh[BMAX];      // round histogram
    __shared__ uint32_t rs[BMAX];      // round start of each bucket in the staging area
    __shared__ uint32_t dbase[BMAX];   // destination of staged slot i is dbase[bucket] + i
    __shared__ uint32_t wtot[NW];
    __shared__ T stg[SCAT_ROUND];
    __shared__ uint16_t stb[SCAT_ROUND];
    __shared__ int32_t stv[KV ? SCAT_ROUND : 1];

    for (uint32_t wi = blockIdx.x; wi < *nchunk; wi += gridDim.x) {
    const ChunkRef c = cr[wi];
    const PartSeg s = ps[c.seg];
    const uint32_t F = 1u << s.logf, B = 2 * F;
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    const Guide g = load_classifier<K, SCAT_P>(s, trees, guides, lsrt, lguide, lt);
    const uint32_t* h = hist + s.hist_base + (uint64_t)c.idx * B;
    for (uint32_t j = t; j < BMAX; j += SCAT_P) {
        cursor[j] = j < B ? (uint32_t)bstart[s.bkt_base + j] + h[j] : 0u;
        rh[j] = 0;
    }
    const T* src = (s.buf ? b : a) + s.begin;
    T* dst = (s.buf ? a : b) + s.begin;
    const int32_t* vsrc = KV ? ((s.buf ? vb : va) + s.begin) : nullptr;
    int32_t* vdst = KV ? ((s.buf ? va : vb) + s.begin) : nullptr;
    const uint64_t lo = (uint64_t)c.idx * s.chunk;
    const uint64_t hi = lo + s.chunk < s.size ? lo + s.chunk : s.size;

    T x[SCAT_E], xn[SCAT_E];
    int32_t xv[SCAT_E], xvn[SCAT_E];
    auto load_round = [&](uint64_t base, T (&dx)[SCAT_E], int32_t (&dv)[SCAT_E]) {
        const uint32_t cnt = (uint32_t)((hi - base) < (uint64_t)SCAT_ROUND ? (hi - base) : (uint64_t)SCAT_ROUND);
#pragma unroll
        for (int e = 0; e < SCAT_E; ++e) {
            const uint32_t i = e * SCAT_P + t;
            const uint64_t ic = base + (i < cnt ? i : cnt - 1);  // in bounds, unconditional
            dx[e] = src[ic];
            if constexpr (KV) dv[e] = vsrc[ic];
        }
    };
    if (lo < hi) load_round(lo, x, xv);
    __syncthreads();

    for (uint64_t base = lo; base < hi; base += SCAT_ROUND) {
        const uint32_t cnt = (uint32_t)((hi - base) < (uint64_t)SCAT_ROUND ? (hi - base) : (uint64_t)SCAT_ROUND);
        if (base + SCAT_ROUND < hi) load_round(base + SCAT_ROUND, xn, xvn);  // prefetch
        uint32_t bk[SCAT_E], lr[SCAT_E];
#pragma unroll
        for (int e = 0; e < SCAT_E; ++e) {
            const uint32_t i = e * SCAT_P + t;
            bk[e] = classify_seg(s, lsrt, lguide, g, lt, Tr::enc(x[e]));
            if (i < cnt) lr[e] = atomicAdd(&rh[bk[e]], 1u);
        }
        __syncthreads();  // 1
        // exclusive scan of rh: thread t owns buckets 2t, 2t+1
        const uint32_t j0 = 2 * t, j1 = 2 * t + 1;
        const uint32_t h0 = rh[j0], h1 = rh[j1];
        uint32_t v = h0 + h1, incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += y;
        }
        if (lane == 63) wtot[wid] = incl;
        __syncthreads();  // 2
        uint32_t woff = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) woff += (w < (int)wid) ? wtot[w] : 0u;
        const uint32_t p0 = woff + incl - v, p1 = p0 + h0;
        rs[j0] = p0;
        rs[j1] = p1;
        dbase[j0] = cursor[j0] - p0;
        dbase[j1] = cursor[j1] - p1;
        cursor[j0] += h0;
        cursor[j1] += h1;
        rh[j0] = 0;
        rh[j1] = 0;
        __syncthreads();  // 3
#pragma unroll
        for (int e = 0; e < SCAT_E; ++e) {
            const uint32_t i = e * SCAT_P + t;
            if (i < cnt) {
                const uint32_t p = rs[bk[e]] + lr[e];
                stg[p] = x[e];
                stb[p] = (uint16_t)bk[e];
                if constexpr (KV) stv[p] = xv[e];
            }
        }
        __syncthreads();  // 4
#pragma unroll
        for (int e = 0; e < SCAT_E; ++e) {
            const uint32_t i = e * SCAT_P + t;  // staged keys are grouped by bucket
            const uint32_t bkt = stb[i] & (BMAX - 1u);  // slots >= cnt: stale, never stored
            const uint64_t d = (uint64_t)dbase[bkt] + i;
            const T y = stg[i];
            int32_t yv = 0;
            if constexpr (KV) yv = stv[i];
            if (i < cnt) {
                dst[d] = y;
                if constexpr (KV) vdst[d] = yv;
            }
        }
#pragma unroll
        for (int e = 0; e < SCAT_E; ++e) {
            x[e] = xn[e];
            if constexpr (KV) xv[e] = xvn[e];
        }
        // the next round's atomics touch rh (cleared before barrier 3) and the
        // staging area is rewritten only after its barrier 3, which every
        // thread reaches after finishing the stores above
    }
    __syncthreads();  // LDS is reloaded for the next chunk
    }
}

// ---------------------------------------------------------------- plan
struct PlanLists {
    TileSeg* tile[8];         // per tile class
    uint32_t* tile_count;     // [8]
    uint64_t* tile_keys;      // [8] keys per class; [8] keys copied back
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
    __shared__ unsigned long long lkeys[NL];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
    const PartSeg s = ps[w];
    const uint32_t B = 2u << s.logf;
    const uint32_t b = blockIdx.y * blockDim.x + threadIdx.x;
    if (threadIdx.x < NL) {
        lcnt[threadIdx.x] = 0;
        lkeys[threadIdx.x] = 0;
    }
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
    }
    __syncthreads();
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
            // a bucket of a learned level with a finer level below keeps the
            // learned classifier: aux = (bucket index << 1) | 1
            const uint32_t aux = (s.mode && s.lshift) ? ((((b >> 1) + s.lbase) << 1) | 1u) : 0u;
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
        for (uint32_t i = threadIdx.x; i < sg.size; i += blockDim.x) {
            a[sg.begin + i] = b[sg.begin + i];
            if constexpr (KV) va[sg.begin + i] = vb[sg.begin + i];
        }
    }
}

}  // namespace s512
