// s512_segsort.hip -- batched segmented sort: many independent arrays of
// uneven sizes, laid out back to back in one key array and described by an
// offsets array (segment i = keys[off[i] .. off[i+1])), sorted by one fixed
// sequence of kernels with no host round trip inside it, so the sequence can
// be captured once in a hipGraph and replayed.  No direct counterpart in the
// reference; it replaces a loop of SmallSort16V over many small arrays
// (sort512perf.cpp:656-661) or of Sort512::Sort (sort512.hpp:5960) per array.
//
// Sequence (every kernel reads its work count from device memory and loops,
// so the grids are fixed at enqueue time):
//   k_zero64        clear the control block (a kernel, so it may be captured)
//   k_seg_classify  bin every segment by size: up to the largest tile -> that
//                   tile class's list; larger -> a partition descriptor whose
//                   chunk / histogram / splitter / sample space is allocated
//                   with workgroup-aggregated atomics (no host scan)
//   k_seg_guard     a workspace overflow (impossible by the capacity bounds,
//                   checked anyway) empties every list so nothing runs
//   one partition level over all large segments: sample, sample sort,
//   splitter tree, count, column scan, scatter, plan (the plan appends the
//   buckets to the same tile lists as the small segments)
//   a second partition level over the buckets the first level's plan left
//   on the "next" list (still larger than the largest tile): k_seg_reset_level
//   clears the level counters and the space pools (the first level's
//   splitters, histograms and bucket tables are dead by then), k_seg_level2
//   turns every such bucket into a partition descriptor, and the same eight
//   kernels run again; their plan appends to the same tile lists
//   copy-back of equality buckets, then the tile sorts of every class
// A bucket still larger than the largest tile after two levels lands on a
// second next list; the host finishes those with the multi-level engine after
// the sequence (and every bucket of the first next list if the second level's
// descriptors do not fit the pools).
#include "s512_engine_internal.hpp"
#include "s512_tile_f64.hpp"
#include "s512_hostio.hpp"
#include "../../include/s512.h"
#include <algorithm>
#include <string.h>
#include <type_traits>
#include <vector>

namespace s512 {

constexpr uint32_t SEG_CHUNK = 32768;  // keys per chunk of a large segment
constexpr uint32_t SEG_SMAX = 8192;    // samples per large segment, at most
constexpr int NTOP = 5;                // space pools: hist, splitters, buckets, guide cells, samples

// Control block: u32 counters [0..63], pool tops u64 [32..36], keys per
// class u64 [40..55] (indices in u64 words).
constexpr int CTL_WORDS64 = 56;
constexpr int CNT_NEXT = 9, CNT_COPY = 8, CNT_NBIG = 10, CNT_NCHUNK = 11, CNT_OVERFLOW = 12, CNT_SPECIAL = 13,
              CNT_PLAN_OVERFLOW = 14, CNT_NEXT2 = 15, CNT_OVERFLOW2 = 16, CNT_NBIG2 = 17,
              CNT_NBIG1 = 18;

struct SegClassify {
    PlanLists pl;
    PartSeg* ps;
    uint32_t* nbig;
    ChunkRef* cr;
    uint32_t* nchunk;
    TileSeg* sl;
    unsigned long long* tops;  // [NTOP]
    uint32_t* overflow;
    uint32_t nbig_cap, nchunk_cap;
    uint64_t top_cap[NTOP];
    uint32_t list_cap[8];
    uint32_t target_leaf, oversample, max_logf, maxcap;
    uint64_t nkeys;  // length of the key array (and of the scratch array)
};

__global__ void k_zero64(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__device__ __forceinline__ int ceil_log2_dev(uint64_t x) { return x <= 1 ? 0 : 64 - __clzll(x - 1); }

__global__ __launch_bounds__(256) void k_seg_classify(const uint64_t* off, uint64_t nseg, SegClassify c) {
    constexpr int NL = 9;  // tile classes 0..7, large = 8
    __shared__ uint32_t lcnt[NL], lbase[NL];
    __shared__ unsigned long long lkeys[8];
    __shared__ unsigned long long ltop[NTOP], btop[NTOP];
    __shared__ uint32_t lch, bch;
    if (threadIdx.x < NL) lcnt[threadIdx.x] = 0;
    if (threadIdx.x < 8) lkeys[threadIdx.x] = 0;
    if (threadIdx.x < NTOP) ltop[threadIdx.x] = 0;
    if (threadIdx.x == 0) lch = 0;
    __syncthreads();
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint64_t begin = 0, sz = 0;
    if (i < nseg) {
        begin = off[i];
        const uint64_t end = off[i + 1];
        if (end < begin || end > c.nkeys) atomicOr(c.overflow, 2u);  // invalid offsets: nothing is sorted
        else sz = end - begin;
    }
    int list = -1;
    uint32_t rank = 0, logf = 0, nch = 0, nsamp = 0, rch = 0;
    unsigned long long rt[NTOP] = {0, 0, 0, 0, 0};
    if (sz > 1 && sz <= c.maxcap) {
        list = 0;
        while (sz > c.pl.tile_cap[list]) ++list;
        rank = atomicAdd(&lcnt[list], 1u);
        atomicAdd(&lkeys[list], (unsigned long long)sz);
    } else if (sz > c.maxcap) {
        list = 8;
        int lf = ceil_log2_dev((sz + c.target_leaf - 1) / c.target_leaf);
        lf = lf < 1 ? 1 : (lf > (int)c.max_logf ? (int)c.max_logf : lf);
        logf = (uint32_t)lf;
        const uint32_t F = 1u << logf, B = 2 * F;
        nch = (uint32_t)((sz + SEG_CHUNK - 1) / SEG_CHUNK);
        nsamp = c.oversample * F < SEG_SMAX ? c.oversample * F : SEG_SMAX;
        rank = atomicAdd(&lcnt[8], 1u);
        rch = atomicAdd(&lch, nch);
        const unsigned long long need[NTOP] = {(unsigned long long)nch * B, F, B,
                                               (unsigned long long)guide_words(F), nsamp};
#pragma unroll
        for (int q = 0; q < NTOP; ++q) rt[q] = atomicAdd(&ltop[q], need[q]);
    }
    __syncthreads();
    if (threadIdx.x < NL && lcnt[threadIdx.x]) {
        const uint32_t l = threadIdx.x;
        uint32_t* cnt = l < 8 ? &c.pl.tile_count[l] : c.nbig;
        lbase[l] = atomicAdd(cnt, lcnt[l]);
        const uint32_t cap = l < 8 ? c.list_cap[l] : c.nbig_cap;
        if ((uint64_t)lbase[l] + lcnt[l] > cap) atomicOr(c.overflow, 1u);
        if (l < 8) atomicAdd((unsigned long long*)&c.pl.tile_keys[l], lkeys[l]);
    }
    if (threadIdx.x < NTOP && ltop[threadIdx.x]) {
        const uint32_t q = threadIdx.x;
        btop[q] = atomicAdd(&c.tops[q], ltop[q]);
        if (btop[q] + ltop[q] > c.top_cap[q]) atomicOr(c.overflow, 1u);
    }
    if (threadIdx.x == 0 && lch) {
        bch = atomicAdd(c.nchunk, lch);
        if ((uint64_t)bch + lch > c.nchunk_cap) atomicOr(c.overflow, 1u);
    }
    __syncthreads();
    if (list < 0) return;
    const uint64_t o = (uint64_t)lbase[list] + rank;
    if (list < 8) {
        if (o < c.list_cap[list]) c.pl.tile[list][o] = TileSeg{begin, (uint32_t)sz, 0u};
        return;
    }
    PartSeg p;
    p.begin = begin;
    p.size = sz;
    p.buf = 0;
    p.logf = logf;
    p.nchunks = nch;
    p.nsample = nsamp;
    p.chunk = SEG_CHUNK;
    p.hist_base = btop[0] + rt[0];
    p.tree_base = btop[1] + rt[1];
    p.bkt_base = btop[2] + rt[2];
    p.guide_base = btop[3] + rt[3];
    p.kmin = 0;
    p.shift = 0;
    p.mode = 0;
    p.lshift = 0;
    p.lbase = 0;
    p.flags = 0;
    p.sample_base = btop[4] + rt[4];
    if (o >= c.nbig_cap) return;
    c.ps[o] = p;
    c.sl[o] = TileSeg{p.sample_base, nsamp, 0u};
    const uint64_t cb = (uint64_t)bch + rch;
    for (uint32_t k = 0; k < nch; ++k)
        if (cb + k < c.nchunk_cap) c.cr[cb + k] = ChunkRef{(uint32_t)o, k};
}

// On overflow every list is emptied, so the kernels that follow do nothing;
// the host then sorts every segment with the multi-level engine instead
// (bit 0: a workspace capacity) or reports the offsets as invalid (bit 1).
__global__ void k_seg_guard(uint32_t* counters) {
    if (counters[CNT_OVERFLOW] && threadIdx.x < 12) counters[threadIdx.x] = 0;
}

// Between the two levels: the partition counters and the space pools start
// again from zero (the first level's descriptors are no longer read).
__global__ void k_seg_reset_level(uint64_t* ctl) {
    uint32_t* cnt = (uint32_t*)ctl;
    if (threadIdx.x < NTOP) ctl[32 + threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        cnt[CNT_NBIG1] = cnt[CNT_NBIG];  // kept for the statistics
        cnt[CNT_NBIG] = 0;
        cnt[CNT_NCHUNK] = 0;
    }
}

// Second level: one partition descriptor per bucket on the first next list,
// from the same pools as the first level's (one thread per bucket; global
// atomics, the list is short).  Keys of such a bucket are in the scratch
// array (buf = 1), so its partition writes them back to the caller's array.
__global__ __launch_bounds__(256) void k_seg_level2(const Seg* next, const uint32_t* nnext, SegClassify c,
                                                    uint32_t* counters) {
    const uint32_t n = *nnext;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Seg sg = next[i];
        int lf = ceil_log2_dev((sg.size + c.target_leaf - 1) / c.target_leaf);
        lf = lf < 1 ? 1 : (lf > (int)c.max_logf ? (int)c.max_logf : lf);
        const uint32_t F = 1u << lf, B = 2 * F;
        const uint32_t nch = (uint32_t)((sg.size + SEG_CHUNK - 1) / SEG_CHUNK);
        const uint32_t nsamp = c.oversample * F < SEG_SMAX ? c.oversample * F : SEG_SMAX;
        const unsigned long long need[NTOP] = {(unsigned long long)nch * B, F, B,
                                               (unsigned long long)guide_words(F), nsamp};
        const uint32_t o = atomicAdd(c.nbig, 1u);
        const uint32_t cb = atomicAdd(c.nchunk, nch);
        unsigned long long rt[NTOP];
        bool fits = o < c.nbig_cap && (uint64_t)cb + nch <= c.nchunk_cap;
#pragma unroll
        for (int q = 0; q < NTOP; ++q) {
            rt[q] = atomicAdd(&c.tops[q], need[q]);
            fits = fits && rt[q] + need[q] <= c.top_cap[q];
        }
        if (!fits) {
            atomicOr(&counters[CNT_OVERFLOW2], 1u);
            continue;
        }
        PartSeg p;
        p.begin = sg.begin;
        p.size = sg.size;
        p.buf = sg.buf;
        p.logf = (uint32_t)lf;
        p.nchunks = nch;
        p.nsample = nsamp;
        p.chunk = SEG_CHUNK;
        p.hist_base = rt[0];
        p.tree_base = rt[1];
        p.bkt_base = rt[2];
        p.guide_base = rt[3];
        p.kmin = 0;
        p.shift = 0;
        p.mode = 0;
        p.lshift = 0;
        p.lbase = 0;
        p.flags = 0;
        p.sample_base = rt[4];
        c.ps[o] = p;
        c.sl[o] = TileSeg{p.sample_base, nsamp, 0u};
        for (uint32_t k = 0; k < nch; ++k) c.cr[cb + k] = ChunkRef{o, k};
        atomicAdd(&counters[CNT_NBIG2], 1u);
    }
}

// A second level whose descriptors did not all fit runs on nothing: the host
// finishes the whole first next list instead.
__global__ void k_seg_guard2(uint32_t* counters) {
    if (counters[CNT_OVERFLOW2] && threadIdx.x == 0) {
        counters[CNT_NBIG] = 0;
        counters[CNT_NCHUNK] = 0;
    }
}

// ---------------------------------------------------------------- workspace
// Device workspace of one segmented sort, carved out of one allocation.
struct SegWs {
    uint64_t* ctl;      // control block (CTL_WORDS64 words)
    uint32_t* counters; // = ctl as u32: tile counts 0..7, copy 8, next 9, nbig 10, nchunk 11, overflow 12,
                        //   special 13, plan overflow 14
    PartSeg* ps;
    ChunkRef* cr;
    TileSeg* sl;
    void* samples;
    void* trees;
    uint32_t* guides;
    uint32_t* hist;
    uint64_t* bstart;
    uint64_t* bsize;
    TileSeg* tile[8];
    TileSeg* copy;
    Seg* next;   // first level's buckets still larger than the largest tile
    Seg* next2;  // the second level's
    uint32_t next_cap;
    SegClassify cls;
    PlanLists pl2;  // the second level's plan: as cls.pl, appending to next2
    size_t bytes;
};

// Capacities are upper bounds, so no list can overflow (per level: the
// second level's segments are disjoint key ranges larger than a tile too):
//   large segments (> largest tile)   nbig   <= nkeys / (maxcap + 1)
//   chunks                                   <= nkeys / SEG_CHUNK + nbig
//   ordinary buckets of one segment   F = 2^ceil(log2(ceil(size / leaf))) < 2 size / leaf + 2,
//                                     summed <= 2 floor(nkeys / leaf) + 2 nbig + 2 = obkt
//   tile list k (segments, then the buckets of both levels)
//                                            <= segments + min(2 obkt, nkeys / (cap[k-1] + 1))
//   copy pieces (equality buckets and single keys of the first level, F of
//   each per segment; the second level writes to the caller's array)
//                                            <= 2 obkt + nkeys / COPY_PIECE
template <class T>
static SegWs seg_layout(char* base, size_t nseg, size_t nkeys, const EngineParams& prm, int ncls,
                        const uint32_t* caps) {
    using K = typename KeyTraits<T>::K;
    SegWs w;
    memset(&w, 0, sizeof(w));
    const uint32_t maxcap = caps[ncls - 1];
    const uint64_t nbig_cap = std::max<uint64_t>(1, nkeys / (maxcap + 1));
    const uint64_t nchunk_cap = nkeys / SEG_CHUNK + nbig_cap;
    const uint64_t obkt = std::min<uint64_t>(nbig_cap * FMAX, 2 * (nkeys / prm.target_leaf) + 2 * nbig_cap + 2);
    const uint64_t s_cap = std::min<uint64_t>(nbig_cap * SEG_SMAX, (uint64_t)prm.oversample * obkt);
    Arena ar{base, 0};
    w.ctl = ar.take<uint64_t>(CTL_WORDS64);
    w.counters = (uint32_t*)w.ctl;
    w.ps = ar.take<PartSeg>(nbig_cap);
    w.cr = ar.take<ChunkRef>(nchunk_cap);
    w.sl = ar.take<TileSeg>(nbig_cap);
    w.samples = ar.take<K>(s_cap);
    w.trees = ar.take<K>(obkt);
    w.guides = ar.take<uint32_t>(17 * obkt + 4 * nbig_cap);
    w.hist = ar.take<uint32_t>(nchunk_cap * BMAX);
    w.bstart = ar.take<uint64_t>(2 * obkt);
    w.bsize = ar.take<uint64_t>(2 * obkt);
    SegClassify& c = w.cls;
    for (int k = 0; k < ncls; ++k) {
        const uint64_t from_buckets = std::min<uint64_t>(2 * obkt, nkeys / (k ? caps[k - 1] + 1 : 2));
        const uint64_t cap = std::min<uint64_t>(nseg + from_buckets, 0xFFFFFFFFull);
        w.tile[k] = ar.take<TileSeg>(cap);
        c.list_cap[k] = (uint32_t)cap;
        c.pl.tile[k] = w.tile[k];
        c.pl.tile_cap[k] = caps[k];
    }
    const uint64_t copy_cap = 2 * obkt + nkeys / COPY_PIECE + 1;
    w.copy = ar.take<TileSeg>(copy_cap);
    w.next_cap = (uint32_t)std::min<uint64_t>(obkt, 0xFFFFFFFFull);
    w.next = ar.take<Seg>(w.next_cap);
    w.next2 = ar.take<Seg>(w.next_cap);
    w.bytes = ar.off;
    c.pl.nclass = (uint32_t)ncls;
    c.pl.tile_count = w.counters;
    c.pl.tile_keys = w.ctl + 40;
    c.pl.copy = w.copy;
    c.pl.copy_count = w.counters + CNT_COPY;
    c.pl.next = w.next;
    c.pl.next_count = w.counters + CNT_NEXT;
    for (int k = 0; k < ncls; ++k) c.pl.cap[k] = c.list_cap[k];
    c.pl.cap[8] = (uint32_t)std::min<uint64_t>(copy_cap, 0xFFFFFFFFull);
    c.pl.cap[9] = w.next_cap;
    c.pl.overflow = w.counters + CNT_PLAN_OVERFLOW;
    c.ps = w.ps;
    c.nbig = w.counters + CNT_NBIG;
    c.cr = w.cr;
    c.nchunk = w.counters + CNT_NCHUNK;
    c.sl = w.sl;
    c.tops = (unsigned long long*)(w.ctl + 32);
    c.overflow = w.counters + CNT_OVERFLOW;
    c.nbig_cap = (uint32_t)nbig_cap;
    c.nchunk_cap = (uint32_t)nchunk_cap;
    c.top_cap[0] = nchunk_cap * BMAX;
    c.top_cap[1] = obkt;
    c.top_cap[2] = 2 * obkt;
    c.top_cap[3] = 17 * obkt + 4 * nbig_cap;  // guide_words(F) <= 17 F + 4
    c.top_cap[4] = s_cap;
    c.target_leaf = prm.target_leaf;
    c.oversample = prm.oversample;
    c.max_logf = (uint32_t)prm.max_logf;
    c.maxcap = maxcap;
    c.nkeys = nkeys;
    w.pl2 = c.pl;
    w.pl2.next = w.next2;
    w.pl2.next_count = w.counters + CNT_NEXT2;
    return w;
}

static uint32_t grid_of(uint64_t cap, uint32_t mx) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, mx));
}

// Enqueue the whole segmented sort on st: fixed grids, no host round trip.
template <class T>
static s512_status seg_enqueue(T* a, T* b, const uint64_t* off, size_t nseg, const SegWs& w, hipStream_t st,
                               const EngineParams& prm) {
    using K = typename KeyTraits<T>::K;
    const uint32_t nbig_cap = w.cls.nbig_cap, nchunk_cap = w.cls.nchunk_cap;
    k_zero64<<<1, 64, 0, st>>>(w.ctl, CTL_WORDS64);
    S512_CHECK_LAUNCH();
    k_seg_classify<<<(uint32_t)((nseg + 255) / 256), 256, 0, st>>>(off, nseg, w.cls);
    S512_CHECK_LAUNCH();
    k_seg_guard<<<1, 64, 0, st>>>(w.counters);
    S512_CHECK_LAUNCH();
    const uint32_t* nbig = w.counters + CNT_NBIG;
    const uint32_t* nchunk = w.counters + CNT_NCHUNK;
    // one partition level over the descriptors in w.ps (count in nbig)
    auto level = [&](const PlanLists& pl) -> s512_status {
        k_sample<T><<<grid_of(nbig_cap, 2048), 256, 0, st>>>(w.ps, nbig, a, b, (K*)w.samples, SEG_SMAX, prm.seed);
        S512_CHECK_LAUNCH();
        // each sample by the smallest of four networks (1024..8192 keys) that holds it
        {
            const uint32_t g = grid_of(nbig_cap, 2048);
            K* smp = (K*)w.samples;
            k_tile_sort_range<K, 512, 2><<<g, 512, 0, st>>>(w.sl, nbig, smp, smp, 0u);
            S512_CHECK_LAUNCH();
            k_tile_sort_range<K, 512, 4><<<g, 512, 0, st>>>(w.sl, nbig, smp, smp, 1024u);
            S512_CHECK_LAUNCH();
            k_tile_sort_range<K, 512, 8><<<g, 512, 0, st>>>(w.sl, nbig, smp, smp, 2048u);
            S512_CHECK_LAUNCH();
            k_tile_sort_range<K, 512, 16><<<g, 512, 0, st>>>(w.sl, nbig, smp, smp, 4096u);
        }
        S512_CHECK_LAUNCH();
        k_build_tree<K><<<grid_of(nbig_cap, 2048), FMAX, 0, st>>>(w.ps, nbig, (K*)w.samples, SEG_SMAX, (K*)w.trees,
                                                                  w.guides);
        S512_CHECK_LAUNCH();
        launch_count<T>(prm.count_shape, grid_of(nchunk_cap, 8192), st, w.ps, w.cr, nchunk, a, b, (K*)w.trees,
                        w.guides, w.hist, w.counters + CNT_SPECIAL, nullptr);
        S512_CHECK_LAUNCH();
        k_colscan<4><<<grid_of((uint64_t)nbig_cap * CS_NSLICE, 8192), 256, 0, st>>>(w.ps, nbig, w.hist, w.bsize);
        S512_CHECK_LAUNCH();
        k_bucket_scan<<<grid_of(nbig_cap, 2048), BMAX, 0, st>>>(w.ps, nbig, w.bsize, w.bstart);
        S512_CHECK_LAUNCH();
        launch_scatter<T, false>(prm.scatter_shape, false, grid_of(nchunk_cap, 4096), st, w.ps, w.cr, a, b, nullptr,
                                 nullptr, (K*)w.trees, w.guides, w.hist, w.bstart, nchunk, nullptr);
        S512_CHECK_LAUNCH();
        k_plan<<<dim3(grid_of(nbig_cap, 2048), BMAX / 256), 256, 0, st>>>(w.ps, nbig, w.bstart, w.bsize, pl);
        S512_CHECK_LAUNCH();
        return S512_OK;
    };
    s512_status rc = level(w.cls.pl);
    if (rc) return rc;
    // the second level, over the first level's buckets larger than a tile
    k_seg_reset_level<<<1, 64, 0, st>>>(w.ctl);
    S512_CHECK_LAUNCH();
    k_seg_level2<<<grid_of((w.next_cap + 255) / 256, 256), 256, 0, st>>>(w.next, w.counters + CNT_NEXT, w.cls,
                                                                          w.counters);
    S512_CHECK_LAUNCH();
    k_seg_guard2<<<1, 64, 0, st>>>(w.counters);
    S512_CHECK_LAUNCH();
    if ((rc = level(w.pl2))) return rc;
    k_copy_back<T, false><<<2048, 256, 0, st>>>(w.copy, w.counters + CNT_COPY, a, b, nullptr, nullptr);
    S512_CHECK_LAUNCH();
    for (uint32_t k = 0; k < w.cls.pl.nclass; ++k) {
        if constexpr (std::is_same<T, double>::value)
            launch_tile_f64_auto((int)k, prm.tile_plan, w.tile[k], w.counters + k, grid_of(w.cls.list_cap[k], 32768),
                                 a, b, st);
        else
            launch_tile_fixed<T>((int)k, prm.tile_plan, w.tile[k], w.counters + k,
                                 grid_of(w.cls.list_cap[k], 32768), a, b, st);
        S512_CHECK_LAUNCH();
    }
    return S512_OK;
}

// After the sequence: finish what it could not (buckets still larger than a
// tile; every segment after a workspace overflow).
template <class T>
static s512_status seg_finish(T* a, T* b, const uint64_t* off, size_t nseg, size_t nkeys, const SegWs& w,
                              hipStream_t st, const EngineParams& prm, DeviceState& ds, s512_stats* stats) {
    s512_status rc;
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    uint64_t* h = (uint64_t*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h, w.ctl, CTL_WORDS64 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipStreamSynchronize(st));
    const uint32_t* hc = (const uint32_t*)h;
    // after a completed second level the host finishes only the second next
    // list; if the second level did not fit, the whole first one
    const bool l2 = hc[CNT_OVERFLOW2] == 0;
    const uint32_t overflow = hc[CNT_OVERFLOW], nnext = l2 ? hc[CNT_NEXT2] : hc[CNT_NEXT];
    const Seg* dnext = l2 ? w.next2 : w.next;
    if (stats) {
        stats->levels = hc[CNT_NBIG1] ? (hc[CNT_NBIG2] && l2 ? 2 : 1) : 0;
        for (uint32_t k = 0; k < w.cls.pl.nclass; ++k) {
            stats->tiles += hc[k];
            stats->keys[S512_PHASE_TILE] += h[40 + k];
        }
        stats->launches[S512_PHASE_TILE] = w.cls.pl.nclass;
    }
    if (hc[CNT_PLAN_OVERFLOW]) {
        set_error("segmented sort: plan lists overflowed");
        return S512_ERR_INTERNAL;
    }
    if (overflow & 2u) {
        set_error("invalid offsets: a segment ends before it begins or past the key array (%zu keys)", nkeys);
        return S512_ERR_INVALID;
    }
    std::vector<Seg> segs;
    if (overflow) {
        std::vector<uint64_t> ho(nseg + 1);
        S512_HIP(hipMemcpyAsync(ho.data(), off, (nseg + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < nseg; ++i)
            if (ho[i + 1] > ho[i] + 1) segs.push_back(Seg{ho[i], ho[i + 1] - ho[i], 0u, 0u});
    } else if (nnext) {
        if (nnext > w.next_cap) {
            set_error("segmented sort planned %u segments for %u slots", nnext, w.next_cap);
            return S512_ERR_INTERNAL;
        }
        if ((rc = ds.pinned2.reserve((size_t)nnext * sizeof(Seg), st))) return rc;
        Seg* hs = (Seg*)ds.pinned2.ptr;
        S512_HIP(hipMemcpyAsync(hs, dnext, (size_t)nnext * sizeof(Seg), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
        segs.assign(hs, hs + nnext);
    }
    if (stats) {
        stats->learned_fallback = overflow ? 1 : 0;  // reused: 1 = workspace overflow path taken
        stats->learned_levels = (int32_t)segs.size();
    }
    if (segs.empty()) return S512_OK;
    Profiler pf;
    pf.ds = &ds;
    LevelState ls;
    return run_levels<T, false>(a, nullptr, b, nullptr, nkeys, std::move(segs), st, prm, nullptr, pf, ds,
                                LearnedModelDev{nullptr, nullptr}, ls);
}

template <class T>
static s512_status segsort_dev(T* a, const uint64_t* off, size_t nseg, hipStream_t st, s512_stats* stats,
                               const s512_options* opt = nullptr) {
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nseg == 0) return S512_OK;
    if (!a || !off) { set_error("null pointer"); return S512_ERR_INVALID; }
    if (nseg >= (1ull << 31)) { set_error("%zu segments exceed 2^31-1", nseg); return S512_ERR_TOO_LARGE; }
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    uint64_t* h = (uint64_t*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h, off, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(h + 1, off + nseg, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipStreamSynchronize(st));
    if (h[1] < h[0]) { set_error("offsets decrease: off[0] = %llu > off[nseg] = %llu",
                                 (unsigned long long)h[0], (unsigned long long)h[1]); return S512_ERR_INVALID; }
    const size_t nkeys = h[1];  // keys addressed: [0, off[nseg])
    if (nkeys >= (1ull << 32)) { set_error("%zu keys exceed 2^32-1", nkeys); return S512_ERR_TOO_LARGE; }
    const EngineParams prm = params_for(opt, sizeof(T) == 8 ? TUNE_SEG_F64 : TUNE_SEG_I32, nkeys);
    int ncls;
    const uint32_t* caps = tile_caps<T>(&ncls);
    const SegWs probe = seg_layout<T>(nullptr, nseg, nkeys, prm, ncls, caps);
    if ((rc = ds.arena.reserve(probe.bytes, st))) return rc;
    if ((rc = ds.data.reserve(std::max<size_t>(nkeys, 1) * sizeof(T), st))) return rc;
    const SegWs w = seg_layout<T>((char*)ds.arena.ptr, nseg, nkeys, prm, ncls, caps);
    T* b = (T*)ds.data.ptr;
    if ((rc = seg_enqueue<T>(a, b, off, nseg, w, st, prm))) return rc;
    return seg_finish<T>(a, b, off, nseg, nkeys, w, st, prm, ds, stats);
}

// ---------------------------------------------------------------- graph plan
struct SegPlan {
    int dtype;
    void* keys;
    const uint64_t* off;
    size_t nseg, nkeys;
    void* ws;       // workspace
    void* scratch;  // second key array
    SegWs w;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int dev;
};

template <class T>
static s512_status plan_create(SegPlan* p) {
    s512_status rc = current_device(&p->dev);
    if (rc) return rc;
    const EngineParams prm = params_for(nullptr, sizeof(T) == 8 ? TUNE_SEG_F64 : TUNE_SEG_I32, p->nkeys);
    int ncls;
    const uint32_t* caps = tile_caps<T>(&ncls);
    const SegWs probe = seg_layout<T>(nullptr, p->nseg, p->nkeys, prm, ncls, caps);
    S512_HIP(hipMalloc(&p->ws, probe.bytes));
    S512_HIP(hipMalloc(&p->scratch, std::max<size_t>(p->nkeys, 1) * sizeof(T)));
    p->w = seg_layout<T>((char*)p->ws, p->nseg, p->nkeys, prm, ncls, caps);
    if (p->nseg == 0) return S512_OK;
    hipStream_t cs;
    S512_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
        return S512_ERR_HIP;
    }
    rc = seg_enqueue<T>((T*)p->keys, (T*)p->scratch, p->off, p->nseg, p->w, cs, prm);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) {
        set_error("graph capture failed: %s", hipGetErrorString(e));
        return S512_ERR_HIP;
    }
    p->graph = g;
    S512_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    return S512_OK;
}

template <class T>
static s512_status plan_launch(SegPlan* p, hipStream_t st, s512_stats* stats) {
    if (stats) memset(stats, 0, sizeof(*stats));
    if (p->nseg == 0) return S512_OK;
    S512_HIP(hipGraphLaunch(p->exec, st));
    DeviceState& ds = device_state(p->dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    const EngineParams prm = params_for(nullptr, sizeof(T) == 8 ? TUNE_SEG_F64 : TUNE_SEG_I32, p->nkeys);
    return seg_finish<T>((T*)p->keys, (T*)p->scratch, p->off, p->nseg, p->nkeys, p->w, st, prm, ds, stats);
}

// Host arrays: stage keys and offsets on the device (pipelined copies).
template <class T>
static s512_status segsort_host(T* keys, const uint64_t* off, size_t nseg) {
    if (nseg == 0) return S512_OK;
    if (!keys || !off) { set_error("null pointer"); return S512_ERR_INVALID; }
    const size_t nkeys = off[nseg];
    if (nkeys < off[0]) { set_error("offsets decrease"); return S512_ERR_INVALID; }
    HostStage hs;
    void* dk = nullptr;
    void* doff = nullptr;
    if (s512_status rc = hs.in(0, keys, nkeys * sizeof(T), &dk)) return rc;
    if (s512_status rc = hs.in(2, off, (nseg + 1) * sizeof(uint64_t), &doff)) return rc;
    if (s512_status rc = segsort_dev<T>((T*)dk, (const uint64_t*)doff, nseg, (hipStream_t)0, nullptr)) return rc;
    return hs.out(0, keys, nkeys * sizeof(T));
}

}  // namespace s512

using namespace s512;

extern "C" s512_status s512_segsort_i32_dev(int32_t* d_keys, const uint64_t* d_offsets, size_t nseg,
                                            s512_stats* stats, void* stream) {
    return segsort_dev<int32_t>(d_keys, d_offsets, nseg, (hipStream_t)stream, stats);
}
extern "C" s512_status s512_segsort_f64_dev(double* d_keys, const uint64_t* d_offsets, size_t nseg,
                                            s512_stats* stats, void* stream) {
    return segsort_dev<double>(d_keys, d_offsets, nseg, (hipStream_t)stream, stats);
}
extern "C" s512_status s512_segsort_i32_dev_ex(int32_t* d_keys, const uint64_t* d_offsets, size_t nseg,
                                               const s512_options* opt, s512_stats* stats, void* stream) {
    return segsort_dev<int32_t>(d_keys, d_offsets, nseg, (hipStream_t)stream, stats, opt);
}
extern "C" s512_status s512_segsort_f64_dev_ex(double* d_keys, const uint64_t* d_offsets, size_t nseg,
                                               const s512_options* opt, s512_stats* stats, void* stream) {
    return segsort_dev<double>(d_keys, d_offsets, nseg, (hipStream_t)stream, stats, opt);
}
extern "C" s512_status s512_segsort_i32(int32_t* keys, const uint64_t* offsets, size_t nseg) {
    return segsort_host<int32_t>(keys, offsets, nseg);
}
extern "C" s512_status s512_segsort_f64(double* keys, const uint64_t* offsets, size_t nseg) {
    return segsort_host<double>(keys, offsets, nseg);
}

struct s512_segsort_plan_t {
    SegPlan p;
};

extern "C" s512_status s512_segsort_plan_destroy(s512_segsort_plan plan) {
    if (!plan) return S512_OK;
    if (plan->p.exec) (void)hipGraphExecDestroy(plan->p.exec);
    if (plan->p.graph) (void)hipGraphDestroy(plan->p.graph);
    if (plan->p.ws) (void)hipFree(plan->p.ws);
    if (plan->p.scratch) (void)hipFree(plan->p.scratch);
    delete plan;
    return S512_OK;
}

extern "C" s512_status s512_segsort_plan_create(int dtype, void* d_keys, const uint64_t* d_offsets, size_t nseg,
                                                size_t nkeys, s512_segsort_plan* out) {
    if (!out) { set_error("null plan pointer"); return S512_ERR_INVALID; }
    *out = nullptr;
    if (dtype != S512_DTYPE_I32 && dtype != S512_DTYPE_F64) { set_error("unknown dtype %d", dtype); return S512_ERR_INVALID; }
    if (nseg && (!d_keys || !d_offsets)) { set_error("null pointer"); return S512_ERR_INVALID; }
    if (nseg >= (1ull << 31)) { set_error("%zu segments exceed 2^31-1", nseg); return S512_ERR_TOO_LARGE; }
    if (nkeys >= (1ull << 32)) { set_error("%zu keys exceed 2^32-1", nkeys); return S512_ERR_TOO_LARGE; }
    s512_segsort_plan_t* h = new s512_segsort_plan_t();
    memset(&h->p, 0, sizeof(h->p));
    h->p.dtype = dtype;
    h->p.keys = d_keys;
    h->p.off = d_offsets;
    h->p.nseg = nseg;
    h->p.nkeys = nkeys;
    const s512_status rc = dtype == S512_DTYPE_I32 ? plan_create<int32_t>(&h->p) : plan_create<double>(&h->p);
    if (rc) {
        s512_segsort_plan_destroy(h);
        return rc;
    }
    *out = h;
    return S512_OK;
}

extern "C" s512_status s512_segsort_plan_launch(s512_segsort_plan plan, s512_stats* stats, void* stream) {
    if (!plan) { set_error("null plan"); return S512_ERR_INVALID; }
    return plan->p.dtype == S512_DTYPE_I32 ? plan_launch<int32_t>(&plan->p, (hipStream_t)stream, stats)
                                           : plan_launch<double>(&plan->p, (hipStream_t)stream, stats);
}
