// s512_sort.hip -- the sort engine (multiway sample sort + bitonic tiles) and
// its C entry points: Sort512::Sort (sort512.hpp:5960) and Sort512kv::Sort
// (sort512kv.hpp:6132).
//
// Host loop, one iteration per partition level:
//   1. describe every open segment (fan-out, sample size, chunks, offsets)
//   2. upload the descriptors, run sample -> sample tile sort -> tree ->
//      count -> column scan -> scatter -> plan on the stream
//   3. read back the plan counters; launch the copy-back and the tile sorts of
//      finished buckets; read back the segments that need another level.
// Tile sorts of one level overlap the partition kernels of the next (they work
// on disjoint ranges and are ordered on the same stream).
#include "s512_kernels.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"
#include "s512_tile_f64.hpp"
#include "s512_learned_host.hpp"
#include "s512_engine_internal.hpp"
#include "s512_hostio.hpp"
#include <algorithm>
#include <cmath>
#include <type_traits>

namespace s512 {

// ---------------------------------------------------------------- tile classes
static const uint32_t kCaps32[] = {256, 512, 1024, 2048, 4096, 8192, 16384};
static const uint32_t kCaps64[] = {256, 512, 1024, 2048, 4096, 8192, 16384};  // 8-byte keys
static const uint32_t kCapsKV[] = {256, 512, 1024, 2048, 4096, 8192};         // key/value pairs (and segmented 8-byte keys)

// Tile shapes (threads P x keys per thread E) per class and plan.  Plan 0
// keeps 8 keys per thread up to 2048-key tiles; plan 1 uses 16; plan 2 uses 32
// (a 2048-key tile is one wave: the stages on lane bits of the first three
// merge phases above the registers are DPP moves, every later window one
// wave-local LDS round trip, no workgroup barrier).  Every shape's LDS layout
// is conflict-free (BlockBitonic::phys).
constexpr int TILE_PLANS = 3;

template <class T, class Pol>
static void launch_tile_keys_pol(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* a,
                                 const T* b, hipStream_t st) {
    using K = typename Pol::key_t;
    if (count == 0) return;
    const uint32_t gw = (count + 3) / 4;
#define S512_TILE(PP, EE) k_tile_sort<T, PP, EE, Pol><<<count, PP, 0, st>>>(segs, dcount, a, b)
    if constexpr (sizeof(K) == 4) {
        if (cls == 0) { k_wave_sort<T, 4, Pol><<<gw, 256, 0, st>>>(segs, dcount, a, b); return; }
        if (cls == 1) { k_wave_sort<T, 8, Pol><<<gw, 256, 0, st>>>(segs, dcount, a, b); return; }
        switch (plan * 8 + cls) {
            case 2: S512_TILE(128, 8); break;
            case 3: S512_TILE(256, 8); break;
            case 4: S512_TILE(256, 16); break;
            case 5: S512_TILE(512, 16); break;
            case 8 + 2: S512_TILE(64, 16); break;
            case 8 + 3: S512_TILE(128, 16); break;
            case 8 + 4: S512_TILE(256, 16); break;
            case 8 + 5: S512_TILE(512, 16); break;
            case 16 + 2: S512_TILE(64, 16); break;
            case 16 + 3: S512_TILE(64, 32); break;
            case 16 + 4: S512_TILE(128, 32); break;
            case 16 + 5: S512_TILE(256, 32); break;
            case 16 + 6: S512_TILE(512, 32); break;
            default: S512_TILE(1024, 16); break;  // class 6 of plans 0, 1
        }
    } else {
        if (cls == 0) { k_wave_sort<T, 4, Pol><<<gw, 256, 0, st>>>(segs, dcount, a, b); return; }
        switch (plan * 8 + cls) {
            case 1: S512_TILE(128, 4); break;
            case 2: S512_TILE(128, 8); break;
            case 3: S512_TILE(256, 8); break;
            case 4: S512_TILE(256, 16); break;
            case 8 + 1: S512_TILE(64, 8); break;
            case 8 + 2: S512_TILE(64, 16); break;
            case 8 + 3: S512_TILE(128, 16); break;
            case 8 + 4: S512_TILE(256, 16); break;
            case 16 + 1: S512_TILE(64, 8); break;
            case 16 + 2: S512_TILE(64, 16); break;
            case 16 + 3: S512_TILE(64, 32); break;
            case 16 + 4: S512_TILE(128, 32); break;
            case 16 + 5: S512_TILE(256, 32); break;
            case 6:
            case 8 + 6: S512_TILE(1024, 16); break;  // 128 KiB of LDS
            case 16 + 6: S512_TILE(512, 32); break;
            default: S512_TILE(512, 16); break;  // class 5 of plans 0, 1
        }
    }
#undef S512_TILE
}

template <class T>
static void launch_tile_keys(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* a,
                             const T* b, hipStream_t st) {
    launch_tile_keys_pol<T, DefaultOrder<T>>(cls, plan, segs, dcount, count, a, b, st);
}

static void launch_tile_kv(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count,
                           int32_t* ka, int32_t* va, const int32_t* kb, const int32_t* vb, hipStream_t st) {
    if (count == 0) return;
    const uint32_t gw = (count + 3) / 4;
#define S512_TKV(PP, EE) k_tile_sort_kv<PP, EE><<<count, PP, 0, st>>>(segs, dcount, ka, va, kb, vb)
    if (cls == 0) { k_wave_sort_kv<4><<<gw, 256, 0, st>>>(segs, dcount, ka, va, kb, vb); return; }
    switch (plan * 8 + cls) {
        case 1: S512_TKV(128, 4); break;
        case 2: S512_TKV(128, 8); break;
        case 3: S512_TKV(256, 8); break;
        case 4: S512_TKV(256, 16); break;
        case 8 + 1: S512_TKV(64, 8); break;
        case 8 + 2: S512_TKV(64, 16); break;
        case 8 + 3: S512_TKV(128, 16); break;
        case 8 + 4: S512_TKV(256, 16); break;
        case 16 + 1: S512_TKV(64, 8); break;
        case 16 + 2: S512_TKV(64, 16); break;
        case 16 + 3: S512_TKV(64, 32); break;
        case 16 + 4: S512_TKV(128, 32); break;
        case 16 + 5: S512_TKV(256, 32); break;
        default: S512_TKV(512, 16); break;
    }
#undef S512_TKV
}

// Key/value tiles with 64-bit keys (one shape per class).
template <class T>
static void launch_tile_kv64(int cls, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* ka,
                             int32_t* va, const T* kb, const int32_t* vb, hipStream_t st) {
    if (count == 0) return;
#define S512_TKV64(PP, EE) k_tile_sort_kv64<T, PP, EE><<<count, PP, 0, st>>>(segs, dcount, ka, va, kb, vb)
    switch (cls) {
        case 0:
        case 1: S512_TKV64(64, 8); break;
        case 2: S512_TKV64(64, 16); break;
        case 3: S512_TKV64(128, 16); break;
        case 4: S512_TKV64(256, 16); break;
        default: S512_TKV64(512, 16); break;
    }
#undef S512_TKV64
}

template <class T, bool KV>
struct TileCfg {
    // key/value tiles stop at 8192 pairs (LDS); plain keys of either width at 16384
    static const uint32_t* caps() { return KV ? kCapsKV : sizeof(T) == 8 ? kCaps64 : kCaps32; }
    static int nclass() { return KV ? 6 : 7; }
    static uint32_t maxcap() { return caps()[nclass() - 1]; }
    static int cls_of(uint64_t n) {
        for (int c = 0; c < nclass(); ++c)
            if (n <= caps()[c]) return c;
        return -1;
    }
};

template <class T, bool KV>
static void launch_tile(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* a,
                        int32_t* va, const T* b, const int32_t* vb, hipStream_t st, bool f64_fast = false) {
    if constexpr (KV && sizeof(T) == 8) launch_tile_kv64<T>(cls, segs, dcount, count, a, va, b, vb, st);
    else if constexpr (KV) launch_tile_kv(cls, plan, segs, dcount, count, a, va, b, vb, st);
    else if constexpr (std::is_same<T, double>::value) {
        if (f64_fast) launch_tile_f64_fast(cls, plan, segs, dcount, count, a, b, st);
        else launch_tile_keys<T>(cls, plan, segs, dcount, count, a, b, st);
    } else launch_tile_keys<T>(cls, plan, segs, dcount, count, a, b, st);
}

static inline int ceil_log2(uint64_t x) {
    int l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

// The tile launches and classes of the engine, for the one-pass levels.
template <class T, bool KV>
void launch_tile_list(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t count, T* a,
                      int32_t* va, const T* b, const int32_t* vb, hipStream_t st, bool f64_fast) {
    launch_tile<T, KV>(cls, plan, segs, dcount, count, a, va, b, vb, st, f64_fast);
}
template <class T, bool KV>
const uint32_t* tile_class_caps(int* ncls) {
    *ncls = TileCfg<T, KV>::nclass();
    return TileCfg<T, KV>::caps();
}
#define S512_TILE_LIST(T, KV)                                                                                    \
    template void launch_tile_list<T, KV>(int, int, const TileSeg*, const uint32_t*, uint32_t, T*, int32_t*,     \
                                          const T*, const int32_t*, hipStream_t, bool);                          \
    template const uint32_t* tile_class_caps<T, KV>(int*);
S512_TILE_LIST(int32_t, false)
S512_TILE_LIST(int32_t, true)
S512_TILE_LIST(double, false)
S512_TILE_LIST(int64_t, false)
S512_TILE_LIST(uint64_t, false)
S512_TILE_LIST(uint32_t, false)
#undef S512_TILE_LIST

// Tile classes of plain keys for the segmented sort.  For 8-byte keys they
// stop at 8192: a segment of 8193..16384 keys partitions there, which measured
// faster than the 16384-key tile (whose registers spill; the sort engine uses
// that tile only for the rare bucket just above 8192 keys).
template <class T>
const uint32_t* tile_caps(int* ncls) {
    if (sizeof(T) == 8) {
        *ncls = 6;
        return kCapsKV;
    }
    *ncls = TileCfg<T, false>::nclass();
    return TileCfg<T, false>::caps();
}
template <class T>
void launch_tile_fixed(int cls, int plan, const TileSeg* segs, const uint32_t* dcount, uint32_t grid_items, T* a,
                       const T* b, hipStream_t st) {
    launch_tile_keys<T>(cls, plan, segs, dcount, grid_items, a, b, st);
}
template const uint32_t* tile_caps<int32_t>(int*);
template const uint32_t* tile_caps<double>(int*);
template void launch_tile_fixed<int32_t>(int, int, const TileSeg*, const uint32_t*, uint32_t, int32_t*,
                                         const int32_t*, hipStream_t);
template void launch_tile_fixed<double>(int, int, const TileSeg*, const uint32_t*, uint32_t, double*, const double*,
                                        hipStream_t);

// ---------------------------------------------------------------- engine
// The partition level loop over an initial list of segments of the arrays
// a (buf 0) and b (buf 1), both n keys long (key/value: va, vb).  Synchronises
// the stream between levels; returns with every segment sorted into a.
template <class T, bool KV>
s512_status run_levels(T* a, int32_t* va, T* b, int32_t* vb, size_t n, std::vector<Seg> segs, hipStream_t st,
                       const EngineParams& prm, s512_stats* stats, Profiler& pf, DeviceState& ds,
                       const LearnedModelDev& model, LevelState& lstate) {
    using K = typename KeyTraits<T>::K;
    using TC = TileCfg<T, KV>;
    s512_status rc;
    const int ncls = TC::nclass();
    bool& use_learned = lstate.use_learned;
    const uint32_t lf1 = lstate.lf1;
    std::vector<PartSeg> ps;
    std::vector<ChunkRef> crs;
    const uint32_t smax = 8192;
    int level = 0;
    uint64_t tiles = 0;
    uint64_t keys_ph[S512_NPHASE] = {0};
    bool f64_fast = false;
    uint16_t* d_ids = nullptr;
    if (prm.ids) {
        if ((rc = ds.ids.reserve(n * sizeof(uint16_t), st))) return rc;
        d_ids = (uint16_t*)ds.ids.ptr;
    }
    while (!segs.empty()) {
        if (++level > prm.max_levels) {
            set_error("partition did not converge within %d levels", prm.max_levels);
            return S512_ERR_INTERNAL;
        }
        if (level == 2 && lstate.pending) {
            // the network trained beside the first level: its knots place the
            // second level, unless training failed (then sampled splitters)
            lstate.pending = false;
            S512_HIP(hipEventSynchronize(lstate.trained));
            const float loss = *lstate.loss_host;
            if (stats) stats->model_loss = loss;
            if (use_learned && (!std::isfinite(loss) || loss > lstate.max_loss)) {
                use_learned = false;
                if (stats) {
                    stats->learned_fallback = 1;
                    stats->learned_levels = 1;
                }
            }
            S512_HIP(hipStreamWaitEvent(st, lstate.trained, 0));
        }
        const uint32_t m = (uint32_t)segs.size();
        uint64_t total = 0;
        for (auto& s : segs) total += s.size;
        uint64_t chunk = (total + prm.target_chunks - 1) / prm.target_chunks;
        // chunks of 512 KiB or more: four times as many (each still >= 128 KiB).
        // With ~4 chunks per CU the count and scatter grids end on their
        // slowest CU; 4x finer chunks measured 2-3% faster for 2^26..2^28
        // fp64 keys and 2^28 int32 keys, while smaller chunks (2^24 fp64,
        // 2^26 int32) measured 4-5% slower.
        // The line-buffered scatter stores a (chunk, bucket) run's first and
        // last line partially: it takes chunks of LB_MIN_CHUNK keys or more.
        const bool lb = !KV && prm.scatter_shape == SCATTER_LB;
        if (lb) chunk = std::max<uint64_t>(chunk, LB_MIN_CHUNK);
        else if (chunk * (sizeof(T) + (KV ? sizeof(int32_t) : 0)) >= (512u << 10)) chunk = (chunk + 3) / 4;
        chunk = (chunk + SCAT_ROUND - 1) / SCAT_ROUND * SCAT_ROUND;
        chunk = std::max<uint64_t>(SCAT_ROUND, std::min<uint64_t>(chunk, prm.max_chunk));
        ps.resize(m);
        crs.clear();
        uint64_t histN = 0, sumB = 0, maxB = 0, sumF = 0, sumG = 0;
        uint32_t max_nchunks = 0;
        for (uint32_t i = 0; i < m; ++i) {
            PartSeg& p = ps[i];
            p.begin = segs[i].begin;
            p.size = segs[i].size;
            p.buf = segs[i].buf;
            int lf = ceil_log2((p.size + prm.target_leaf - 1) / prm.target_leaf);
            lf = std::max(1, std::min(lf, prm.max_logf));
            p.mode = 0;
            p.lshift = 0;
            p.lbase = 0;
            if (use_learned && (segs[i].aux & 1u)) {
                // the model's CDF table places every key: in the root level
                // (2^lf1 buckets), then inside each root bucket j (fan-out
                // from the bucket's size)
                p.mode = 2;
                if (level == 1) {
                    lf = (int)lf1;
                } else {
                    p.lshift = lf1;
                    p.lbase = segs[i].aux >> 1;
                }
            }
            p.logf = (uint32_t)lf;
            const uint32_t F = 1u << lf, B = 2 * F;
            p.nsample = p.mode ? 0u : std::min<uint32_t>(smax, prm.oversample * F);
            p.chunk = chunk;
            p.nchunks = (uint32_t)((p.size + chunk - 1) / chunk);
            max_nchunks = std::max(max_nchunks, p.nchunks);
            p.hist_base = histN;
            p.tree_base = sumF;
            p.bkt_base = sumB;
            p.guide_base = sumG;
            p.kmin = 0;
            p.shift = 0;
            p.flags = 0;
            p.sample_base = (uint64_t)i * smax;
            sumF += F;
            if (!p.mode) sumG += guide_words(F);
            histN += (uint64_t)p.nchunks * B;
            sumB += B;
            maxB = std::max<uint64_t>(maxB, B);
            for (uint32_t c = 0; c < p.nchunks; ++c) crs.push_back(ChunkRef{i, c});
        }
        // learned segments: after the guides, one copy of the root's CDF table
        // (level 1) or of the root buckets' local tables (level 2)
        const uint64_t tab_off = sumG;
        bool level_learned = false;
        for (auto& p : ps)
            if (p.mode) {
                p.guide_base = tab_off + (p.lshift ? (uint64_t)p.lbase * LT_WORDS : 0ull);
                level_learned = true;
            }
        if (level_learned) sumG += level == 1 ? CT_WORDS : ((size_t)LT_WORDS << lf1);
        const uint32_t tc = (uint32_t)crs.size();
        const uint64_t copy_cap = sumB + total / COPY_PIECE + 1;
        // ---- arena layout: the uploaded prefix first
        SizeCounter sc;
        sc.take<PartSeg>(m);
        sc.take<ChunkRef>(tc);
        sc.take<TileSeg>(m);
        sc.take<uint32_t>(32);
        sc.take<uint64_t>(16);
        const size_t up_bytes = sc.off;
        sc.take<K>((size_t)m * smax);
        sc.take<K>(sumF);
        sc.take<uint32_t>(sumG);
        sc.take<uint32_t>(histN);
        sc.take<uint64_t>(sumB);
        sc.take<uint64_t>(sumB);
        for (int c = 0; c < ncls; ++c) sc.take<TileSeg>(sumB);
        sc.take<TileSeg>(copy_cap);
        sc.take<Seg>(sumB);
        if ((rc = ds.arena.reserve(sc.off, st))) return rc;
        if ((rc = ds.pinned.reserve(up_bytes, st))) return rc;
        // pinned2: the level's counts (first 4 KiB), then the next level's segments
        const uint64_t pre_next = std::min<uint64_t>(sumB, 4096);  // segments read with the counts
        // (+ the learned root level's bucket sizes, for the model's rejection test)
        const size_t root_sizes_off = (4096 + std::max<size_t>(sumB, 1) * sizeof(Seg) + 63) & ~(size_t)63;
        const bool root_check = lstate.check_root && level == 1;
        if ((rc = ds.pinned2.reserve(root_sizes_off + (root_check ? sumB * sizeof(uint64_t) : 0), st))) return rc;

        Arena ar{(char*)ds.arena.ptr, 0}, hp{(char*)ds.pinned.ptr, 0};
        PartSeg* d_ps = ar.take<PartSeg>(m);
        ChunkRef* d_cr = ar.take<ChunkRef>(tc);
        TileSeg* d_sl = ar.take<TileSeg>(m);
        uint32_t* d_cnt = ar.take<uint32_t>(32);
        uint64_t* d_cnt64 = ar.take<uint64_t>(16);
        K* d_smp = ar.take<K>((size_t)m * smax);
        K* d_tree = ar.take<K>(sumF);
        uint32_t* d_guide = ar.take<uint32_t>(sumG);
        uint32_t* d_hist = ar.take<uint32_t>(histN);
        uint64_t* d_bstart = ar.take<uint64_t>(sumB);
        uint64_t* d_bsize = ar.take<uint64_t>(sumB);
        PlanLists pl;
        memset(&pl, 0, sizeof(pl));
        for (int c = 0; c < ncls; ++c) {
            pl.tile[c] = ar.take<TileSeg>(sumB);
            pl.tile_cap[c] = TC::caps()[c];
        }
        pl.nclass = (uint32_t)ncls;
        pl.tile_count = d_cnt;
        pl.tile_keys = d_cnt64;
        pl.copy = ar.take<TileSeg>(copy_cap);
        pl.copy_count = d_cnt + 8;
        pl.next = ar.take<Seg>(sumB);
        pl.next_count = d_cnt + 9;
        for (int c = 0; c < ncls; ++c) pl.cap[c] = (uint32_t)sumB;
        pl.cap[8] = (uint32_t)copy_cap;
        pl.cap[9] = (uint32_t)sumB;
        pl.overflow = d_cnt + 14;

        PartSeg* h_ps = hp.take<PartSeg>(m);
        ChunkRef* h_cr = hp.take<ChunkRef>(tc);
        TileSeg* h_sl = hp.take<TileSeg>(m);
        uint32_t* h_cnt = hp.take<uint32_t>(32);
        uint64_t* h_cnt64 = hp.take<uint64_t>(16);
        memcpy(h_ps, ps.data(), m * sizeof(PartSeg));
        memcpy(h_cr, crs.data(), tc * sizeof(ChunkRef));
        uint32_t nsplit = 0;  // sample lists of the segments partitioned by splitters
        for (uint32_t i = 0; i < m; ++i)
            if (!ps[i].mode) h_sl[nsplit++] = TileSeg{(uint64_t)i * smax, ps[i].nsample, 0};
        memset(h_cnt, 0, 32 * sizeof(uint32_t));
        memset(h_cnt64, 0, 16 * sizeof(uint64_t));
        h_cnt[10] = nsplit;
        h_cnt[12] = m;   // segments of this level
        h_cnt[13] = tc;  // chunks of this level
        S512_HIP(hipMemcpyAsync(ds.arena.ptr, ds.pinned.ptr, up_bytes, hipMemcpyHostToDevice, st));

        const uint64_t seed = prm.seed + 0x9E37ull * (uint64_t)level;
        bool any_split = false, any_learned = false;
        for (auto& p : ps) (p.mode ? any_learned : any_split) = true;
        if (any_learned && any_split) {
            // the count and scatter builds carry one classifier each
            set_error("level %d mixes learned and sampled segments", level);
            return S512_ERR_INTERNAL;
        }
        if (any_split) {
            pf.begin(S512_PHASE_SAMPLE);
            k_sample<T><<<m, 256, 0, st>>>(d_ps, d_cnt + 12, a, b, d_smp, smax, seed);
            S512_CHECK_LAUNCH();
            pf.end();
            pf.begin(S512_PHASE_SAMPLE_SORT);
            // the smallest tile that holds every segment's sample: a level of
            // many small segments (fan-out below 512) samples 16F < 8192 keys each
            uint32_t max_ns = 0;
            for (auto& p : ps)
                if (!p.mode) max_ns = std::max(max_ns, p.nsample);
            if (max_ns <= 1024) k_tile_sort<K, 64, 16><<<nsplit, 64, 0, st>>>(d_sl, d_cnt + 10, d_smp, d_smp);
            else if (max_ns <= 2048) k_tile_sort<K, 128, 16><<<nsplit, 128, 0, st>>>(d_sl, d_cnt + 10, d_smp, d_smp);
            else if (max_ns <= 4096) k_tile_sort<K, 256, 16><<<nsplit, 256, 0, st>>>(d_sl, d_cnt + 10, d_smp, d_smp);
            else k_tile_sort<K, 512, 16><<<nsplit, 512, 0, st>>>(d_sl, d_cnt + 10, d_smp, d_smp);
            S512_CHECK_LAUNCH();
            pf.end();
        }
        if (any_learned) {
            // the first level uses the sample's table, the second the network's
            pf.begin(S512_PHASE_SAMPLE);
            const size_t words = level == 1 ? CT_WORDS : ((size_t)LT_WORDS << lf1);
            S512_HIP(hipMemcpyAsync(d_guide + tab_off, level == 1 ? lstate.tabs.tab1 : lstate.tabs.tab2,
                                    words * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            pf.end();
        }
        if (any_split) {
            pf.begin(S512_PHASE_TREE);
            k_build_tree<K><<<m, FMAX, 0, st>>>(d_ps, d_cnt + 12, d_smp, smax, d_tree, d_guide);
            S512_CHECK_LAUNCH();
            pf.end();
        }
        pf.begin(S512_PHASE_COUNT);
        // NaN / -0.0 keys are looked for in the first level only (it sees every key)
        launch_count<T>(prm.count_shape, tc, st, d_ps, d_cr, d_cnt + 13, a, b, d_tree, d_guide, d_hist,
                        level == 1 ? d_cnt + 11 : nullptr, d_ids, any_learned);
        S512_CHECK_LAUNCH();
        pf.end();
        pf.begin(S512_PHASE_SCAN);
        // d_cnt[15], d_cnt[16]: sampled runs of the level and their keys, for the
        // full-line scatter's gate
        uint32_t* d_runs = (!KV && prm.scatter_shape == SCATTER_FL && d_ids) ? d_cnt + 15 : nullptr;
        if (max_nchunks >= 64)
            k_colscan<16><<<m * CS_NSLICE, 1024, 0, st>>>(d_ps, d_cnt + 12, d_hist, d_bsize, d_runs);
        else k_colscan<4><<<m * CS_NSLICE, 256, 0, st>>>(d_ps, d_cnt + 12, d_hist, d_bsize, d_runs);
        S512_CHECK_LAUNCH();
        k_bucket_scan<<<m, BMAX, 0, st>>>(d_ps, d_cnt + 12, d_bsize, d_bstart);
        S512_CHECK_LAUNCH();
        pf.end();
        pf.begin(S512_PHASE_SCATTER);
        launch_scatter<T, KV>(prm.scatter_shape, d_ids != nullptr, tc, st, d_ps, d_cr, a, b, va, vb, d_tree,
                              d_guide, d_hist, d_bstart, d_cnt + 13, d_ids, any_learned, d_runs);
        S512_CHECK_LAUNCH();
        pf.end();
        pf.begin(S512_PHASE_PLAN);
        k_plan<<<dim3(m, (uint32_t)((maxB + 255) / 256)), 256, 0, st>>>(d_ps, d_cnt + 12, d_bstart, d_bsize, pl);
        S512_CHECK_LAUNCH();
        pf.end();
        keys_ph[S512_PHASE_COUNT] += total;
        keys_ph[S512_PHASE_SCATTER] += total;
        uint32_t* h_out = (uint32_t*)ds.pinned2.ptr;
        Seg* h_next = (Seg*)((char*)ds.pinned2.ptr + 4096);
        S512_HIP(hipMemcpyAsync(h_out, d_cnt, 32 * sizeof(uint32_t) + 256, hipMemcpyDeviceToHost, st));
        // the first segments of the next level come back with the counts: one
        // host round trip per level instead of two
        S512_HIP(hipMemcpyAsync(h_next, pl.next, pre_next * sizeof(Seg), hipMemcpyDeviceToHost, st));
        uint64_t* h_sizes = (uint64_t*)((char*)ds.pinned2.ptr + root_sizes_off);
        if (root_check)
            S512_HIP(hipMemcpyAsync(h_sizes, d_bsize, sumB * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
        uint32_t counts[32];
        uint64_t counts64[16];
        memcpy(counts, h_out, sizeof(counts));
        memcpy(counts64, (char*)h_out + ((char*)d_cnt64 - (char*)d_cnt), sizeof(counts64));
        if (counts[14]) {
            set_error("plan lists overflowed at level %d", level);
            return S512_ERR_INTERNAL;
        }
        if (root_check) {
            // the learned model's rejection test, now that the stream has
            // passed its flags and the root level's counts: a degenerate
            // sample, repeated values it cannot split, or an ordinary bucket
            // far above the mean.  A rejected model leaves the next level to
            // sampled splitters; the root level's buckets are ordered key
            // ranges all the same.
            lstate.check_root = false;
            const LearnedModel& md = *lstate.model;
            uint64_t mx = 0;
            for (uint64_t j = 0; j < sumB; j += 2) mx = std::max(mx, h_sizes[j]);
            const double mean = (double)n / (double)(1u << lf1);
            const bool degenerate = md.flags_host[1] <= md.flags_host[0];
            if (stats) stats->learned_max_bucket_ratio = (float)((double)mx / mean);
            if (learned_model_rejected(md) || (double)mx > prm.imbalance * mean) {
                use_learned = false;
                if (stats && !degenerate) stats->learned_fallback = 1;
                if (stats) stats->learned_levels = 1;
            } else if (stats) {
                stats->learned_levels = 2;
            }
        }
        if (any_learned && level == 2 && stats) {
            uint64_t lk = 0, lb = 0;  // keys and buckets of the model's second level
            for (auto& p : ps)
                if (p.mode) {
                    lk += p.size;
                    lb += 1ull << p.logf;
                }
            if (lk) stats->learned_leaf_max_ratio = (float)((double)counts64[10] * (double)lb / (double)lk);
        }
        // level 1 saw every key: fp64 leaves may use min/max when no key is NaN or -0.0
        if (level == 1) f64_fast = !KV && std::is_same<T, double>::value && prm.f64_minmax && counts[11] == 0;
        if (counts[8]) {
            pf.begin(S512_PHASE_COPY);
            k_copy_back<T, KV><<<std::min<uint32_t>(counts[8], 16384), 256, 0, st>>>(pl.copy, pl.copy_count, a, b, va,
                                                                                     vb);
            S512_CHECK_LAUNCH();
            pf.end();
            keys_ph[S512_PHASE_COPY] += counts64[8];
        }
        for (int c = 0; c < ncls; ++c) {
            if (!counts[c]) continue;
            pf.begin(S512_PHASE_TILE);
            launch_tile<T, KV>(c, prm.tile_plan, pl.tile[c], d_cnt + c, counts[c], a, va, b, vb, st, f64_fast);
            S512_CHECK_LAUNCH();
            pf.end();
            tiles += counts[c];
            keys_ph[S512_PHASE_TILE] += counts64[c];
        }
        const uint32_t nn = counts[9];
        segs.clear();
        if (nn) {
            if (nn > sumB) {
                set_error("plan produced %u segments for %llu buckets", nn, (unsigned long long)sumB);
                return S512_ERR_INTERNAL;
            }
            if (nn > pre_next) {
                S512_HIP(hipMemcpyAsync(h_next + pre_next, pl.next + pre_next, (nn - pre_next) * sizeof(Seg),
                                        hipMemcpyDeviceToHost, st));
                S512_HIP(hipStreamSynchronize(st));
            }
            segs.assign(h_next, h_next + nn);
        }
    }
    S512_HIP(hipStreamSynchronize(st));
    if (stats) {
        stats->levels = level;
        stats->classifier = use_learned ? 1 : 0;
        stats->tiles = tiles;
        memcpy(stats->keys, keys_ph, sizeof(keys_ph));
        pf.collect(stats->kernel_ms, stats->launches);
    }
    return S512_OK;
}

// Arrays of 2^32 keys or more: the partition levels keep segment-relative
// offsets in 32 bits, so such an array is sorted as pieces of SORT_PIECE keys
// (each one ordinary engine sort), which merge-path rounds then merge.
constexpr uint64_t SORT_PIECE = 1ull << 31;

template <class T>
static s512_status sort_pieces_locked(T* a, size_t n, hipStream_t st, const EngineParams& prm, s512_stats* stats,
                                      bool profile, DeviceState& ds) {
    s512_stats ps;
    for (uint64_t b = 0; b < n; b += SORT_PIECE) {
        const uint64_t len = std::min<uint64_t>(SORT_PIECE, n - b);
        s512_status rc = sort_engine_locked<T, false>(a + b, nullptr, len, st, prm, stats ? &ps : nullptr, profile, ds);
        if (rc) return rc;
        if (!stats) continue;
        if (b == 0) {
            *stats = ps;  // the first piece's model and fallback figures
            continue;
        }
        stats->levels = std::max(stats->levels, ps.levels);
        stats->tiles += ps.tiles;
        for (int j = 0; j < S512_NPHASE; ++j) {
            stats->kernel_ms[j] += ps.kernel_ms[j];
            stats->launches[j] += ps.launches[j];
            stats->keys[j] += ps.keys[j];
        }
    }
    return merge_runs_locked<T>(a, n, SORT_PIECE, st, ds);
}

template <class T, bool KV>
s512_status sort_engine(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm, s512_stats* stats,
                        bool profile) {
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n <= 1) return S512_OK;
    if (KV && n >= (1ull << 32)) {
        set_error("n = %zu: key/value sorts take fewer than 2^32 pairs", n);
        return S512_ERR_TOO_LARGE;
    }
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    if constexpr (!KV) {
        if (n >= (1ull << 32)) return sort_pieces_locked<T>(a, n, st, prm, stats, profile, ds);
    }
    return sort_engine_locked<T, KV>(a, va, n, st, prm, stats, profile, ds);
}

template <class T, bool KV>
s512_status sort_engine_locked(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm,
                               s512_stats* stats, bool profile, DeviceState& ds) {
    using TC = TileCfg<T, KV>;
    if (n <= 1) return S512_OK;
    s512_status rc;
    Profiler pf;
    pf.on = profile && stats;
    pf.st = st;
    pf.ds = &ds;

    // ---- one tile: no partition level at all
    if (n <= TC::maxcap()) {
        SizeCounter sc;
        sc.take<TileSeg>(1);
        sc.take<uint32_t>(1);
        if ((rc = ds.arena.reserve(sc.off, st))) return rc;
        if ((rc = ds.pinned.reserve(sc.off, st))) return rc;
        Arena ar{(char*)ds.arena.ptr, 0}, hp{(char*)ds.pinned.ptr, 0};
        TileSeg* dseg = ar.take<TileSeg>(1);
        uint32_t* dcnt = ar.take<uint32_t>(1);
        TileSeg* hseg = hp.take<TileSeg>(1);
        uint32_t* hcnt = hp.take<uint32_t>(1);
        *hseg = TileSeg{0, (uint32_t)n, 0};
        *hcnt = 1;
        S512_HIP(hipMemcpyAsync(ds.arena.ptr, ds.pinned.ptr, sc.off, hipMemcpyHostToDevice, st));
        pf.begin(S512_PHASE_TILE);
        launch_tile<T, KV>(TC::cls_of(n), prm.tile_plan, dseg, dcnt, 1, a, va, a, va, st);
        S512_CHECK_LAUNCH();
        pf.end();
        S512_HIP(hipStreamSynchronize(st));  // pinned staging is reused by the next call
        if (stats) {
            stats->tiles = 1;
            stats->keys[S512_PHASE_TILE] = n;
            pf.collect(stats->kernel_ms, stats->launches);
        }
        return S512_OK;
    }

    // ---- learned partitioner: model of the key CDF for the first two levels
    LearnedModel model;
    memset(&model, 0, sizeof(model));
    bool use_learned = false;
    uint32_t lf1 = 0;
    // every return waits for a training still running on the training stream
    // (it writes the model buffers the next call reuses)
    struct TrainWait {
        hipEvent_t ev = nullptr;
        ~TrainWait() {
            if (ev) (void)hipEventSynchronize(ev);
        }
    } train_wait;
    if constexpr (!KV && (std::is_same<T, double>::value || std::is_same<T, int32_t>::value)) {
        if (prm.learned) {
            if ((rc = build_learned_model<T>(a, n, st, ds, prm, &model, stats))) return rc;
            if (model.training) train_wait.ev = model.trained;
            // the first level runs on the sample's table at once; whether the
            // model is rejected (a degenerate sample, repeated values it cannot
            // split, a bucket far above the mean) is read with that level's
            // counts, and a rejected model leaves the second level to sampled
            // splitters (a network that fails to train as well)
            use_learned = true;
            lf1 = model.lf1;
        }
    }
    // ---- one-pass levels, when asked for (not for the learned model's levels,
    // nor 64-bit keys with values); they hand back with nothing changed when
    // the root level's regions overflow, and the two-pass levels below run
    if constexpr (!(KV && sizeof(T) == 8)) {
        const bool want = prm.engine == S512_ENGINE_ONE_PASS && (uint64_t)n >= prm.onepass_min;
        if (want && !use_learned) {
            bool done = false;
            if ((rc = onepass_engine<T, KV>(a, va, n, st, prm, stats, pf, ds, &done))) return rc;
            if (done) return S512_OK;
        }
    }
    // ---- scratch arrays (same size as the input)
    const size_t dbytes = n * sizeof(T) + (KV ? n * sizeof(int32_t) : 0);
    if ((rc = ds.data.reserve(dbytes, st))) return rc;
    T* b = (T*)ds.data.ptr;
    int32_t* vb = KV ? (int32_t*)((char*)ds.data.ptr + n * sizeof(T)) : nullptr;

    std::vector<Seg> segs{Seg{0, n, 0, use_learned ? 1u : 0u}};
    LevelState ls;
    ls.use_learned = use_learned;
    ls.model = use_learned ? &model : nullptr;
    ls.check_root = use_learned;
    ls.lf1 = lf1;
    ls.tabs = model.tabs;
    ls.trained = model.trained;
    ls.loss_host = model.loss_host;
    ls.max_loss = prm.max_loss;
    ls.pending = use_learned && model.training;
    rc = run_levels<T, KV>(a, va, b, vb, n, segs, st, prm, stats, pf, ds, model.tabs, ls);
    if (stats) stats->engine = S512_ENGINE_TWO_PASS;
    if (model.ev0) {
        // the sample and its table, timed for the statistics (the stream has
        // passed them: every level ends with a synchronisation)
        float ms = 0.f;
        if (rc == S512_OK && stats && hipEventElapsedTime(&ms, model.ev0, model.ev1) == hipSuccess) {
            stats->kernel_ms[S512_PHASE_TRAIN] += ms;
            stats->launches[S512_PHASE_TRAIN] += 1;
            stats->keys[S512_PHASE_TRAIN] += model.ns;
        }
        (void)hipEventDestroy(model.ev0);
        (void)hipEventDestroy(model.ev1);
    }
    if (rc == S512_OK && stats && model.training && ls.pending) {
        // the sort finished within one level: report the loss all the same
        (void)hipEventSynchronize(model.trained);
        stats->model_loss = *model.loss_host;
    }
    return rc;
}

template s512_status run_levels<int32_t, false>(int32_t*, int32_t*, int32_t*, int32_t*, size_t, std::vector<Seg>,
                                                hipStream_t, const EngineParams&, s512_stats*, Profiler&, DeviceState&,
                                                const LearnedModelDev&, LevelState&);
template s512_status run_levels<double, false>(double*, int32_t*, double*, int32_t*, size_t, std::vector<Seg>,
                                               hipStream_t, const EngineParams&, s512_stats*, Profiler&, DeviceState&,
                                               const LearnedModelDev&, LevelState&);
template s512_status sort_engine<int32_t, false>(int32_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine_locked<uint64_t, false>(uint64_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool, DeviceState&);
template s512_status sort_engine_locked<uint32_t, false>(uint32_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool, DeviceState&);
template s512_status sort_engine<double, false>(double*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<int32_t, true>(int32_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<double, true>(double*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<int64_t, true>(int64_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<uint64_t, true>(uint64_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<int64_t, false>(int64_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<uint64_t, false>(uint64_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);
template s512_status sort_engine<uint32_t, false>(uint32_t*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, bool);

EngineParams default_params(bool wide) {
    EngineParams p;
    p.target_leaf = 1024;  // fastest measured for int32 and double (tiles of ~1024 keys)
    p.oversample = 16;
    p.target_chunks = 1024;
    p.max_chunk = 1u << 20;
    p.max_levels = 48;
    p.max_logf = FMAX_LOG;
    p.seed = 0x5EEDull;
    p.f64_minmax = true;
    p.learned = false;
    p.sample_permille = 10;
    p.train_steps = 10;  // over before the first level's scatter starts (30 ran into it)
    p.learn_rate = 2e-3f;
    p.max_loss = 1e-3f;  // RMS CDF error ~3%: the network did not train
    p.imbalance = 4.0f;
    p.scatter_shape = 0;
    p.ids = true;
    p.count_shape = 2;  // 256 threads x 8 keys, 8 classified together (fastest measured)
    p.tile_plan = 2;    // 32 keys per thread above 1024-key tiles (fastest measured)
    p.engine = S512_ENGINE_AUTO;
    p.onepass_min = 1;  // any array with a partition level
    p.region_z = 5.0f;  // P(a region overflows) ~ 3e-7 per (ordinary, equality) pair
    return p;
}

EngineParams params_from(const s512_options* opt, bool wide) {
    EngineParams p = default_params(wide);
    if (!opt) return p;
    if (opt->target_leaf) p.target_leaf = opt->target_leaf;
    if (opt->oversample) p.oversample = opt->oversample;
    if (opt->max_fanout_log2) p.max_logf = std::max(1, std::min<int>((int)opt->max_fanout_log2, FMAX_LOG));
    if (opt->seed) p.seed = opt->seed;
    p.learned = opt->classifier == S512_CLASSIFY_LEARNED;
    if (opt->sample_permille) p.sample_permille = opt->sample_permille;
    if (opt->train_steps) p.train_steps = opt->train_steps;
    if (opt->imbalance_pct) p.imbalance = (float)opt->imbalance_pct / 100.0f;
    if (opt->scatter_variant) {
        const uint32_t v = opt->scatter_variant - 1;
        if (v == 2 * SCATTER_SHAPES + 2) {  // 11: full lines, stored ids
            p.scatter_shape = SCATTER_FL;
            p.ids = true;
        } else if (v >= 2 * SCATTER_SHAPES) {  // 9, 10: line-buffered, classified again / stored ids
            p.scatter_shape = SCATTER_LB;
            p.ids = v == 2 * SCATTER_SHAPES + 1;
        } else {
            p.scatter_shape = (int)(v % SCATTER_SHAPES);
            p.ids = v < SCATTER_SHAPES;
        }
    }
    if (opt->count_variant) p.count_shape = (int)opt->count_variant - 1;
    if (opt->tile_variant) p.tile_plan = (int)opt->tile_variant - 1;
    p.engine = (int)opt->engine;
    if (opt->region_slack) p.region_z = (float)opt->region_slack / 10.0f;
    return p;
}

// ---------------------------------------------------------------- tuned defaults
// One autotuned entry per workload kind and size class: n in [2^lg_lo, 2^lg_hi).
struct TuneEntry {
    int kind, lg_lo, lg_hi;
    int count_shape, scatter_shape, ids, tile_plan;
    uint32_t target_leaf;
    int max_logf;
};
#include "s512_tune_table.inc"  // kTune[]: generated by avx-512-sort_amd/s512_tune.py

static void apply_tuned(EngineParams& p, int kind, uint64_t n) {
    int lg = 0;
    while (lg < 63 && (1ull << (lg + 1)) <= n) ++lg;  // floor(log2 n)
    for (const TuneEntry& e : kTune) {
        if (e.kind != kind || lg < e.lg_lo || lg >= e.lg_hi) continue;
        p.count_shape = e.count_shape;
        p.scatter_shape = e.scatter_shape;
        p.ids = e.ids != 0;
        p.tile_plan = e.tile_plan;
        p.target_leaf = e.target_leaf;
        p.max_logf = e.max_logf;
        return;
    }
}

EngineParams params_for(const s512_options* opt, int kind, uint64_t n) {
    EngineParams p = default_params(kind == TUNE_F64 || kind == TUNE_KV || kind == TUNE_SEG_F64);
    apply_tuned(p, kind, n);
    if (!opt) return p;
    const EngineParams d = p;
    p = params_from(opt, false);
    // params_from starts from the untuned defaults: keep the tuned value of
    // every knob the caller left at zero
    if (!opt->target_leaf) p.target_leaf = d.target_leaf;
    if (!opt->max_fanout_log2) p.max_logf = d.max_logf;
    if (!opt->scatter_variant) {
        p.scatter_shape = d.scatter_shape;
        p.ids = d.ids;
    }
    if (!opt->count_variant) p.count_shape = d.count_shape;
    if (!opt->tile_variant) p.tile_plan = d.tile_plan;
    return p;
}

}  // namespace s512

using namespace s512;

static s512_status check_opts(const s512_options* opt) {
    if (opt && opt->oversample > 512) {
        set_error("oversample %u too large (max 512)", opt->oversample);
        return S512_ERR_INVALID;
    }
    if (opt && opt->classifier > S512_CLASSIFY_LEARNED) {
        set_error("unknown classifier %u", opt->classifier);
        return S512_ERR_INVALID;
    }
    if (opt && (opt->sample_permille > 1000 || opt->train_steps > 100000)) {
        set_error("sample_permille (<= 1000) or train_steps (<= 100000) out of range");
        return S512_ERR_INVALID;
    }
    if (opt && opt->engine > S512_ENGINE_ONE_PASS) {
        set_error("engine %u out of range (max %d)", opt->engine, S512_ENGINE_ONE_PASS);
        return S512_ERR_INVALID;
    }
    if (opt && opt->tile_variant > TILE_PLANS) {
        set_error("tile_variant %u out of range (max %d)", opt->tile_variant, TILE_PLANS);
        return S512_ERR_INVALID;
    }
    if (opt && opt->count_variant > COUNT_SHAPES) {
        set_error("count_variant %u out of range (max %d)", opt->count_variant, COUNT_SHAPES);
        return S512_ERR_INVALID;
    }
    if (opt && opt->scatter_variant > 2 * SCATTER_SHAPES + 3) {
        set_error("scatter_variant %u out of range (max %d)", opt->scatter_variant, 2 * SCATTER_SHAPES + 3);
        return S512_ERR_INVALID;
    }
    if (opt && opt->target_leaf && opt->target_leaf < 64) {
        set_error("target_leaf %u too small (min 64)", opt->target_leaf);
        return S512_ERR_INVALID;
    }
    return S512_OK;
}

extern "C" s512_status s512_sort_i32_dev_ex(int32_t* d_a, size_t n, const s512_options* opt, s512_stats* stats,
                                            void* stream) {
    if (n > 1 && !d_a) { set_error("null array"); return S512_ERR_INVALID; }
    if (s512_status rc = check_opts(opt)) return rc;
    return sort_engine<int32_t, false>(d_a, nullptr, n, (hipStream_t)stream, params_for(opt, TUNE_I32, n), stats,
                                       opt && opt->profile);
}
extern "C" s512_status s512_sort_f64_dev_ex(double* d_a, size_t n, const s512_options* opt, s512_stats* stats,
                                            void* stream) {
    if (n > 1 && !d_a) { set_error("null array"); return S512_ERR_INVALID; }
    if (s512_status rc = check_opts(opt)) return rc;
    return sort_engine<double, false>(d_a, nullptr, n, (hipStream_t)stream, params_for(opt, TUNE_F64, n), stats,
                                      opt && opt->profile);
}
extern "C" s512_status s512kv_sort_i32_dev_ex(int32_t* d_k, int32_t* d_v, size_t n, const s512_options* opt,
                                              s512_stats* stats, void* stream) {
    if (n > 1 && (!d_k || !d_v)) { set_error("null array"); return S512_ERR_INVALID; }
    if (s512_status rc = check_opts(opt)) return rc;
    return sort_engine<int32_t, true>(d_k, d_v, n, (hipStream_t)stream, params_for(opt, TUNE_KV, n), stats,
                                      opt && opt->profile);
}
extern "C" s512_status s512_sort_i32_dev(int32_t* d_a, size_t n, void* stream) {
    return s512_sort_i32_dev_ex(d_a, n, nullptr, nullptr, stream);
}
extern "C" s512_status s512_sort_f64_dev(double* d_a, size_t n, void* stream) {
    return s512_sort_f64_dev_ex(d_a, n, nullptr, nullptr, stream);
}
extern "C" s512_status s512_sort_i32(int32_t* a, size_t n) {
    if (n <= 1) return S512_OK;
    return with_staged_array(a, n, [n](int32_t* d, hipStream_t st) { return s512_sort_i32_dev(d, n, st); });
}
extern "C" s512_status s512_sort_f64(double* a, size_t n) {
    if (n <= 1) return S512_OK;
    return with_staged_array(a, n, [n](double* d, hipStream_t st) { return s512_sort_f64_dev(d, n, st); });
}
// 64-bit and unsigned keys (no reference counterpart): the same engine, tuned
// as the key width's reference type.
#define S512_WIDE_KEYS(NAME, T, KIND)                                                                         \
    extern "C" s512_status s512_sort_##NAME##_dev_ex(T* d_a, size_t n, const s512_options* opt, s512_stats* stats, \
                                                     void* stream) {                                          \
        if (n > 1 && !d_a) { set_error("null array"); return S512_ERR_INVALID; }                             \
        if (s512_status rc = check_opts(opt)) return rc;                                                      \
        if (opt && opt->classifier == S512_CLASSIFY_LEARNED) {                                                \
            set_error("the learned classifier covers int32 and double keys");                                 \
            return S512_ERR_INVALID;                                                                          \
        }                                                                                                     \
        return sort_engine<T, false>(d_a, nullptr, n, (hipStream_t)stream, params_for(opt, KIND, n), stats,   \
                                     opt && opt->profile);                                                    \
    }                                                                                                         \
    extern "C" s512_status s512_sort_##NAME##_dev(T* d_a, size_t n, void* stream) {                          \
        return s512_sort_##NAME##_dev_ex(d_a, n, nullptr, nullptr, stream);                                   \
    }                                                                                                         \
    extern "C" s512_status s512_sort_##NAME(T* a, size_t n) {                                                \
        if (n <= 1) return S512_OK;                                                                           \
        return with_staged_array(a, n, [n](T* d, hipStream_t st) { return s512_sort_##NAME##_dev(d, n, st); }); \
    }
S512_WIDE_KEYS(i64, int64_t, TUNE_F64)
S512_WIDE_KEYS(u64, uint64_t, TUNE_F64)
S512_WIDE_KEYS(u32, uint32_t, TUNE_I32)
#undef S512_WIDE_KEYS

// Key/value sorts with 64-bit keys and argsort (no reference counterpart;
// Sort512kv covers int keys).  64-bit keys take the fp64 tuned table with the
// scatter shapes whose LDS holds 8-byte keys with their values.
static EngineParams kv64_params(size_t n) {
    EngineParams p = params_for(nullptr, TUNE_F64, n);
    if (p.scatter_shape == 3) p.scatter_shape = 0;  // no deferred-line shape for wide pairs
    return p;
}
#define S512_KV64(NAME, T)                                                                                  \
    extern "C" s512_status s512kv_sort_##NAME##_dev(T* d_k, int32_t* d_v, size_t n, void* stream) {        \
        if (n > 1 && (!d_k || !d_v)) { set_error("null array"); return S512_ERR_INVALID; }                 \
        return sort_engine<T, true>(d_k, d_v, n, (hipStream_t)stream, kv64_params(n), nullptr, false);      \
    }                                                                                                       \
    extern "C" s512_status s512kv_sort_##NAME(T* k, int32_t* v, size_t n) {                                 \
        if (n <= 1) return S512_OK;                                                                         \
        return with_staged_pair(k, v, n, [n](T* dk, int32_t* dv, hipStream_t st) {                         \
            return s512kv_sort_##NAME##_dev(dk, dv, n, st);                                                 \
        });                                                                                                 \
    }
S512_KV64(f64, double)
S512_KV64(i64, int64_t)
S512_KV64(u64, uint64_t)
#undef S512_KV64

__global__ void k_iota(int32_t* idx, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        idx[i] = (int32_t)i;
}

// idx[i] = position in keys of the i-th smallest key.  The keys are copied
// to the per-device staging buffer (slot 0) and sorted there with the
// positions as values, so the caller's keys are left unchanged.
template <class T, bool KV_WIDE>
static s512_status argsort_any(const T* keys, int32_t* idx, size_t n, hipStream_t st) {
    if (n == 0) return S512_OK;
    if (!keys || !idx) { set_error("null array"); return S512_ERR_INVALID; }
    if (n > 0x7FFFFFFFull) { set_error("argsort of %zu keys: positions are int32", n); return S512_ERR_TOO_LARGE; }
    const bool kdev = is_device_pointer(keys), idev = is_device_pointer(idx);
    HostStage hs;
    void* kc = nullptr;
    if (s512_status rc = hs.in(0, keys, n * sizeof(T), &kc, !kdev)) return rc;
    if (kdev) S512_HIP(hipMemcpyAsync(kc, keys, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    void* di = idx;
    if (!idev)
        if (s512_status rc = hs.in(1, nullptr, n * sizeof(int32_t), &di, false)) return rc;
    k_iota<<<(unsigned)std::min<size_t>((n + 255) / 256, 65536), 256, 0, st>>>((int32_t*)di, n);
    S512_CHECK_LAUNCH();
    s512_status rc;
    if constexpr (KV_WIDE) rc = sort_engine<T, true>((T*)kc, (int32_t*)di, n, st, kv64_params(n), nullptr, false);
    else rc = sort_engine<T, true>((T*)kc, (int32_t*)di, n, st, params_for(nullptr, TUNE_KV, n), nullptr, false);
    if (rc) return rc;
    S512_HIP(hipStreamSynchronize(st));  // the staging buffer is reused by the next call
    if (!idev) return hs.out(1, idx, n * sizeof(int32_t));
    return S512_OK;
}
// Key/value sorts with an 8-byte value payload: the keys are sorted with their
// int32 positions (the kv engine), then the values are gathered through that
// permutation -- about the HBM traffic of carrying 8-byte values through every
// partition pass, with the tested 4-byte-value machinery.
__global__ void k_gather64(int64_t* out, const int64_t* in, const int32_t* idx, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[idx[i]];
}

template <class T, bool KV_WIDE>
static s512_status kv_sort_v64(T* keys, int64_t* vals, size_t n, hipStream_t st) {
    if (n <= 1) return S512_OK;
    if (!keys || !vals) { set_error("null array"); return S512_ERR_INVALID; }
    if (n > 0x7FFFFFFFull) { set_error("%zu pairs: the permutation is int32", n); return S512_ERR_TOO_LARGE; }
    const bool kdev = is_device_pointer(keys), vdev = is_device_pointer(vals);
    HostStage hs;
    void* dk = keys;
    void* dv = vals;
    if (!kdev)
        if (s512_status rc = hs.in(0, keys, n * sizeof(T), &dk)) return rc;
    if (!vdev)
        if (s512_status rc = hs.in(1, vals, n * sizeof(int64_t), &dv)) return rc;
    const size_t ib = (n * sizeof(int32_t) + 255) & ~(size_t)255;
    void* ws = nullptr;  // positions, then the gathered values
    if (s512_status rc = hs.in(2, nullptr, ib + n * sizeof(int64_t), &ws, false)) return rc;
    int32_t* idx = (int32_t*)ws;
    int64_t* tmp = (int64_t*)((char*)ws + ib);
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    k_iota<<<grid, 256, 0, st>>>(idx, n);
    S512_CHECK_LAUNCH();
    s512_status rc;
    if constexpr (KV_WIDE) rc = sort_engine<T, true>((T*)dk, idx, n, st, kv64_params(n), nullptr, false);
    else rc = sort_engine<T, true>((T*)dk, idx, n, st, params_for(nullptr, TUNE_KV, n), nullptr, false);
    if (rc) return rc;
    k_gather64<<<grid, 256, 0, st>>>(tmp, (const int64_t*)dv, idx, n);
    S512_CHECK_LAUNCH();
    S512_HIP(hipMemcpyAsync(dv, tmp, n * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    S512_HIP(hipStreamSynchronize(st));  // the staging buffers are reused by the next call
    if (!kdev)
        if (s512_status rc2 = hs.out(0, keys, n * sizeof(T))) return rc2;
    if (!vdev)
        if (s512_status rc2 = hs.out(1, vals, n * sizeof(int64_t))) return rc2;
    return S512_OK;
}
extern "C" s512_status s512kv_sort_i32_v64(int32_t* k, int64_t* v, size_t n, void* stream) {
    return kv_sort_v64<int32_t, false>(k, v, n, (hipStream_t)stream);
}
extern "C" s512_status s512kv_sort_f64_v64(double* k, int64_t* v, size_t n, void* stream) {
    return kv_sort_v64<double, true>(k, v, n, (hipStream_t)stream);
}
extern "C" s512_status s512kv_sort_i64_v64(int64_t* k, int64_t* v, size_t n, void* stream) {
    return kv_sort_v64<int64_t, true>(k, v, n, (hipStream_t)stream);
}
extern "C" s512_status s512kv_sort_u64_v64(uint64_t* k, int64_t* v, size_t n, void* stream) {
    return kv_sort_v64<uint64_t, true>(k, v, n, (hipStream_t)stream);
}

extern "C" s512_status s512_argsort_i32(const int32_t* keys, int32_t* idx, size_t n, void* stream) {
    return argsort_any<int32_t, false>(keys, idx, n, (hipStream_t)stream);
}
extern "C" s512_status s512_argsort_f64(const double* keys, int32_t* idx, size_t n, void* stream) {
    return argsort_any<double, true>(keys, idx, n, (hipStream_t)stream);
}
extern "C" s512_status s512_argsort_i64(const int64_t* keys, int32_t* idx, size_t n, void* stream) {
    return argsort_any<int64_t, true>(keys, idx, n, (hipStream_t)stream);
}
extern "C" s512_status s512_argsort_u64(const uint64_t* keys, int32_t* idx, size_t n, void* stream) {
    return argsort_any<uint64_t, true>(keys, idx, n, (hipStream_t)stream);
}

extern "C" s512_status s512kv_sort_i32_dev(int32_t* d_k, int32_t* d_v, size_t n, void* stream) {
    return s512kv_sort_i32_dev_ex(d_k, d_v, n, nullptr, nullptr, stream);
}
extern "C" s512_status s512kv_sort_i32(int32_t* k, int32_t* v, size_t n) {
    if (n <= 1) return S512_OK;
    if (!k || !v) { set_error("null array"); return S512_ERR_INVALID; }
    return with_staged_pair(k, v, n, [n](int32_t* dk, int32_t* dv, hipStream_t st) {
        return s512kv_sort_i32_dev(dk, dv, n, st);
    });
}

// SmallSort16V (sort512.hpp:2255 double <= 128 keys, 5312 int <= 256 keys;
// sort512kv.hpp:5235 pairs <= 256): one wave sorts the array in registers.
// Lengths above the reference's limit are rejected (undefined there).
extern "C" s512_status s512_smallsort16v_i32(int32_t* a, size_t len) {
    if (len > 256) { set_error("SmallSort16V(int) takes at most 256 keys, got %zu", len); return S512_ERR_INVALID; }
    return s512_sort_i32(a, len);
}
extern "C" s512_status s512_smallsort16v_f64(double* a, size_t len) {
    if (len > 128) { set_error("SmallSort16V(double) takes at most 128 keys, got %zu", len); return S512_ERR_INVALID; }
    return s512_sort_f64(a, len);
}
extern "C" s512_status s512kv_smallsort16v_i32(int32_t* k, int32_t* v, size_t len) {
    if (len > 256) { set_error("SmallSort16V(int,int) takes at most 256 pairs, got %zu", len); return S512_ERR_INVALID; }
    return s512kv_sort_i32(k, v, len);
}
