// s512_onepass.hip -- host side of the one-pass partition levels
// (s512_onepass.hpp): Sort512::Sort (sort512.hpp:5960) and Sort512kv::Sort
// (sort512kv.hpp:6132) on large arrays.
//
//   root level     A -> B.  All chunks resident at once (small arrays): exact,
//                  final positions in B.  Otherwise gapped: pair regions in B
//                  sized from a capacity sample of up to 2^18 keys (B holds
//                  about 1.23 n keys).  An overflowing region leaves A as it
//                  was and the caller sorts with the two-pass engine instead.
//   routing        equality buckets -> A (written from their first key);
//                  buckets up to a tile -> tile sorts into A; larger buckets
//                  -> the second level
//   second level   B -> A, exact, final positions, every segment whose chunks
//                  all fit on the GPU at once
//   leftovers      buckets still above a tile after the second level, and
//                  second-level segments too large for the exact mode: the
//                  multi-level engine (run_levels) in A with B as its scratch
#include "s512_onepass.hpp"
#include "s512_engine_internal.hpp"
#include <algorithm>
#include <cmath>
#include <type_traits>

#ifndef S512_OP_NB
#define S512_OP_NB 4
#endif
#ifndef S512_OP_WPE
#define S512_OP_WPE 1
#endif
#ifndef S512_OP_P
#define S512_OP_P 512
#endif
#ifndef S512_OP_E
#define S512_OP_E 16
#endif

namespace s512 {

// Workgroup shape per key type: P threads x E keys in registers per chunk,
// NB keys classified together, at least WPE waves per SIMD.
template <class T, bool KV>
struct OpCfg {
    static constexpr int P = S512_OP_P;
    static constexpr int E = S512_OP_E;
    static constexpr int NB = S512_OP_NB;
    static constexpr int WPE = S512_OP_WPE;
    static constexpr uint32_t CH = (uint32_t)(P * E);
    static constexpr int LINE = 128 / (int)sizeof(T);  // keys per 128-byte line
};

static inline int op_ceil_log2(uint64_t x) {
    int l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

struct OpSegIn {
    uint64_t src;   // first key in the source buffer
    uint64_t size;
    uint64_t fin;   // first key's position in sorted order
    uint64_t dst;   // exact: first key's position in the destination; gapped: region base
};

struct OpResult {
    uint32_t counts[32];
    uint64_t keys[16];
    std::vector<OpNext> next;
    std::vector<Seg> left;
    bool overflow = false, timeout = false, special = false;
};

constexpr int OPC_TICKET = 16, OPC_NSEG = 17, OPC_NCHUNK = 18, OPC_NSPLIT = 19, OPC_FLAGS = 20;

template <class T, bool KV, bool EXACT, class W>
static const void* op_kernel() {
    using C = OpCfg<T, KV>;
    return (const void*)k_onepass<T, KV, C::P, C::E, C::NB, C::WPE, EXACT, W>;
}

// Workgroups of the one-pass kernel resident on the whole device at once.
template <class T, bool KV, bool EXACT, class W>
static s512_status op_resident(uint32_t* out, uint32_t* per_cu) {
    int dev = 0, cus = 0, occ = 0;
    S512_HIP(hipGetDevice(&dev));
    S512_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    S512_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, op_kernel<T, KV, EXACT, W>(), OpCfg<T, KV>::P, 0));
    if (occ < 1) occ = 1;
    *out = (uint32_t)(occ * cus);
    if (per_cu) *per_cu = (uint32_t)occ;
    return S512_OK;
}

// Largest chunk count of a segment that the exact mode takes: the resident
// workgroups less a margin for work of other streams (the learned
// partitioner's trainer) that may hold a few compute units.
template <class T, bool KV>
static s512_status op_exact_limit(uint32_t* out) {
    uint32_t res = 0, occ = 0;
    if (s512_status rc = op_resident<T, KV, true, uint32_t>(&res, &occ)) return rc;
    *out = res > 8 * occ ? res - 8 * occ : 1;
    return S512_OK;
}

__global__ void k_op_zero(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// One one-pass level over `in` (sources in src/vsrc, outputs in dst/vdst; A is
// the caller's array, B the scratch one), then its routing, and -- unless the
// level overflowed or timed out -- the moves and tile sorts of its finished
// buckets.  `pre` pieces are moved B -> A first.  Synchronises the stream.
template <class T, bool KV, bool EXACT, bool OUT_B, class W>
static s512_status op_level(const std::vector<OpSegIn>& in, const std::vector<MovePiece>& pre, T* A, int32_t* VA,
                            T* B, int32_t* VB, uint64_t dst_limit, hipStream_t st, const EngineParams& prm,
                            uint32_t ncap, bool detect_special, bool f64_fast, int level, Profiler& pf,
                            DeviceState& ds, OpResult& res) {
    using K = typename KeyTraits<T>::K;
    using C = OpCfg<T, KV>;
    int ncls = 0;
    const uint32_t* caps = tile_class_caps<T, KV>(&ncls);
    const uint32_t smax = 8192;
    const uint32_t m = (uint32_t)in.size();
    const uint32_t src_buf = OUT_B ? 0u : 1u;  // root level reads A, the second level B
    T* src = OUT_B ? A : B;
    T* dst = OUT_B ? B : A;
    const int32_t* vsrc = OUT_B ? VA : VB;
    int32_t* vdst = OUT_B ? VB : VA;

    std::vector<PartSeg> ps(m);
    std::vector<ChunkRef> crs;
    uint64_t sumF = 0, sumB = 0, sumG = 0, sumSt = 0, total = 0;
    uint32_t max_ns = 0;
    for (uint32_t i = 0; i < m; ++i) {
        PartSeg& p = ps[i];
        memset(&p, 0, sizeof(p));
        p.begin = in[i].src;
        p.size = in[i].size;
        p.buf = src_buf;
        int lf = op_ceil_log2((p.size + prm.target_leaf - 1) / prm.target_leaf);
        lf = std::max(1, std::min(lf, prm.max_logf));
        p.logf = (uint32_t)lf;
        const uint32_t F = 1u << lf;
        p.nsample = std::min<uint32_t>(smax, prm.oversample * F);
        max_ns = std::max(max_ns, p.nsample);
        p.chunk = C::CH;
        p.nchunks = (uint32_t)((p.size + C::CH - 1) / C::CH);
        p.tree_base = sumF;
        p.bkt_base = sumB;
        p.guide_base = sumG;
        p.sample_base = (uint64_t)i * smax;
        p.dst = in[i].dst;
        p.fin = in[i].fin;
        p.st_base = sumSt;
        sumF += F;
        sumB += 2 * F;
        sumG += guide_words(F);
        sumSt += (uint64_t)p.nchunks * 2 * F;
        total += p.size;
        for (uint32_t c = 0; c < p.nchunks; ++c) crs.push_back(ChunkRef{i, c});
    }
    const uint32_t tc = (uint32_t)crs.size();
    const uint64_t move_cap = sumB + total / COPY_PIECE + 1;
    const uint32_t npre = (uint32_t)pre.size();

    SizeCounter sc;
    sc.take<PartSeg>(m);
    sc.take<ChunkRef>(tc);
    sc.take<TileSeg>(m);
    sc.take<uint32_t>(32);
    sc.take<uint64_t>(16);
    sc.take<MovePiece>(npre);
    sc.take<uint32_t>(1);
    const size_t up_bytes = sc.off;
    sc.take<K>((size_t)m * smax);
    sc.take<K>(sumF);
    sc.take<uint32_t>(sumG);
    sc.take<uint32_t>(sumF);
    sc.take<uint64_t>(sumF);
    sc.take<uint64_t>(sumF);
    sc.take<uint64_t>(sumB);
    sc.take<W>(sumSt);
    for (int c = 0; c < ncls; ++c) sc.take<TileSeg>(sumB);
    sc.take<MovePiece>(move_cap);
    sc.take<OpNext>(sumB);
    sc.take<Seg>(sumB);
    s512_status rc;
    if ((rc = ds.arena.reserve(sc.off, st))) return rc;
    if ((rc = ds.pinned.reserve(up_bytes, st))) return rc;
    const uint64_t pre_back = std::min<uint64_t>(sumB, 4096);  // list entries read back with the counts
    if ((rc = ds.pinned2.reserve(4096 + std::max<uint64_t>(sumB, 1) * std::max(sizeof(OpNext), sizeof(Seg)), st)))
        return rc;

    Arena ar{(char*)ds.arena.ptr, 0}, hp{(char*)ds.pinned.ptr, 0};
    PartSeg* d_ps = ar.take<PartSeg>(m);
    ChunkRef* d_cr = ar.take<ChunkRef>(tc);
    TileSeg* d_sl = ar.take<TileSeg>(m);
    uint32_t* d_cnt = ar.take<uint32_t>(32);
    uint64_t* d_cnt64 = ar.take<uint64_t>(16);
    MovePiece* d_pre = ar.take<MovePiece>(npre);
    uint32_t* d_npre = ar.take<uint32_t>(1);
    K* d_smp = ar.take<K>((size_t)m * smax);
    K* d_tree = ar.take<K>(sumF);
    uint32_t* d_guide = ar.take<uint32_t>(sumG);
    uint32_t* d_capcnt = ar.take<uint32_t>(sumF);
    uint64_t* d_pstart = ar.take<uint64_t>(sumF);
    uint64_t* d_pcap = ar.take<uint64_t>(sumF);
    uint64_t* d_bsize = ar.take<uint64_t>(sumB);
    W* d_status = ar.take<W>(sumSt);
    OpPlan pl;
    memset(&pl, 0, sizeof(pl));
    for (int c = 0; c < ncls; ++c) {
        pl.tile[c] = ar.take<TileSeg>(sumB);
        pl.tile_cap[c] = caps[c];
        pl.cap[c] = (uint32_t)sumB;
    }
    pl.nclass = (uint32_t)ncls;
    pl.counts = d_cnt;
    pl.keys = d_cnt64;
    pl.move = ar.take<MovePiece>(move_cap);
    pl.next = ar.take<OpNext>(sumB);
    pl.left = ar.take<Seg>(sumB);
    pl.cap[8] = (uint32_t)std::min<uint64_t>(move_cap, 0xFFFFFFFFu);
    pl.cap[9] = (uint32_t)sumB;
    pl.cap[10] = (uint32_t)sumB;

    PartSeg* h_ps = hp.take<PartSeg>(m);
    ChunkRef* h_cr = hp.take<ChunkRef>(tc);
    TileSeg* h_sl = hp.take<TileSeg>(m);
    uint32_t* h_cnt = hp.take<uint32_t>(32);
    uint64_t* h_cnt64 = hp.take<uint64_t>(16);
    MovePiece* h_pre = hp.take<MovePiece>(npre);
    uint32_t* h_npre = hp.take<uint32_t>(1);
    memcpy(h_ps, ps.data(), m * sizeof(PartSeg));
    memcpy(h_cr, crs.data(), tc * sizeof(ChunkRef));
    for (uint32_t i = 0; i < m; ++i) h_sl[i] = TileSeg{(uint64_t)i * smax, ps[i].nsample, 0};
    memset(h_cnt, 0, 32 * sizeof(uint32_t));
    memset(h_cnt64, 0, 16 * sizeof(uint64_t));
    h_cnt[OPC_NSEG] = m;
    h_cnt[OPC_NCHUNK] = tc;
    h_cnt[OPC_NSPLIT] = m;
    if (npre) memcpy(h_pre, pre.data(), npre * sizeof(MovePiece));
    *h_npre = npre;
    S512_HIP(hipMemcpyAsync(ds.arena.ptr, ds.pinned.ptr, up_bytes, hipMemcpyHostToDevice, st));

    if (npre) {
        pf.begin(S512_PHASE_COPY);
        k_move<T, KV><<<std::min<uint32_t>(npre, 16384), 256, 0, st>>>(d_pre, d_npre, B, A, VB, VA);
        S512_CHECK_LAUNCH();
        pf.end();
    }
    const uint64_t seed = prm.seed + 0x9E37ull * (uint64_t)level;
    pf.begin(S512_PHASE_SAMPLE);
    k_sample<T><<<m, 256, 0, st>>>(d_ps, d_cnt + OPC_NSEG, A, B, d_smp, smax, seed);
    S512_CHECK_LAUNCH();
    pf.end();
    pf.begin(S512_PHASE_SAMPLE_SORT);
    const uint32_t* nsl = d_cnt + OPC_NSPLIT;
    if (max_ns <= 1024) k_tile_sort<K, 64, 16><<<m, 64, 0, st>>>(d_sl, nsl, d_smp, d_smp);
    else if (max_ns <= 2048) k_tile_sort<K, 128, 16><<<m, 128, 0, st>>>(d_sl, nsl, d_smp, d_smp);
    else if (max_ns <= 4096) k_tile_sort<K, 256, 16><<<m, 256, 0, st>>>(d_sl, nsl, d_smp, d_smp);
    else k_tile_sort<K, 512, 16><<<m, 512, 0, st>>>(d_sl, nsl, d_smp, d_smp);
    S512_CHECK_LAUNCH();
    pf.end();
    pf.begin(S512_PHASE_TREE);
    k_build_tree<K><<<m, FMAX, 0, st>>>(d_ps, d_cnt + OPC_NSEG, d_smp, smax, d_tree, d_guide);
    S512_CHECK_LAUNCH();
    pf.end();
    if constexpr (!EXACT) {
        pf.begin(S512_PHASE_SAMPLE);
        k_op_zero<<<(unsigned)std::min<uint64_t>((sumF + 255) / 256, 1024), 256, 0, st>>>(d_capcnt, sumF);
        S512_CHECK_LAUNCH();
        const uint32_t parts = std::max<uint32_t>(1, std::min<uint32_t>(256, ncap / 2048));
        k_cap_count<T><<<dim3(parts, m), 256, 0, st>>>(d_ps, A, B, d_tree, d_guide, ncap, seed, d_capcnt);
        S512_CHECK_LAUNCH();
        k_cap_plan<C::LINE><<<m, FMAX, 0, st>>>(d_ps, d_cnt + OPC_NSEG, d_capcnt, ncap, (double)prm.region_z, d_pstart, d_pcap,
                                                d_cnt + OPC_FLAGS);
        S512_CHECK_LAUNCH();
        pf.end();
    }
    pf.begin(S512_PHASE_SCAN);
    if constexpr (EXACT) {  // look-back words start unpublished
        const uint64_t per = std::max<uint64_t>(1, sumSt / std::max<uint32_t>(m, 1));
        const uint32_t gx = (uint32_t)std::min<uint64_t>((per + 255) / 256, 4096);
        k_status_clear<W><<<dim3(gx, std::min<uint32_t>(m, 65535)), 256, 0, st>>>(d_ps, d_cnt + OPC_NSEG, d_status);
        S512_CHECK_LAUNCH();
    } else {  // reservation cursors (the bucket totals) start at zero
        k_op_zero<<<(unsigned)std::min<uint64_t>((2 * sumB + 255) / 256, 1024), 256, 0, st>>>((uint32_t*)d_bsize,
                                                                                           2 * sumB);
        S512_CHECK_LAUNCH();
    }
    pf.end();
    OpArgs oa;
    oa.ps = d_ps;
    oa.cr = d_cr;
    oa.nchunk = d_cnt + OPC_NCHUNK;
    oa.ticket = d_cnt + OPC_TICKET;
    oa.status = d_status;
    oa.pstart = d_pstart;
    oa.pcap = d_pcap;
    oa.bsize = d_bsize;
    oa.flags = d_cnt + OPC_FLAGS;
    oa.dst_limit = dst_limit;
    oa.detect_special = detect_special ? 1u : 0u;
    uint32_t resident = 0;
    if ((rc = op_resident<T, KV, EXACT, W>(&resident, nullptr))) return rc;
    pf.begin(S512_PHASE_SCATTER);
    k_onepass<T, KV, C::P, C::E, C::NB, C::WPE, EXACT, W><<<std::min(tc, resident), C::P, 0, st>>>(oa, src, dst, vsrc, vdst,
                                                                                         d_tree, d_guide);
    S512_CHECK_LAUNCH();
    pf.end();
    pf.begin(S512_PHASE_PLAN);
    k_op_plan<EXACT, OUT_B><<<m, BMAX, 0, st>>>(d_ps, d_cnt + OPC_NSEG, d_pstart, d_pcap, d_bsize, pl);
    S512_CHECK_LAUNCH();
    pf.end();

    char* h_out = (char*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h_out, d_cnt, 32 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(h_out + 256, d_cnt64, 16 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    char* h_list = h_out + 4096;
    if (OUT_B) S512_HIP(hipMemcpyAsync(h_list, pl.next, pre_back * sizeof(OpNext), hipMemcpyDeviceToHost, st));
    else S512_HIP(hipMemcpyAsync(h_list, pl.left, pre_back * sizeof(Seg), hipMemcpyDeviceToHost, st));
    S512_HIP(hipStreamSynchronize(st));
    memcpy(res.counts, h_out, sizeof(res.counts));
    memcpy(res.keys, h_out + 256, sizeof(res.keys));
    const uint32_t* fl = res.counts + OPC_FLAGS;
    res.overflow = fl[OP_FLAG_OVERFLOW] != 0 || res.counts[15] != 0;
    res.timeout = fl[OP_FLAG_TIMEOUT] != 0;
    res.special = fl[OP_FLAG_SPECIAL] != 0;
    if (res.overflow || res.timeout) return S512_OK;  // the caller redoes this level
    if (fl[OP_FLAG_RANGE]) {
        set_error("one-pass level %d: a destination outside its region", level);
        return S512_ERR_INTERNAL;
    }
    if (res.counts[14]) {
        set_error("one-pass level %d: routing lists overflowed", level);
        return S512_ERR_INTERNAL;
    }
    if (res.overflow || res.timeout) return S512_OK;  // the caller redoes this level

    const uint32_t nl = OUT_B ? res.counts[9] : res.counts[10];
    if (nl > sumB) {
        set_error("one-pass level %d: %u segments for %llu buckets", level, nl, (unsigned long long)sumB);
        return S512_ERR_INTERNAL;
    }
    if (nl > pre_back) {
        if (OUT_B) S512_HIP(hipMemcpyAsync(h_list + pre_back * sizeof(OpNext), pl.next + pre_back,
                                           (nl - pre_back) * sizeof(OpNext), hipMemcpyDeviceToHost, st));
        else S512_HIP(hipMemcpyAsync(h_list + pre_back * sizeof(Seg), pl.left + pre_back,
                                     (nl - pre_back) * sizeof(Seg), hipMemcpyDeviceToHost, st));
        S512_HIP(hipStreamSynchronize(st));
    }
    if (OUT_B) res.next.assign((OpNext*)h_list, (OpNext*)h_list + nl);
    else res.left.assign((Seg*)h_list, (Seg*)h_list + nl);

    const bool fast = f64_fast || (detect_special && !res.special);
    if (res.counts[8]) {
        pf.begin(S512_PHASE_COPY);
        k_move<T, KV><<<std::min<uint32_t>(res.counts[8], 16384), 256, 0, st>>>(pl.move, d_cnt + 8, B, A, VB, VA);
        S512_CHECK_LAUNCH();
        pf.end();
    }
    for (int c = 0; c < ncls; ++c) {
        if (!res.counts[c]) continue;
        pf.begin(S512_PHASE_TILE);
        launch_tile_list<T, KV>(c, prm.tile_plan, pl.tile[c], d_cnt + c, res.counts[c], A, VA, B, VB, st,
                                std::is_same<T, double>::value && !KV && prm.f64_minmax && fast);
        S512_CHECK_LAUNCH();
        pf.end();
    }
    return S512_OK;
}

template <class T, bool KV>
s512_status onepass_engine(T* a, int32_t* va, size_t n, hipStream_t st, const EngineParams& prm, s512_stats* stats,
                           Profiler& pf, DeviceState& ds, bool* done) {
    using C = OpCfg<T, KV>;
    *done = false;
    s512_status rc;
    uint32_t gmax = 0;
    if ((rc = op_exact_limit<T, KV>(&gmax))) return rc;
    int lf = op_ceil_log2((n + prm.target_leaf - 1) / prm.target_leaf);
    lf = std::max(1, std::min(lf, prm.max_logf));
    const uint32_t F1 = 1u << lf;
    const uint64_t nch1 = (n + C::CH - 1) / C::CH;
    const bool exact1 = nch1 <= gmax;
    const uint32_t ncap = exact1 ? 0u : (uint32_t)std::min<uint64_t>(1u << 18, std::max<uint64_t>(n / 64, 1u << 14));
    const uint64_t region = exact1 ? n : op_region_bound(n, ncap, F1, C::LINE, (double)prm.region_z);
    const uint64_t bcap = std::max<uint64_t>(region, n);
    const size_t kbytes = (bcap * sizeof(T) + 255) & ~(size_t)255;
    if ((rc = ds.data.reserve(kbytes + (KV ? bcap * sizeof(int32_t) : 0), st))) return rc;
    T* b = (T*)ds.data.ptr;
    int32_t* vb = KV ? (int32_t*)((char*)ds.data.ptr + kbytes) : nullptr;
    const bool fp64 = std::is_same<T, double>::value && !KV;

    // ---- root level
    OpResult r1;
    std::vector<OpSegIn> l1{OpSegIn{0, n, 0, 0}};
    const std::vector<MovePiece> none;
    if (exact1) rc = op_level<T, KV, true, true, uint32_t>(l1, none, a, va, b, vb, bcap, st, prm, 0, fp64, false, 1,
                                                           pf, ds, r1);
    else if (n >= 0x7FFFFFF0ull) rc = op_level<T, KV, false, true, uint64_t>(l1, none, a, va, b, vb, bcap, st, prm,
                                                                             ncap, fp64, false, 1, pf, ds, r1);
    else rc = op_level<T, KV, false, true, uint32_t>(l1, none, a, va, b, vb, bcap, st, prm, ncap, fp64, false, 1, pf,
                                                     ds, r1);
    if (rc) return rc;
    if (r1.overflow || r1.timeout) {  // a is untouched: the caller sorts with two passes
        if (stats) stats->region_overflow = 1;
        return S512_OK;
    }
    const bool fast = fp64 && !r1.special;
    uint64_t tiles = 0, keys_tile = 0, keys_copy = r1.keys[8];
    for (int c = 0; c < 8; ++c) {
        tiles += r1.counts[c];
        keys_tile += r1.keys[c];
    }

    // ---- second level
    std::vector<OpSegIn> l2;
    std::vector<MovePiece> pre;
    std::vector<Seg> left;
    auto move_all = [&](uint64_t from, uint64_t to, uint64_t size) {
        for (uint64_t p = 0; p < size; p += COPY_PIECE)
            pre.push_back(MovePiece{from + p, to + p, (uint32_t)std::min<uint64_t>(COPY_PIECE, size - p), 0u});
    };
    for (const OpNext& x : r1.next) {
        if ((x.size + C::CH - 1) / C::CH <= gmax) {
            l2.push_back(OpSegIn{x.src, x.size, x.dst, x.dst});
        } else {
            move_all(x.src, x.dst, x.size);
            left.push_back(Seg{x.dst, x.size, 0u, 0u});
        }
    }
    int levels = 1;
    uint64_t keys_pass = n;
    if (!l2.empty()) {
        OpResult r2;
        if ((rc = op_level<T, KV, true, false, uint32_t>(l2, pre, a, va, b, vb, n, st, prm, 0, false, fast, 2, pf, ds,
                                                         r2)))
            return rc;
        pre.clear();
        levels = 2;
        for (const OpSegIn& s : l2) keys_pass += s.size;
        if (r2.timeout) {
            // nothing of the second level is trusted: its segments go to the
            // multi-level engine from their unchanged copies in b
            for (const OpSegIn& s : l2) {
                move_all(s.src, s.fin, s.size);
                left.push_back(Seg{s.fin, s.size, 0u, 0u});
            }
        } else {
            left.insert(left.end(), r2.left.begin(), r2.left.end());
            for (int c = 0; c < 8; ++c) {
                tiles += r2.counts[c];
                keys_tile += r2.keys[c];
            }
        }
    }
    if (!pre.empty()) {
        // moves that no level launch carried: one small upload and copy
        SizeCounter sc;
        sc.take<MovePiece>(pre.size());
        sc.take<uint32_t>(1);
        if ((rc = ds.arena.reserve(sc.off, st))) return rc;
        if ((rc = ds.pinned.reserve(sc.off, st))) return rc;
        Arena ar{(char*)ds.arena.ptr, 0}, hp{(char*)ds.pinned.ptr, 0};
        MovePiece* d_pc = ar.take<MovePiece>(pre.size());
        uint32_t* d_n = ar.take<uint32_t>(1);
        MovePiece* h_pc = hp.take<MovePiece>(pre.size());
        uint32_t* h_n = hp.take<uint32_t>(1);
        memcpy(h_pc, pre.data(), pre.size() * sizeof(MovePiece));
        *h_n = (uint32_t)pre.size();
        S512_HIP(hipMemcpyAsync(ds.arena.ptr, ds.pinned.ptr, sc.off, hipMemcpyHostToDevice, st));
        pf.begin(S512_PHASE_COPY);
        k_move<T, KV><<<std::min<uint32_t>((uint32_t)pre.size(), 16384), 256, 0, st>>>(d_pc, d_n, b, a, vb, va);
        S512_CHECK_LAUNCH();
        pf.end();
        S512_HIP(hipStreamSynchronize(st));  // pinned staging is reused below
    }

    // ---- what is left: the multi-level engine, in a with b as scratch
    s512_stats sub;
    memset(&sub, 0, sizeof(sub));
    if (!left.empty()) {
        LevelState ls;
        if ((rc = run_levels<T, KV>(a, va, b, vb, n, std::move(left), st, prm, stats ? &sub : nullptr, pf, ds,
                                    LearnedModelDev{nullptr, nullptr}, ls)))
            return rc;
    } else {
        S512_HIP(hipStreamSynchronize(st));
        if (stats) pf.collect(stats->kernel_ms, stats->launches);
    }
    if (stats) {
        if (sub.levels) {
            memcpy(stats->kernel_ms, sub.kernel_ms, sizeof(sub.kernel_ms));
            memcpy(stats->launches, sub.launches, sizeof(sub.launches));
            for (int i = 0; i < S512_NPHASE; ++i) stats->keys[i] += sub.keys[i];
        }
        stats->levels = levels + sub.levels;
        stats->tiles = tiles + sub.tiles;
        stats->keys[S512_PHASE_SCATTER] += keys_pass;
        stats->keys[S512_PHASE_TILE] += keys_tile;
        stats->keys[S512_PHASE_COPY] += keys_copy;
        stats->classifier = 0;
        stats->engine = S512_ENGINE_ONE_PASS;
    }
    *done = true;
    return S512_OK;
}

#define S512_OP_ENGINE(T, KV)                                                                                      \
    template s512_status onepass_engine<T, KV>(T*, int32_t*, size_t, hipStream_t, const EngineParams&, s512_stats*, \
                                               Profiler&, DeviceState&, bool*);
S512_OP_ENGINE(int32_t, false)
S512_OP_ENGINE(int32_t, true)
S512_OP_ENGINE(double, false)
S512_OP_ENGINE(int64_t, false)
S512_OP_ENGINE(uint64_t, false)
S512_OP_ENGINE(uint32_t, false)
#undef S512_OP_ENGINE

}  // namespace s512
