// s512_onepass.hpp -- one-pass partition levels: every key of a level is read
// once, classified once and written once.
//
// The reference partitions with one read and one write of every key per pass
// (Partition512's compress-store loop, sort512.hpp:5748-5772 int, 5840-5864
// double).  The two-pass level of s512_kernels.hpp reads every key twice
// (k_count, then k_scatter).  Here one kernel does both for a chunk of
// P x E keys held in registers:
//
//   A  load the chunk, classify every key, rank it inside its bucket with
//      LDS atomics (the chunk histogram)
//   B  find where the chunk's run of each bucket goes (below)
//   C  stage the chunk in LDS in bucket order and write each bucket's run to
//      its destination with consecutive threads on consecutive addresses.
// Chunks are taken in order from an atomic ticket.
//
// Where a bucket's keys go needs the sizes of all buckets before it.  Two
// modes provide them:
//   gapped  every (ordinary bucket, equality bucket) pair j has a region of
//           pcap[j] slots sized from a second, larger random sample; the
//           ordinary bucket fills it from the front, the equality bucket from
//           the back.  A chunk reserves its run of each bucket with one
//           device-scope atomic on the bucket's cursor, so the runs of a
//           bucket land in reservation order (the next level sorts them).  A
//           pair whose totals outgrow its region is reported as overflowed
//           and the caller redoes the level from its unchanged source.  Used
//           for the root level, whose thousands of chunks cannot all be
//           resident at once.
//   exact   decoupled look-back (Merrill & Garland 2016, "Single-pass
//           Parallel Prefix Scan with Decoupled Look-back"): a chunk
//           publishes its counts per bucket, adds up the counts of earlier
//           chunks of its segment until one that has published its
//           inclusive prefix, and publishes its own; then it waits for the
//           inclusive prefix of the segment's last chunk (the bucket totals)
//           and writes final positions, in the deterministic chunk order.
//           Only for segments whose chunks all fit on the GPU at once
//           (nchunks <= the kernel's resident workgroups), so the wait always
//           ends.
//
// Measured on MI355X (268,435,456 fp64 keys, two levels): 5.2 ms, against
// 4.0 ms for the two-pass levels (count, then scatter).  With one 8-wave
// workgroup per CU (the registers of 8192 keys in flight allow no more) the
// phases of a chunk do not overlap with other work: a chunk spends about
// 20 us where its 64 KB read and 64 KB write need 6.5 us of the CU's share of
// HBM, and every device-scope round trip costs about 2 us while the memory
// system streams.  The engine therefore runs these levels only when asked
// (S512_ENGINE_ONE_PASS).
// Every wait is bounded (a flag that any timed-out wait sets ends all the
// others), and every store is checked against its buffer, so a kernel that
// meets a missing status word finishes, writes nothing out of bounds and
// reports it.
#pragma once
#include "s512_kernels.hpp"

namespace s512 {

constexpr uint64_t OP_NOWRITE = ~0ull;
constexpr uint32_t OP_SPIN_LIMIT = 1u << 20;  // polls of one status word before giving up
constexpr uint32_t OP_FLAG_OVERFLOW = 0, OP_FLAG_TIMEOUT = 1, OP_FLAG_SPECIAL = 2, OP_FLAG_RANGE = 3;

struct MovePiece {  // copy size keys from[src..] -> to[dst..]; fill: all keys equal from[src]
    uint64_t src, dst;
    uint32_t size, fill;
};
struct OpNext {  // a bucket that needs another level
    uint64_t src, dst, size;
};

struct OpArgs {
    const PartSeg* ps;
    const ChunkRef* cr;       // chunks, segment by segment, in order
    const uint32_t* nchunk;   // chunks of the level (device)
    uint32_t* ticket;         // next chunk to take (zero before the launch)
    void* status;             // look-back words
    const uint64_t* pstart;   // gapped: first slot of pair j's region (tree_base + j)
    const uint64_t* pcap;     // gapped: slots of pair j's region
    uint64_t* bsize;          // bucket totals (bkt_base + b), written by each segment's last chunk
    uint32_t* flags;          // OP_FLAG_*
    uint64_t dst_limit;       // keys in the destination buffer (bounds of every store)
    uint32_t detect_special;  // look for NaN / -0.0 keys (fp64 leaves by min/max otherwise)
};

template <class W>
__device__ __forceinline__ W op_load(const W* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class W>
__device__ __forceinline__ void op_store(W* p, W v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Status word: 0 = not published, else ((v + 1) << 1) | inclusive.  Each word
// is self-contained (no other data is published with it), so relaxed
// device-scope loads and stores suffice.
template <class W>
__device__ __forceinline__ W op_pack(uint64_t v, bool incl) {
    return (W)(((v + 1) << 1) | (incl ? 1u : 0u));
}
__device__ __forceinline__ uint64_t op_val(uint64_t w) { return (w >> 1) - 1; }

// Poll *p until it is published (inclusive only, if `incl`).  Gives up after
// OP_SPIN_LIMIT polls or as soon as another wait has given up; then sets
// *tmo and returns an inclusive zero.
template <class W>
__device__ __forceinline__ W op_wait(const W* p, bool incl, uint32_t* flags, bool& tmo) {
    for (uint32_t it = 1;; ++it) {
        const W w = op_load(p);
        if (w != 0 && (!incl || (w & 1))) return w;
        if ((it & 255u) == 0) {
            if (it >= OP_SPIN_LIMIT) atomicOr(&flags[OP_FLAG_TIMEOUT], 1u);
            if (op_load(&flags[OP_FLAG_TIMEOUT])) {
                tmo = true;
                return op_pack<W>(0, true);
            }
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Exclusive scan, in place, of the first B entries of a BMAX-entry LDS array
// (entries B.. are read as zero and left alone); thread t owns entries
// t*BMAX/P .. .  `wt` is LDS scratch for P/64 values.  Ends with a barrier.
template <int P, class V>
__device__ __forceinline__ void op_bucket_scan(V* arr, uint32_t B, V* wt) {
    constexpr int BPT = BMAX / P;
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    V v[BPT], sum = 0;
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const uint32_t i = t * BPT + k;
        v[k] = i < B ? arr[i] : (V)0;
        sum += v[k];
    }
    V incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        V y;
        if constexpr (sizeof(V) == 8) {
            const uint32_t lo = __shfl_up((uint32_t)incl, d), hi = __shfl_up((uint32_t)(incl >> 32), d);
            y = ((V)hi << 32) | lo;
        } else {
            y = __shfl_up(incl, d);
        }
        if (lane >= (uint32_t)d) incl += y;
    }
    if (lane == 63) wt[wid] = incl;
    __syncthreads();
    V off = 0;
#pragma unroll
    for (int w = 0; w < P / 64; ++w) off += (w < (int)wid) ? wt[w] : (V)0;
    V run = off + incl - sum;
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const uint32_t i = t * BPT + k;
        if (i < B) arr[i] = run;
        run += v[k];
    }
    __syncthreads();
}

// ---------------------------------------------------------------- capacity sample
// Pair counts of a random sample of ncap keys per segment (positions hashed
// with a seed of their own), classified with the level's classifier:
// cnt[tree_base + j] = sample keys in pair j (ordinary bucket j and its
// equality bucket).  grid = (parts, segments).
template <class T>
__global__ __launch_bounds__(256) void k_cap_count(const PartSeg* ps, const T* a, const T* b,
                                                   const typename KeyTraits<T>::K* trees, const uint32_t* guides,
                                                   uint32_t ncap, uint64_t seed, uint32_t* cnt) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    __shared__ K lsrt[FMAX];
    __shared__ __align__(16) uint32_t lguide[LDS_TABLE_WORDS];
    __shared__ uint32_t lh[FMAX];
    const PartSeg s = ps[blockIdx.y];
    const uint32_t F = 1u << s.logf;
    const Guide g = load_classifier<K, 256, false>(s, trees, guides, lsrt, lguide);
    for (uint32_t j = threadIdx.x; j < F; j += 256) lh[j] = 0;
    __syncthreads();
    const T* src = (s.buf ? b : a) + s.begin;
    const uint32_t per = (ncap + gridDim.x - 1) / gridDim.x;
    const uint32_t i0 = blockIdx.x * per, i1 = i0 + per < ncap ? i0 + per : ncap;
    const uint64_t hs = seed ^ (s.begin * 0x632BE59BD9B4E019ull) ^ 0xCA9Aull;
    for (uint32_t i = i0 + threadIdx.x * CLS_NB; i < i1; i += 256 * CLS_NB) {
        K x[CLS_NB];
        uint32_t bk[CLS_NB];
#pragma unroll
        for (int u = 0; u < CLS_NB; ++u) {
            const uint32_t q = i + u < i1 ? i + u : i;
            x[u] = Tr::enc(src[__umul64hi(mix64(hs, q), s.size)]);
        }
        classify_batch<K, CLS_NB>(lsrt, lguide, g, x, bk);
#pragma unroll
        for (int u = 0; u < CLS_NB; ++u)
            if (i + u < i1) atomicAdd(&lh[bk[u] >> 1], 1u);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < F; j += 256)
        if (lh[j]) atomicAdd(&cnt[s.tree_base + j], lh[j]);
}

// Regions of the pairs of each segment: pair j holds up to
//   size * (c_j + z sqrt(c_j + 1) + 2) / ncap
// keys (c_j = its sample count, z the margin in standard deviations of that
// estimate), rounded up to a whole number of 128-byte lines; regions are
// consecutive from the segment's region base s.dst.  The sum over pairs is at
// most op_region_bound() (Cauchy-Schwarz on the square roots), which the host
// reserved.
__host__ __device__ inline uint64_t op_region_bound(uint64_t size, uint32_t ncap, uint32_t F, uint32_t line,
                                                    double z) {
    const double r =
        (double)size * ((double)ncap + z * sqrt((double)F * (double)(ncap + F)) + 2.0 * F) / (double)ncap;
    return (uint64_t)r + (uint64_t)F * (line + 1) + 4096;
}

template <int LINE>
__global__ __launch_bounds__(FMAX) void k_cap_plan(const PartSeg* ps, const uint32_t* nseg, const uint32_t* cnt,
                                                   uint32_t ncap, double z, uint64_t* pstart, uint64_t* pcap,
                                                   uint32_t* flags) {
    __shared__ uint64_t arr[BMAX];
    __shared__ uint64_t wt[FMAX / 64];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        const PartSeg s = ps[w];
        const uint32_t F = 1u << s.logf, t = threadIdx.x;
        uint64_t cap = 0;
        if (t < F) {
            const double c = (double)cnt[s.tree_base + t];
            const double est = (double)s.size * (c + z * sqrt(c + 1.0) + 2.0) / (double)ncap;
            cap = ((uint64_t)ceil(est) + LINE - 1) / LINE * LINE;
        }
        arr[t] = cap;
        arr[t + FMAX] = 0;
        __syncthreads();
        op_bucket_scan<FMAX, uint64_t>(arr, F, wt);
        if (t < F) {
            pstart[s.tree_base + t] = s.dst + arr[t];
            pcap[s.tree_base + t] = cap;
            if (t == F - 1 && arr[t] + cap > op_region_bound(s.size, ncap, F, LINE, z))
                atomicOr(&flags[OP_FLAG_RANGE], 1u);
        }
        __syncthreads();
    }
}

// Zero the look-back words of every chunk of every segment (the first F
// columns of each row, all 2F when the segment has equality buckets).
template <class W>
__global__ __launch_bounds__(256) void k_status_clear(const PartSeg* ps, const uint32_t* nseg, W* status) {
    for (uint32_t w = blockIdx.y; w < *nseg; w += gridDim.y) {
        const PartSeg s = ps[w];
        const uint32_t F = 1u << s.logf;
        const uint32_t used = (s.flags & GUIDE_EQBUCKETS) ? 2 * F : F;
        const uint64_t rows = s.nchunks;
        for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < rows * used; i += gridDim.x * 256ull) {
            const uint64_t r = i / used, c = i - r * used;
            status[s.st_base + r * (2 * F) + c] = 0;
        }
    }
}

// ---------------------------------------------------------------- the one-pass kernel
template <class T, bool KV, int P, int CH>
struct OpLds {
    using K = typename KeyTraits<T>::K;
    K lsrt[FMAX];
    __align__(16) uint32_t lguide[LDS_TABLE_WORDS];
    uint32_t hist[BMAX];    // chunk counts, then each bucket's first staging slot
    uint64_t dbase[BMAX];   // destination of the bucket's first key of this chunk (OP_NOWRITE: none)
    T stg[CH];              // the chunk in bucket order
    uint16_t stb[CH];
    int32_t stv[KV ? CH : 1];
    uint64_t wt[P / 64];
    uint32_t dummy[DUMMY_SLOTS];
    uint32_t tk, bad;
};

// The keys (and values) of chunk c of segment s, striped over the threads:
// thread t holds keys t, t + P, ...; positions past the segment's end repeat
// its last key (in bounds; they are never counted or stored).
template <class T, bool KV, int P, int E>
__device__ __forceinline__ void op_load_chunk(typename KeyTraits<T>::K (&xb)[E], int32_t (&xv)[KV ? E : 1],
                                              const T* src, const int32_t* vsrc, const ChunkRef& c, const PartSeg& s,
                                              uint32_t t) {
    constexpr int CH = P * E;
    const uint64_t lo = (uint64_t)c.idx * CH;
    const uint32_t n = (uint32_t)((s.size - lo) < (uint64_t)CH ? (s.size - lo) : (uint64_t)CH);
    const T* sp = src + s.begin + lo;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = e * P + t, ic = i < n ? i : n - 1;
        xb[e] = KeyTraits<T>::bits(sp[ic]);
        if constexpr (KV) xv[e] = vsrc[s.begin + lo + ic];
    }
}

// P threads x E keys per chunk, NB keys classified together, at least WPE
// waves per SIMD (register budget).
//
// Each workgroup pipelines its chunks so that the round trips to the status
// words (about two microseconds each while the memory system streams) and
// the key loads overlap work:
//   - the next chunk's ticket is taken one chunk ahead;
//   - the first look-back step is issued right after the chunk's counts are
//     published and consumed only after the chunk is staged in LDS;
//   - once staged, the chunk's registers are free, and the next chunk's keys
//     (and, for another segment, its classifier tables) load while this
//     chunk finishes its look-back and writes.
template <class T, bool KV, int P, int E, int NB, int WPE, bool EXACT, class W>
__global__ __launch_bounds__(P, WPE) void k_onepass(OpArgs A, const T* src, T* dst, const int32_t* vsrc,
                                                    int32_t* vdst, const typename KeyTraits<T>::K* trees,
                                                    const uint32_t* guides) {
    using Tr = KeyTraits<T>;
    using K = typename Tr::K;
    constexpr int CH = P * E;
    using L = OpLds<T, KV, P, CH>;
    constexpr int PPT = (FMAX + P - 1) / P;             // pairs per thread in the look-back
    constexpr int GPT = (LDS_TABLE_WORDS + P - 1) / P;  // guide words per thread (table prefetch)
    constexpr int SPT = (FMAX + P - 1) / P;             // splitters per thread
    static_assert(CH <= (1 << 14) && BMAX <= (1 << 10), "slot and bucket share one word");
    __shared__ L sm;
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t total = *A.nchunk;
    if (t == 0) sm.tk = atomicAdd(A.ticket, 1u);
    __syncthreads();
    const uint32_t wk0 = sm.tk;
    if (wk0 >= total) return;  // uniform
    ChunkRef c = A.cr[wk0];
    PartSeg s = A.ps[c.seg];
    Guide g = load_classifier<K, P, false>(s, trees, guides, sm.lsrt, sm.lguide);
    K xb[E];
    int32_t xv[KV ? E : 1];
    op_load_chunk<T, KV, P, E>(xb, xv, src, vsrc, c, s, t);
    // Gapped mode takes the next ticket one chunk ahead: every wait there is
    // for a chunk with a smaller index, so the smallest unfinished chunk is
    // always some workgroup's current chunk and can go on.  Exact mode waits
    // for its segment's last chunk, which may need a chunk that a waiting
    // workgroup would hold unstarted: it takes the next ticket only once its
    // own waits are over.
    uint32_t tk_next = 0;
    if (!EXACT && t == 0) tk_next = atomicAdd(A.ticket, 1u);
    for (;;) {
        for (uint32_t j = t; j < BMAX; j += P) sm.hist[j] = 0;
        if (t == 0) {
            sm.bad = 0;
            if (!EXACT) sm.tk = tk_next;
        }
        __syncthreads();  // classifier tables, cleared counts and the next ticket visible
        const uint32_t wn_ahead = sm.tk;
        const uint64_t lo = (uint64_t)c.idx * CH;
        const uint32_t cnt = (uint32_t)((s.size - lo) < (uint64_t)CH ? (s.size - lo) : (uint64_t)CH);

        // ---- A: classify, rank
        uint32_t br[E];  // bucket | rank << 10, later bucket | slot << 10
        {
            uint32_t bk[E], rk[E];
            bool odd = false;
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += NB) {
                K kx[NB];
                uint32_t kb[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    kx[u] = Tr::enc_bits(xb[e0 + u]);
                    odd |= nan_or_negzero(kx[u]);
                }
                classify_batch<K, NB>(sm.lsrt, sm.lguide, g, kx, kb);
#pragma unroll
                for (int u = 0; u < NB; ++u) bk[e0 + u] = kb[u];
            }
            if (cnt == CH) {
                wave_rank_many<E>(sm.hist, sm.dummy, bk, rk);
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) rk[e] = (e * P + t < cnt) ? atomicAdd(&sm.hist[bk[e]], 1u) : 0u;
            }
#pragma unroll
            for (int e = 0; e < E; ++e) br[e] = bk[e] | (rk[e] << 10);
            // clamped duplicates of the last key are the same key: odd is exact
            if (A.detect_special && __any(odd) && lane == 0) atomicOr(&A.flags[OP_FLAG_SPECIAL], 1u);
        }
        __syncthreads();  // counts complete

        // ---- B1: publish the counts, issue the first look-back step
        const uint32_t F = 1u << s.logf;
        const bool eq = (s.flags & GUIDE_EQBUCKETS) != 0;
        const uint32_t cols = 2 * F;
        W* sbase = (W*)A.status + s.st_base;
        W* me = sbase + (uint64_t)c.idx * cols;
        const bool first = c.idx == 0, last = c.idx + 1 == s.nchunks;
        uint32_t cnt0[PPT], cnt1[PPT];
        W w0[PPT], w1[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = t + k * P;
            cnt0[k] = cnt1[k] = 0;
            w0[k] = w1[k] = 0;
            if (j >= F) continue;
            cnt0[k] = sm.hist[2 * j];
            cnt1[k] = eq ? sm.hist[2 * j + 1] : 0u;
            if constexpr (EXACT) {
                op_store(me + j, op_pack<W>(cnt0[k], first));
                if (eq) op_store(me + F + j, op_pack<W>(cnt1[k], first));
                if (!first) {
                    const W* row = sbase + (uint64_t)(c.idx - 1) * cols;
                    w0[k] = op_load(row + j);
                    if (eq) w1[k] = op_load(row + F + j);
                }
            } else {
                // gapped: reserve the chunk's runs in the pair's region with
                // one atomic per bucket (no chain of prefixes to wait for);
                // the runs of a bucket land in the order the chunks reserve
                // them, which the next level's sort makes irrelevant.  The
                // cursors end as the bucket totals.
                unsigned long long* cur = (unsigned long long*)A.bsize + s.bkt_base + 2 * j;
                w0[k] = cnt0[k] ? (W)atomicAdd(cur, (unsigned long long)cnt0[k]) : (W)0;
                w1[k] = cnt1[k] ? (W)atomicAdd(cur + 1, (unsigned long long)cnt1[k]) : (W)0;
            }
        }
        // staging slots: exclusive scan of the chunk counts (bucket order)
        op_bucket_scan<P, uint32_t>(sm.hist, cols, reinterpret_cast<uint32_t*>(sm.wt));
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t bk = br[e] & 1023u;
            br[e] = bk | ((sm.hist[bk] + (br[e] >> 10)) << 10);
            if (e * P + t < cnt) {
                const uint32_t q = br[e] >> 10;
                sm.stg[q] = __builtin_bit_cast(T, xb[e]);
                sm.stb[q] = (uint16_t)bk;
                if constexpr (KV) sm.stv[q] = xv[e];
            }
        }
        __syncthreads();  // staged: the key registers are free

        // ---- the next chunk: its keys and, for another segment, its tables
        bool more = false, newseg = false;
        ChunkRef cn = c;
        PartSeg sn = s;
        uint32_t gpf[GPT];
        K spf[SPT];
        auto prefetch = [&](uint32_t wn) {
            more = wn < total;
            if (!more) return;
            cn = A.cr[wn];
            sn = A.ps[cn.seg];
            op_load_chunk<T, KV, P, E>(xb, xv, src, vsrc, cn, sn, t);
            newseg = cn.seg != c.seg;
            if (newseg) {
                const uint32_t Fn = 1u << sn.logf, Gn = guide_words(Fn);
#pragma unroll
                for (int k = 0; k < GPT; ++k) {
                    const uint32_t i = t + k * P;
                    gpf[k] = i < Gn ? guides[sn.guide_base + i] : 0u;
                }
#pragma unroll
                for (int k = 0; k < SPT; ++k) {
                    const uint32_t i = t + k * P;
                    spf[k] = i < Fn ? trees[sn.tree_base + i] : (K)0;
                }
            }
        };
        if constexpr (!EXACT) {
            prefetch(wn_ahead);
            if (t == 0 && more) tk_next = atomicAdd(A.ticket, 1u);
        }

        // ---- B2: finish the look-back (exact) / take the reservations (gapped)
        bool tmo = false;
        uint64_t pre0[PPT], pre1[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = t + k * P;
            pre0[k] = pre1[k] = 0;
            if (j >= F) continue;
            uint64_t p0 = 0, p1 = 0;
            if constexpr (EXACT) {
                if (!first) {
                    bool d0 = false, d1 = !eq;
                    for (uint32_t q = c.idx; q > 0 && !(d0 && d1);) {
                        --q;
                        const W* row = sbase + (uint64_t)q * cols;
                        W a0 = 0, a1 = 0;
                        if (q + 1 == c.idx) {  // the step issued before staging
                            a0 = w0[k];
                            a1 = w1[k];
                        } else {
                            if (!d0) a0 = op_load(row + j);
                            if (!d1) a1 = op_load(row + F + j);
                        }
                        if (!d0) {
                            if (a0 == 0) a0 = op_wait(row + j, false, A.flags, tmo);
                            p0 += op_val(a0);
                            d0 = (a0 & 1) != 0;
                        }
                        if (!d1) {
                            if (a1 == 0) a1 = op_wait(row + F + j, false, A.flags, tmo);
                            p1 += op_val(a1);
                            d1 = (a1 & 1) != 0;
                        }
                    }
                    op_store(me + j, op_pack<W>(p0 + cnt0[k], true));
                    if (eq) op_store(me + F + j, op_pack<W>(p1 + cnt1[k], true));
                }
                const uint64_t i0 = p0 + cnt0[k], i1 = p1 + cnt1[k];
                if (last) {
                    A.bsize[s.bkt_base + 2 * j] = i0;
                    A.bsize[s.bkt_base + 2 * j + 1] = i1;
                }
                // the segment's bucket totals: the last chunk's inclusive prefix
                uint64_t t0 = i0, t1 = i1;
                if (!last) {
                    const W* lr = sbase + (uint64_t)(s.nchunks - 1) * cols;
                    t0 = op_val(op_wait(lr + j, true, A.flags, tmo));
                    t1 = eq ? op_val(op_wait(lr + F + j, true, A.flags, tmo)) : 0u;
                }
                sm.dbase[2 * j] = t0;
                sm.dbase[2 * j + 1] = t1;
            } else {
                // a run that would leave the region is not written; the
                // routing kernel compares the totals with the capacity
                p0 = w0[k];
                p1 = w1[k];
                const uint64_t cap = A.pcap[s.tree_base + j], ps0 = A.pstart[s.tree_base + j];
                const uint64_t i0 = p0 + cnt0[k], i1 = p1 + cnt1[k];
                if (i0 > cap || i1 > cap) {
                    atomicOr(&A.flags[OP_FLAG_OVERFLOW], 1u);
                    sm.dbase[2 * j] = sm.dbase[2 * j + 1] = OP_NOWRITE;
                } else {
                    sm.dbase[2 * j] = ps0 + p0;
                    sm.dbase[2 * j + 1] = ps0 + cap - i1;
                }
            }
            pre0[k] = p0;
            pre1[k] = p1;
        }
        if (tmo) sm.bad = 1;
        uint32_t tk = 0;
        if (EXACT && t == 0) tk = atomicAdd(A.ticket, 1u);  // this chunk's waits are over
        __syncthreads();
        // exact: bucket starts inside the segment from the totals
        if constexpr (EXACT) op_bucket_scan<P, uint64_t>(sm.dbase, cols, sm.wt);
        const bool bad = sm.bad != 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const uint32_t j = t + k * P;
            if (j >= F) continue;
            if (bad) {
                sm.dbase[2 * j] = sm.dbase[2 * j + 1] = OP_NOWRITE;
            } else if constexpr (EXACT) {
                sm.dbase[2 * j] = s.dst + sm.dbase[2 * j] + pre0[k];
                sm.dbase[2 * j + 1] = s.dst + sm.dbase[2 * j + 1] + pre1[k];
            }
        }
        if (EXACT && t == 0) sm.tk = tk;
        __syncthreads();
        if constexpr (EXACT) prefetch(sm.tk);  // loads in flight while this chunk writes

        // ---- C: write each bucket's run
        const uint64_t seg_end = EXACT ? s.dst + s.size : A.dst_limit;
        for (uint32_t i = t; i < cnt; i += P) {
            const uint32_t bkt = sm.stb[i];
            const uint64_t d = sm.dbase[bkt];
            if (d != OP_NOWRITE) {
                const uint64_t o = d + (i - sm.hist[bkt]);
                if (o < seg_end && o < A.dst_limit) {
                    dst[o] = sm.stg[i];
                    if constexpr (KV) vdst[o] = sm.stv[i];
                } else {
                    atomicOr(&A.flags[OP_FLAG_RANGE], 1u);
                }
            }
        }
        if (!more) break;
        __syncthreads();  // this chunk is done with LDS
        if (newseg) {
            const uint32_t Fn = 1u << sn.logf, Gn = guide_words(Fn);
#pragma unroll
            for (int k = 0; k < GPT; ++k) {
                const uint32_t i = t + k * P;
                if (i < Gn) sm.lguide[i] = gpf[k];
            }
#pragma unroll
            for (int k = 0; k < SPT; ++k) {
                const uint32_t i = t + k * P;
                if (i < Fn) sm.lsrt[i] = spf[k];
            }
            g = make_guide(sn.kmin, sn.shift, Fn, sn.flags);
        }
        c = cn;
        s = sn;
    }
}

// ---------------------------------------------------------------- routing
struct OpPlan {
    TileSeg* tile[8];     // tile sorts per class
    uint32_t tile_cap[8]; // class capacities (keys)
    uint32_t nclass;
    uint32_t* counts;     // [0..7] tiles per class, [8] move pieces, [9] next, [10] leftover, [14] list overflow
    uint64_t* keys;       // [0..7] keys per tile class, [8] keys moved
    MovePiece* move;
    OpNext* next;
    Seg* left;
    uint32_t cap[11];     // entries each list holds
};

// Route every bucket of a one-pass level.  A bucket's final position is the
// segment's s.fin plus the sizes of the buckets before it.  Output in B
// (OUT_B) or in the caller's array A:
//   gapped (out B): equality / single key -> move-fill to A; ordinary up to a
//                   tile -> move to A, tile in place; larger -> next level
//   exact, out B:   equality / single -> move-fill to A; ordinary up to a tile
//                   -> tile B -> A; larger -> next level
//   exact, out A:   equality / single -> done; ordinary up to a tile -> tile
//                   in place; larger -> leftover (multi-level engine)
template <bool EXACT, bool OUT_B>
__global__ __launch_bounds__(BMAX) void k_op_plan(const PartSeg* ps, const uint32_t* nseg, const uint64_t* pstart,
                                                  const uint64_t* pcap, const uint64_t* bsize, OpPlan pl) {
    constexpr int NL = 11;  // tile classes 0..7, move 8, next 9, leftover 10
    __shared__ uint64_t arr[BMAX];
    __shared__ uint64_t wt[BMAX / 64];
    __shared__ uint32_t lcnt[NL], lbase[NL];
    __shared__ unsigned long long lkeys[NL];
    for (uint32_t w = blockIdx.x; w < *nseg; w += gridDim.x) {
        const PartSeg s = ps[w];
        const uint32_t B = 2u << s.logf, b = threadIdx.x;
        if (b < NL) {
            lcnt[b] = 0;
            lkeys[b] = 0;
        }
        const uint64_t sz = b < B ? bsize[s.bkt_base + b] : 0;
        arr[b] = sz;
        // gapped: a pair whose two buckets together outgrew the region had
        // runs written over each other (runs past the region were not written)
        if (!EXACT && b < B && !(b & 1) && sz + bsize[s.bkt_base + b + 1] > pcap[s.tree_base + (b >> 1)])
            atomicOr(&pl.counts[15], 1u);
        __syncthreads();
        op_bucket_scan<BMAX, uint64_t>(arr, B, wt);
        const uint64_t fin = s.fin + arr[b];
        uint64_t src = fin;
        if (!EXACT && b < B) {
            const uint64_t p0 = pstart[s.tree_base + (b >> 1)];
            src = (b & 1) ? p0 + pcap[s.tree_base + (b >> 1)] - sz : p0;
        }
        int l1 = -1, l2 = -1;  // up to two lists per bucket (move + tile)
        uint32_t np = 1;
        if (b < B && sz) {
            const bool small = !(b & 1) && sz > 1;
            int cls = -1;
            if (small)
                for (uint32_t k = 0; k < pl.nclass; ++k)
                    if (sz <= pl.tile_cap[k]) { cls = (int)k; break; }
            if (!small) {
                if (!EXACT || OUT_B) l1 = 8;
            } else if (cls >= 0) {
                l2 = cls;
                if (!EXACT) l1 = 8;
            } else {
                l1 = (EXACT && !OUT_B) ? 10 : 9;
            }
            if (l1 == 8) np = (uint32_t)((sz + COPY_PIECE - 1) / COPY_PIECE);
        }
        uint32_t r1 = 0, r2 = 0;
        if (l1 >= 0) {
            r1 = atomicAdd(&lcnt[l1], l1 == 8 ? np : 1u);
            if (l1 == 8) atomicAdd(&lkeys[8], (unsigned long long)sz);
        }
        if (l2 >= 0) {
            r2 = atomicAdd(&lcnt[l2], 1u);
            atomicAdd(&lkeys[l2], (unsigned long long)sz);
        }
        __syncthreads();
        if (b < NL && lcnt[b]) {
            lbase[b] = atomicAdd(&pl.counts[b], lcnt[b]);
            if (b <= 8) atomicAdd((unsigned long long*)&pl.keys[b], lkeys[b]);
            if ((uint64_t)lbase[b] + lcnt[b] > pl.cap[b]) atomicOr(&pl.counts[14], 1u);
        }
        __syncthreads();
        if (l1 == 8 && (uint64_t)lbase[8] + r1 + np <= pl.cap[8]) {
            const uint32_t fill = (b & 1) || sz == 1 ? 1u : 0u;
            for (uint32_t p = 0; p < np; ++p) {
                const uint64_t pb = (uint64_t)p * COPY_PIECE;
                const uint32_t len = (uint32_t)((sz - pb) < COPY_PIECE ? (sz - pb) : COPY_PIECE);
                pl.move[lbase[8] + r1 + p] = MovePiece{src + pb, fin + pb, len, fill};
            }
        } else if (l1 == 9 && lbase[9] + r1 < pl.cap[9]) {
            pl.next[lbase[9] + r1] = OpNext{src, fin, sz};
        } else if (l1 == 10 && lbase[10] + r1 < pl.cap[10]) {
            pl.left[lbase[10] + r1] = Seg{fin, sz, 0u, 0u};
        }
        if (l2 >= 0 && lbase[l2] + r2 < pl.cap[l2])
            pl.tile[l2][lbase[l2] + r2] = TileSeg{fin, (uint32_t)sz, (EXACT && OUT_B) ? 1u : 0u};
        __syncthreads();
    }
}

template <class T, bool KV>
__global__ __launch_bounds__(256) void k_move(const MovePiece* pc, const uint32_t* count, const T* from, T* to,
                                              const int32_t* vfrom, int32_t* vto) {
    const uint32_t n = *count;
    for (uint32_t p = blockIdx.x; p < n; p += gridDim.x) {
        const MovePiece m = pc[p];
        if (m.fill) {
            const T key = from[m.src];
            for (uint32_t i = threadIdx.x; i < m.size; i += 256) to[m.dst + i] = key;
        } else {
            for (uint32_t i = threadIdx.x; i < m.size; i += 256) to[m.dst + i] = from[m.src + i];
        }
        if constexpr (KV)
            for (uint32_t i = threadIdx.x; i < m.size; i += 256) vto[m.dst + i] = vfrom[m.src + i];
    }
}

}  // namespace s512
