// s512_partition.hip -- Partition512 (sort512.hpp:5729-5910, sort512kv.hpp:
// 5958-6078): one read and one write of every key, plus one copy of the
// smaller side.
//
// Contract kept from the reference: the inclusive range [left, right] is
// rearranged so that every key `<= pivot` (int32: signed compare; double: the
// IEEE `<=`, so NaN goes right, as _CMP_LE_OQ does) precedes every other key,
// and the function returns left + #{keys <= pivot}.  The arrangement inside
// each side is unspecified in the reference; here it is stable (both sides keep
// their input order), which is one of the arrangements the reference allows.
//
// The reference reads and writes every key once per pass (sort512.hpp:
// 5748-5772, 5840-5864): it compress-stores each vector's low keys at the left
// write front and its high keys at the right one.  On the GPU:
//   k_part_sample    one block: the share of keys <= pivot in 4096 sampled
//                    keys picks the direction (the larger side stays in place)
//   k_part_pass      chunks in ticket order (from the left when the low side
//                    is larger, else from the right): wave ballots rank every
//                    key inside the chunk; a wave-wide decoupled look-back
//                    over two status words per chunk (low and high counts of
//                    the chunks before it) gives the chunk's offsets.  The
//                    larger side is written in place -- a chunk's keys of
//                    that side land at or before (from the right: at or
//                    after) its own position, over keys that every earlier
//                    chunk has already read -- and the smaller side to
//                    scratch.
//   k_part_copyback  the smaller side from scratch to its final place.
// HBM traffic per key: 2 x key bytes, plus 2 x key bytes for the smaller side
// (at most 24 B per fp64 key); count keys are never read twice.
#include "s512_common.hpp"
#include "s512_runtime.hpp"
#include "s512_hostio.hpp"
#include "../../include/s512.h"
#include <algorithm>
#include <map>
#include <memory>

namespace s512 {

constexpr int PT_P = 256;
constexpr int PT_NW = PT_P / 64;
constexpr uint32_t PT_SPIN_LIMIT = 1u << 20;
template <class T>
struct PartCfg {
    static constexpr int E = sizeof(T) == 8 ? 16 : 32;
    static constexpr uint32_t CH = PT_P * E;
};

// Device control block of one call (zeroed before every call).
struct PartCtl {
    uint32_t ticket;
    uint32_t dir;        // 1: from the left (low side in place), 2: from the right
    uint32_t timeout;    // a look-back wait gave up
    uint32_t pad;
    uint64_t total_low;  // written by the last chunk
    uint64_t total_high;
};

__device__ __forceinline__ uint64_t pt_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void pt_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// status word: 0 = not published, else ((v + 1) << 1) | inclusive
__device__ __forceinline__ uint64_t pt_pack(uint64_t v, bool incl) { return ((v + 1) << 1) | (incl ? 1u : 0u); }
__device__ __forceinline__ uint64_t pt_val(uint64_t w) { return (w >> 1) - 1; }
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, d), hi = __shfl_xor((uint32_t)(v >> 32), d);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

template <class T>
__device__ __forceinline__ bool le_pivot(T x, T p) { return x <= p; }

template <class T>
__global__ __launch_bounds__(256) void k_part_sample(const T* a, uint64_t n, T pivot, PartCtl* ctl) {
    __shared__ uint32_t wsum[4];
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < 4096; i += 256) c += le_pivot(a[(uint64_t)i * n / 4096], pivot) ? 1u : 0u;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) ctl->dir = (wsum[0] + wsum[1] + wsum[2] + wsum[3]) * 2 >= 4096 ? 1u : 2u;
}

// Chunk c (ticket order) covers, from the left, [c CH, (c+1) CH) of the
// range; from the right, [n - (c+1) CH, n - c CH).  `out` holds out_cap keys:
// the smaller side, from its start (left pass: high keys) or its end (right
// pass: low keys).  status: two words per chunk, zero before the launch.
template <class T, bool KV>
__global__ __launch_bounds__(PT_P) void k_part_pass(T* a, int32_t* va, uint64_t n, T pivot, T* out, int32_t* vout,
                                                    uint64_t out_cap, uint64_t* status, PartCtl* ctl,
                                                    uint32_t nchunks) {
    constexpr int E = PartCfg<T>::E;
    constexpr uint32_t CH = PartCfg<T>::CH;
    __shared__ uint32_t tab[E][PT_NW];  // low keys per (slot, wave), then their exclusive prefix
    __shared__ uint32_t s_work, s_low;
    __shared__ uint64_t s_pre[2];
    __shared__ uint32_t s_bad;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    const bool fwd = ctl->dir == 1u;
    for (;;) {
        __syncthreads();  // the previous chunk is done with LDS
        if (t == 0) s_work = atomicAdd(&ctl->ticket, 1u);
        __syncthreads();
        const uint32_t c = s_work;
        if (c >= nchunks) break;  // uniform
        uint64_t lo, hi;
        if (fwd) {
            lo = (uint64_t)c * CH;
            hi = lo + CH < n ? lo + CH : n;
        } else {
            hi = n - (uint64_t)c * CH;
            lo = hi > CH ? hi - CH : 0;
        }
        const uint32_t cnt = (uint32_t)(hi - lo);
        T x[E];
        int32_t xv[KV ? E : 1];
        bool f[E];
        uint64_t bal[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * PT_P + t, ic = i < cnt ? i : cnt - 1;  // in bounds, unconditional
            x[e] = a[lo + ic];
            if constexpr (KV) xv[e] = va[lo + ic];
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            f[e] = (e * PT_P + t < cnt) && le_pivot(x[e], pivot);
            bal[e] = __ballot(f[e]);
            if (lane == 0) tab[e][w] = (uint32_t)__popcll(bal[e]);
        }
        if (t == 0) s_bad = 0;
        __syncthreads();  // every key of the chunk is in registers
        if (t == 0) {
            uint32_t run = 0;
            for (int e = 0; e < E; ++e)
                for (int ww = 0; ww < PT_NW; ++ww) {
                    const uint32_t v = tab[e][ww];
                    tab[e][ww] = run;
                    run += v;
                }
            s_low = run;
        }
        __syncthreads();
        const uint32_t low_c = s_low, high_c = cnt - low_c;
        if (w == 0) {
            uint64_t* me = status + 2 * (uint64_t)c;
            uint64_t pl = 0, ph = 0;
            bool tmo = false;
            if (lane == 0) {
                pt_store(me, pt_pack(low_c, c == 0));
                pt_store(me + 1, pt_pack(high_c, c == 0));
            }
            if (c > 0) {
                // lane i reads chunk q - i; the walk ends at the nearest chunk
                // with an inclusive prefix (chunk 0 always publishes one)
                bool dl = false, dh = false;
                for (int64_t q = (int64_t)c - 1; !(dl && dh) && !tmo; q -= 64) {
                    const int64_t r = q - (int64_t)lane;
                    const bool ok = r >= 0;
                    uint64_t wl = 0, wh = 0;
                    if (ok) {
                        wl = pt_load(status + 2 * r);
                        wh = pt_load(status + 2 * r + 1);
                    }
                    uint64_t ml = 0, mh = 0;
                    bool fl = false, fh = false;
                    for (uint32_t it = 1;; ++it) {
                        // lanes up to and including the first inclusive one
                        const uint64_t il = __ballot(!ok || (wl & 1)), ih = __ballot(!ok || (wh & 1));
                        const uint64_t bl = il & (~il + 1), bh = ih & (~ih + 1);
                        fl = il != 0;
                        fh = ih != 0;
                        ml = fl ? (bl | (bl - 1)) : ~0ull;
                        mh = fh ? (bh | (bh - 1)) : ~0ull;
                        const uint64_t pend = (dl ? 0ull : (__ballot(ok && wl == 0) & ml)) |
                                              (dh ? 0ull : (__ballot(ok && wh == 0) & mh));
                        if (!pend) break;
                        if ((it & 255u) == 0) {
                            if (it >= PT_SPIN_LIMIT && lane == 0) atomicOr(&ctl->timeout, 1u);
                            if (__hip_atomic_load(&ctl->timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                                tmo = true;
                                break;
                            }
                        }
                        __builtin_amdgcn_s_sleep(2);
                        if (ok && wl == 0) wl = pt_load(status + 2 * r);
                        if (ok && wh == 0) wh = pt_load(status + 2 * r + 1);
                    }
                    if (tmo) break;
                    if (!dl) {
                        pl += wave_sum64(ok && ((ml >> lane) & 1) ? pt_val(wl) : 0ull);
                        dl = fl;
                    }
                    if (!dh) {
                        ph += wave_sum64(ok && ((mh >> lane) & 1) ? pt_val(wh) : 0ull);
                        dh = fh;
                    }
                }
                if (lane == 0 && !tmo) {
                    pt_store(me, pt_pack(pl + low_c, true));
                    pt_store(me + 1, pt_pack(ph + high_c, true));
                }
            }
            if (lane == 0) {
                s_pre[0] = pl;
                s_pre[1] = ph;
                if (tmo) s_bad = 1;
                if (c + 1 == nchunks) {
                    ctl->total_low = pl + low_c;
                    ctl->total_high = ph + high_c;
                }
            }
        }
        __syncthreads();
        if (s_bad) continue;  // nothing is written by a chunk without its offsets
        const uint64_t PL = s_pre[0], PH = s_pre[1];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * PT_P + t;
            if (i >= cnt) continue;
            const uint32_t lr = tab[e][w] + (uint32_t)__popcll(bal[e] & lt_mask);  // low keys before slot i
            const uint32_t hr = i - lr;                                             // high keys before slot i
            uint64_t d;
            bool in_place;
            if (fwd) {
                in_place = f[e];
                d = f[e] ? PL + lr : PH + hr;
            } else {
                in_place = !f[e];
                d = f[e] ? out_cap - PL - low_c + lr : n - PH - high_c + hr;
            }
            if (in_place) {
                if (d < n) {
                    a[d] = x[e];
                    if constexpr (KV) va[d] = xv[e];
                }
            } else if (d < out_cap) {
                out[d] = x[e];
                if constexpr (KV) vout[d] = xv[e];
            }
        }
    }
}

// The smaller side from scratch to its place: after the low side (left pass)
// or at the start of the range (right pass).
template <class T, bool KV>
__global__ __launch_bounds__(256) void k_part_copyback(T* a, int32_t* va, uint64_t n, const T* out,
                                                       const int32_t* vout, uint64_t out_cap, const PartCtl* ctl) {
    if (ctl->timeout) return;
    const bool fwd = ctl->dir == 1u;
    const uint64_t m = fwd ? ctl->total_high : ctl->total_low;
    const uint64_t src = fwd ? 0 : out_cap - m, dst = fwd ? ctl->total_low : 0;
    if (m > out_cap || dst + m > n) return;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256ull) {
        a[dst + i] = out[src + i];
        if constexpr (KV) va[dst + i] = vout[src + i];
    }
}

__global__ void k_part_zero(uint64_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Bytes of workspace a partition of n keys needs: control block, status
// words, then the scratch side (n keys, and n values for key/value).
template <class T, bool KV>
static size_t part_workspace(uint64_t n, uint32_t* nchunks, size_t* st_off, size_t* key_off, size_t* val_off) {
    const uint32_t nc = (uint32_t)((n + PartCfg<T>::CH - 1) / PartCfg<T>::CH);
    size_t off = 256;  // PartCtl
    *st_off = off;
    off += ((size_t)nc * 2 * sizeof(uint64_t) + 255) & ~(size_t)255;
    *key_off = off;
    off += (n * sizeof(T) + 255) & ~(size_t)255;
    *val_off = off;
    if (KV) off += (n * sizeof(int32_t) + 255) & ~(size_t)255;
    *nchunks = nc;
    return off;
}

// Partition of a[left..right] (device memory) with the workspace `ws` (at
// least part_workspace() bytes); *out_bound = left + #{<= pivot}.  Reads the
// bound back (synchronises the stream); calls no allocator.
template <class T, bool KV>
static s512_status partition_run(T* a, int32_t* va, size_t left, size_t right, T pivot, size_t* out_bound,
                                 hipStream_t st, char* ws, uint64_t* host_word) {
    const uint64_t n = (uint64_t)right - left + 1;
    uint32_t nc;
    size_t st_off, key_off, val_off;
    part_workspace<T, KV>(n, &nc, &st_off, &key_off, &val_off);
    PartCtl* ctl = (PartCtl*)ws;
    uint64_t* status = (uint64_t*)(ws + st_off);
    T* out = (T*)(ws + key_off);
    int32_t* vout = KV ? (int32_t*)(ws + val_off) : nullptr;
    T* src = a + left;
    int32_t* vsrc = KV ? va + left : nullptr;
    int dev = 0, cus = 0, occ = 0;
    S512_HIP(hipGetDevice(&dev));
    S512_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    S512_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_part_pass<T, KV>, PT_P, 0));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nc, (uint64_t)std::max(occ, 1) * cus);
    // control block and status words start at zero (a fill kernel: callers may capture this)
    const uint64_t zw = (st_off + (size_t)nc * 2 * sizeof(uint64_t)) / sizeof(uint64_t);
    k_part_zero<<<(unsigned)std::min<uint64_t>((zw + 255) / 256, 1024), 256, 0, st>>>((uint64_t*)ws, zw);
    S512_CHECK_LAUNCH();
    k_part_sample<T><<<1, 256, 0, st>>>(src, n, pivot, ctl);
    S512_CHECK_LAUNCH();
    k_part_pass<T, KV><<<grid, PT_P, 0, st>>>(src, vsrc, n, pivot, out, vout, n, status, ctl, nc);
    S512_CHECK_LAUNCH();
    k_part_copyback<T, KV><<<(unsigned)std::min<uint64_t>((n + 255) / 256, 4096), 256, 0, st>>>(src, vsrc, n, out,
                                                                                               vout, n, ctl);
    S512_CHECK_LAUNCH();
    uint64_t res[2];
    S512_HIP(hipMemcpyAsync(host_word, &ctl->timeout, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(host_word + 1, &ctl->total_low, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipStreamSynchronize(st));
    memcpy(res, host_word, sizeof(res));
    if ((uint32_t)res[0]) {
        set_error("Partition512: a look-back wait did not end (the range is left partly rearranged)");
        return S512_ERR_INTERNAL;
    }
    *out_bound = left + (size_t)res[1];
    return S512_OK;
}

template <class T, bool KV>
static s512_status partition_dev(T* a, int32_t* va, size_t left, size_t right, T pivot, size_t* out_bound,
                                 hipStream_t st) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }  // empty range (right = left-1)
    if (!a || (KV && !va)) { set_error("null array"); return S512_ERR_INVALID; }
    const uint64_t n = (uint64_t)right - left + 1;
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    uint32_t nc;
    size_t st_off, key_off, val_off;
    const size_t bytes = part_workspace<T, KV>(n, &nc, &st_off, &key_off, &val_off);
    if ((rc = ds.data.reserve(bytes, st))) return rc;
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    return partition_run<T, KV>(a, va, left, right, pivot, out_bound, st, (char*)ds.data.ptr,
                                (uint64_t*)ds.pinned2.ptr);
}

// ---------------------------------------------------------------- no allocation
// A fixed per-device workspace for callers that cannot allocate (code under
// hipGraph capture, serving loops, harnesses that forbid hipMalloc).  The
// array lives in every GPU the library runs on until the process ends; its
// device address is looked up once per device with hipGetSymbolAddress (the
// name itself is a host-side stand-in, not a device address).  Calls on one
// GPU share it: calls that overlap on one GPU race.
constexpr size_t PART_SCRATCH_BYTES = 64u << 20;
static __device__ __align__(256) char scratch[PART_SCRATCH_BYTES];

struct ScratchSlot {
    std::mutex mu;
    char* ptr = nullptr;
    uint64_t* host = nullptr;  // the call's result words (ordinary host memory, owned here)
    uint64_t words[4];
};
static ScratchSlot& scratch_slot(int dev) {
    static std::mutex map_mu;
    static std::map<int, std::unique_ptr<ScratchSlot>> slots;
    std::lock_guard<std::mutex> lk(map_mu);
    auto& p = slots[dev];
    if (!p) p.reset(new ScratchSlot());
    return *p;
}

// The largest range whose workspace fits the fixed array (8,384,480 fp64 keys).
template <class T, bool KV>
static uint64_t part_max_keys() {
    uint64_t lo = 0, hi = PART_SCRATCH_BYTES;  // part_workspace(lo) fits, part_workspace(hi) does not
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        uint32_t nc;
        size_t s, k, v;
        (part_workspace<T, KV>(mid, &nc, &s, &k, &v) <= PART_SCRATCH_BYTES ? lo : hi) = mid;
    }
    return lo;
}

template <class T, bool KV>
static s512_status partition_noalloc(T* a, int32_t* va, size_t left, size_t right, T pivot, size_t* out_bound,
                                     hipStream_t st) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!a || (KV && !va)) { set_error("null array"); return S512_ERR_INVALID; }
    const uint64_t n = (uint64_t)right - left + 1;
    uint32_t nc;
    size_t st_off, key_off, val_off;
    const size_t bytes = part_workspace<T, KV>(n, &nc, &st_off, &key_off, &val_off);
    if (bytes > PART_SCRATCH_BYTES) {
        set_error("Partition512 without allocation: %llu keys need %zu bytes of workspace, the fixed one holds %zu "
                  "(at most %llu keys of this type)",
                  (unsigned long long)n, bytes, PART_SCRATCH_BYTES, (unsigned long long)part_max_keys<T, KV>());
        return S512_ERR_TOO_LARGE;
    }
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    ScratchSlot& sl = scratch_slot(dev);
    std::lock_guard<std::mutex> lk(sl.mu);
    if (!sl.ptr) {
        void* p = nullptr;
        S512_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        sl.ptr = (char*)p;
        sl.host = sl.words;
    }
    return partition_run<T, KV>(a, va, left, right, pivot, out_bound, st, sl.ptr, sl.host);
}

}  // namespace s512

using namespace s512;

extern "C" s512_status s512_partition_i32_dev(int32_t* d_a, size_t left, size_t right, int32_t pivot,
                                              size_t* out_bound, void* stream) {
    return partition_dev<int32_t, false>(d_a, nullptr, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512_partition_f64_dev(double* d_a, size_t left, size_t right, double pivot,
                                              size_t* out_bound, void* stream) {
    return partition_dev<double, false>(d_a, nullptr, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512kv_partition_i32_dev(int32_t* d_k, int32_t* d_v, size_t left, size_t right, int32_t pivot,
                                                size_t* out_bound, void* stream) {
    return partition_dev<int32_t, true>(d_k, d_v, left, right, pivot, out_bound, (hipStream_t)stream);
}
extern "C" s512_status s512_partition_f64_dev_noalloc(double* d_a, size_t left, size_t right, double pivot,
                                                      size_t* out_bound, void* stream) {
    return partition_noalloc<double, false>(d_a, nullptr, left, right, pivot, out_bound, (hipStream_t)stream);
}

// Host arrays: only [left, right] is staged; the bound is shifted back.
extern "C" s512_status s512_partition_i32(int32_t* a, size_t left, size_t right, int32_t pivot, size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    const size_t n = right - left + 1;
    return with_staged_array(a + left, n, [&](int32_t* d, hipStream_t st) {
        size_t b = 0;
        s512_status rc = s512_partition_i32_dev(d, 0, n - 1, pivot, &b, st);
        *out_bound = left + b;
        return rc;
    });
}
extern "C" s512_status s512_partition_f64(double* a, size_t left, size_t right, double pivot, size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!a) { set_error("null array"); return S512_ERR_INVALID; }
    const size_t n = right - left + 1;
    return with_staged_array(a + left, n, [&](double* d, hipStream_t st) {
        size_t b = 0;
        s512_status rc = s512_partition_f64_dev(d, 0, n - 1, pivot, &b, st);
        *out_bound = left + b;
        return rc;
    });
}
extern "C" s512_status s512kv_partition_i32(int32_t* k, int32_t* v, size_t left, size_t right, int32_t pivot,
                                            size_t* out_bound) {
    if (!out_bound) { set_error("null out_bound"); return S512_ERR_INVALID; }
    if (right + 1 <= left) { *out_bound = left; return S512_OK; }
    if (!k || !v) { set_error("null array"); return S512_ERR_INVALID; }
    const size_t n = right - left + 1;
    return with_staged_pair(k + left, v + left, n, [&](int32_t* dk, int32_t* dv, hipStream_t st) {
        size_t b = 0;
        s512_status rc = s512kv_partition_i32_dev(dk, dv, 0, n - 1, pivot, &b, st);
        *out_bound = left + b;
        return rc;
    });
}
