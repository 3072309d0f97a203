// s512_tile.hpp -- register/LDS bitonic sorting networks for wave64 and
// workgroups (the MI355X counterpart of the reference's SmallSort16V networks,
// sort512.hpp:46-2633 and 2635-5688).
//
// Index model.  A tile of N = P*E keys is sorted by a bitonic network; key
// index i has log2(N) bits.  Each thread keeps E keys in VGPRs.  A "layout"
// says which log2(E) consecutive index bits (the window [w, w+log2 E)) live in
// the register number; the remaining bits are the thread id, in order.  A
// compare-exchange along a bit inside the window is a pure register operation;
// moving to another window is one LDS round trip (write, barrier, read).  The
// blocked layout (w = 0) holds keys t*E .. t*E+E-1 in thread t.
//
// Wave-level sorts (N <= 64*E, one wave) exchange across lanes instead:
//   lane xor 1, 2  -> DPP quad_perm          (VALU, no LDS)
//   lane xor 8     -> DPP row_ror:8          (VALU, no LDS)
//   lane xor 4, 16 -> ds_swizzle bit-mask    (LDS crossbar, no memory)
//   lane xor 32    -> ds_bpermute
#pragma once
#include "s512_common.hpp"

namespace s512 {

// ---------------------------------------------------------------- compare
template <class K>
__device__ __forceinline__ void cex_dir(K& a, K& b, bool desc) {
    // afterwards a <= b (ascending pair) or a >= b (descending pair); for
    // a == b swapping or not gives the same pair, so one compare suffices
    const bool sw = (b < a) != desc;
    const K x = sw ? b : a;
    const K y = sw ? a : b;
    a = x;
    b = y;
}
template <class K>
__device__ __forceinline__ void cex_up(K& a, K& b) {
    const bool sw = b < a;
    const K x = sw ? b : a;
    const K y = sw ? a : b;
    a = x;
    b = y;
}

// ---------------------------------------------------------------- lane exchange
template <int M>
__device__ __forceinline__ uint32_t xor_lane32(uint32_t v) {
    if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
    else if constexpr (M == 4 || M == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (M << 10) | 0x1F);
    else {
        static_assert(M == 32, "lane distance");
        const int lane = (int)__lane_id();
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)v);
    }
}
template <int M>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) { return xor_lane32<M>(v); }
template <int M>
__device__ __forceinline__ uint64_t xor_lane(uint64_t v) {
    const uint32_t lo = xor_lane32<M>((uint32_t)v);
    const uint32_t hi = xor_lane32<M>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------- in-register stages
// Stage along register bit PB of all E registers, direction per register given
// by bit q of the element index (i0 | r << w).
template <class K, int E, int PB>
__device__ __forceinline__ void reg_stage(K (&v)[E], uint32_t i0, int w, int q, int logn) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
        if (r & (1 << PB)) continue;
        const uint32_t ir = i0 | ((uint32_t)r << w);
        const bool desc = (q < logn) && ((ir >> q) & 1u);
        cex_dir(v[r], v[r | (1 << PB)], desc);
    }
}
template <class K, int E, int PB>
__device__ __forceinline__ void reg_stage_u(K (&v)[E], bool desc) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
        if (r & (1 << PB)) continue;
        cex_dir(v[r], v[r | (1 << PB)], desc);
    }
}
template <class K, int E>
__device__ __forceinline__ void reg_stage_any(K (&v)[E], int pb, uint32_t i0, int w, int q, int logn) {
    if constexpr (E >= 2) { if (pb == 0) { reg_stage<K, E, 0>(v, i0, w, q, logn); return; } }
    if constexpr (E >= 4) { if (pb == 1) { reg_stage<K, E, 1>(v, i0, w, q, logn); return; } }
    if constexpr (E >= 8) { if (pb == 2) { reg_stage<K, E, 2>(v, i0, w, q, logn); return; } }
    if constexpr (E >= 16) { if (pb == 3) { reg_stage<K, E, 3>(v, i0, w, q, logn); return; } }
    if constexpr (E >= 32) { if (pb == 4) { reg_stage<K, E, 4>(v, i0, w, q, logn); return; } }
}
// Low bits of the blocked layout with a direction uniform over the registers.
template <class K, int E>
__device__ __forceinline__ void reg_low_stages(K (&v)[E], bool desc) {
    if constexpr (E >= 32) reg_stage_u<K, E, 4>(v, desc);
    if constexpr (E >= 16) reg_stage_u<K, E, 3>(v, desc);
    if constexpr (E >= 8) reg_stage_u<K, E, 2>(v, desc);
    if constexpr (E >= 4) reg_stage_u<K, E, 1>(v, desc);
    if constexpr (E >= 2) reg_stage_u<K, E, 0>(v, desc);
}
// Bitonic sort of the E registers of one thread: merge sizes 2..E; the
// direction of merge size 2^q is bit q of the global index t*E + r, i.e. a
// register bit for q < log2 E and bit 0 of `tbit` for q == log2 E.
template <class K, int E>
__device__ __forceinline__ void reg_sort(K (&v)[E], uint32_t tbit) {
    constexpr int LOGE = ilog2(E);
#pragma unroll
    for (int q = 1; q <= LOGE; ++q) {
#pragma unroll
        for (int b = q - 1; b >= 0; --b) {
#pragma unroll
            for (int r = 0; r < E; ++r) {
                if (r & (1 << b)) continue;
                const bool desc = (q < LOGE) ? (((r >> q) & 1) != 0) : ((tbit & 1u) != 0);
                cex_dir(v[r], v[r | (1 << b)], desc);
            }
        }
    }
}

// ---------------------------------------------------------------- workgroup tile
template <class K, int P, int E>
struct BlockBitonic {
    static constexpr int LOGE = ilog2(E);
    static constexpr int LOGP = ilog2(P);
    static constexpr int N = P * E;
    static constexpr int LOGN = LOGE + LOGP;
    static_assert((1 << LOGE) == E && (1 << LOGP) == P, "powers of two");
    static_assert(LOGN >= 2 * LOGE, "window layouts need P >= E");
    static constexpr size_t LDS_BYTES = sizeof(K) * N;

    __device__ static __forceinline__ uint32_t idx(uint32_t t, uint32_t r, int w) {
        const uint32_t lm = (1u << w) - 1u;
        return ((t & ~lm) << LOGE) | (r << w) | (t & lm);
    }
    // XOR swizzle so that the blocked layout does not put a wave on one bank.
    __device__ static __forceinline__ uint32_t phys(uint32_t i) { return i ^ ((i >> LOGE) & (E - 1)); }

    template <int WF, int WT>
    __device__ static __forceinline__ void relayout(K (&v)[E], K* lds, uint32_t t) {
#pragma unroll
        for (int r = 0; r < E; ++r) lds[phys(idx(t, r, WF))] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = lds[phys(idx(t, r, WT))];
        __syncthreads();
    }

    // Stages on index bits B, B-1, .., W of merge size 2^Q while the window
    // [W, W+LOGE) is in the registers.  The direction of a pair is bit Q of
    // its index: a register bit when Q lies inside the window, otherwise bit
    // Q-LOGE of the thread id (or ascending for the final merge Q == LOGN).
    template <int Q, int W, int B>
    __device__ static __forceinline__ void win_stages(K (&v)[E], uint32_t t) {
        if constexpr (B >= W) {
            constexpr int PB = B - W;
            if constexpr (Q < W + LOGE) {
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (r & (1 << PB)) continue;
                    if ((r >> (Q - W)) & 1) cex_up(v[r | (1 << PB)], v[r]);
                    else cex_up(v[r], v[r | (1 << PB)]);
                }
            } else if constexpr (Q == LOGN) {
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (r & (1 << PB)) continue;
                    cex_up(v[r], v[r | (1 << PB)]);
                }
            } else {
                const bool desc = ((t >> (Q - LOGE)) & 1u) != 0;
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (r & (1 << PB)) continue;
                    cex_dir(v[r], v[r | (1 << PB)], desc);
                }
            }
            win_stages<Q, W, B - 1>(v, t);
        }
    }

    // All stages of merge size 2^Q above the register bits, window by window
    // from the top bit down, then back to the blocked layout for the rest.
    template <int Q, int HI, int WPREV>
    __device__ static __forceinline__ void windows(K (&v)[E], K* lds, uint32_t t) {
        if constexpr (HI >= LOGE) {
            constexpr int NW = (HI - LOGE + 1) > LOGE ? (HI - LOGE + 1) : LOGE;
            relayout<WPREV, NW>(v, lds, t);
            win_stages<Q, NW, HI>(v, t);
            windows<Q, NW - 1, NW>(v, lds, t);
        } else {
            relayout<WPREV, 0>(v, lds, t);
            win_stages<Q, 0, LOGE - 1>(v, t);
        }
    }

    template <int Q>
    __device__ static __forceinline__ void merges(K (&v)[E], K* lds, uint32_t t) {
        if constexpr (Q <= LOGN) {
            windows<Q, Q - 1, 0>(v, lds, t);
            merges<Q + 1>(v, lds, t);
        }
    }

    // v holds the blocked layout on entry and on exit; sorted ascending.
    __device__ static void sort(K (&v)[E], K* lds, uint32_t t) {
        reg_sort<K, E>(v, t);
        merges<LOGE + 1>(v, lds, t);
    }

    // striped global -> blocked registers, through LDS
    __device__ static __forceinline__ void from_striped(K (&v)[E], K* lds, uint32_t t) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = lds[phys(t * E + r)];
        __syncthreads();
    }
    __device__ static __forceinline__ void to_striped(const K (&v)[E], K* lds, uint32_t t) {
#pragma unroll
        for (int r = 0; r < E; ++r) lds[phys(t * E + r)] = v[r];
        __syncthreads();
    }
};

// ---------------------------------------------------------------- wave tile
template <class K, int E>
struct WaveBitonic {
    static constexpr int LOGE = ilog2(E);
    static constexpr int N = 64 * E;
    static constexpr int LOGN = LOGE + 6;

    template <int M>
    __device__ static __forceinline__ void lane_stage(K (&v)[E], uint32_t lane, bool desc) {
        const bool upper = (lane & M) != 0;
        const bool keep_min = (upper == desc);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const K p = xor_lane<M>(v[r]);
            const bool take = keep_min ? (p < v[r]) : (v[r] < p);
            v[r] = take ? p : v[r];
        }
    }
    __device__ static __forceinline__ void lane_stage_any(K (&v)[E], int lb, uint32_t lane, bool desc) {
        switch (lb) {
            case 0: lane_stage<1>(v, lane, desc); break;
            case 1: lane_stage<2>(v, lane, desc); break;
            case 2: lane_stage<4>(v, lane, desc); break;
            case 3: lane_stage<8>(v, lane, desc); break;
            case 4: lane_stage<16>(v, lane, desc); break;
            default: lane_stage<32>(v, lane, desc); break;
        }
    }
    // blocked layout: lane l holds keys l*E .. l*E+E-1; sorted ascending.
    __device__ static void sort(K (&v)[E], uint32_t lane) {
        reg_sort<K, E>(v, lane);
#pragma unroll
        for (int q = LOGE + 1; q <= LOGN; ++q) {
            const bool desc = (q < LOGN) && (((lane >> (q - LOGE)) & 1u) != 0);
#pragma unroll
            for (int b = q - 1; b >= LOGE; --b) lane_stage_any(v, b - LOGE, lane, desc);
            reg_low_stages<K, E>(v, desc);
        }
    }
};

}  // namespace s512
