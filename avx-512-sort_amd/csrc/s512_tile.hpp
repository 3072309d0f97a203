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

// ---------------------------------------------------------------- ordering policies
// A policy supplies the compare-exchange (ascending) and the order-reversing
// "flip" used for descending blocks: every merge phase flips the keys of the
// blocks that the bitonic network sorts descending, runs ascending
// compare-exchanges only, and unflips at the end of the phase.
template <class K>
struct UintOrder {  // encoded keys: bitwise NOT reverses the unsigned order
    using key_t = K;
    static constexpr bool SIGNED_LANES = false;
    __device__ static __forceinline__ void ce(K& a, K& b) { cex_up(a, b); }
    __device__ static __forceinline__ K flip(K x, bool f) { return f ? (K)~x : x; }
    __device__ static __forceinline__ bool less(K a, K b) { return a < b; }
};
// Raw doubles that contain no NaN and no -0.0 (the caller checks): IEEE
// min/max give the exact ascending pair, negation reverses the order.  Only
// compiled into translation units built with -fno-honor-nans, where min/max
// lower to single v_min_f64 / v_max_f64 instructions.
struct F64Order {
    using key_t = double;
    // lane stages on negated keys (BlockBitonic::signed_lane_phase)
    static constexpr bool SIGNED_LANES = true;
    __device__ static __forceinline__ double min_neg(double a, double b) { return __builtin_fmin(a, -b); }
    __device__ static __forceinline__ void ce(double& a, double& b) {
        const double lo = __builtin_fmin(a, b), hi = __builtin_fmax(a, b);
        a = lo;
        b = hi;
    }
    __device__ static __forceinline__ double flip(double x, bool f) {
        const uint64_t u = __builtin_bit_cast(uint64_t, x) ^ (f ? 0x8000000000000000ull : 0ull);
        return __builtin_bit_cast(double, u);
    }
    __device__ static __forceinline__ bool less(double a, double b) { return a < b; }
};

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
template <int M>
__device__ __forceinline__ double xor_lane(double v) {
    return __builtin_bit_cast(double, xor_lane<M>(__builtin_bit_cast(uint64_t, v)));
}

// ---------------------------------------------------------------- in-register stages
// Flip every register of this thread (all of them belong to blocks of the same
// direction in the layouts where this is called).
template <class Pol, int E>
__device__ __forceinline__ void flip_all(typename Pol::key_t (&v)[E], bool f) {
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = Pol::flip(v[r], f);
}
// Ascending compare-exchange along register bit PB.
template <class Pol, int E, int PB>
__device__ __forceinline__ void reg_stage_up(typename Pol::key_t (&v)[E]) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
        if (r & (1 << PB)) continue;
        Pol::ce(v[r], v[r | (1 << PB)]);
    }
}
// Stages PB, PB-1, .., 0 ascending.
template <class Pol, int E, int PB>
__device__ __forceinline__ void reg_low_up(typename Pol::key_t (&v)[E]) {
    if constexpr (PB >= 0) {
        reg_stage_up<Pol, E, PB>(v);
        reg_low_up<Pol, E, PB - 1>(v);
    }
}
// Bitonic sort of the E registers of one thread, blocked layout (element
// t*E + r).  Merge sizes 2^q < E have register-determined directions (the
// operands are swapped at compile time); merge size E has direction bit 0 of
// `tbit`, handled by a flip.  On exit the registers are flipped iff
// `flip_last` (the caller's next phase unflips them).
template <class Pol, int E>
__device__ __forceinline__ void reg_sort(typename Pol::key_t (&v)[E], bool flip_last) {
    constexpr int LOGE = ilog2(E);
#pragma unroll
    for (int q = 1; q < LOGE; ++q) {
#pragma unroll
        for (int b = q - 1; b >= 0; --b) {
#pragma unroll
            for (int r = 0; r < E; ++r) {
                if (r & (1 << b)) continue;
                if ((r >> q) & 1) Pol::ce(v[r | (1 << b)], v[r]);
                else Pol::ce(v[r], v[r | (1 << b)]);
            }
        }
    }
    flip_all<Pol, E>(v, flip_last);
    reg_low_up<Pol, E, LOGE - 1>(v);
}

// Ascending stage between lanes l and l^M, every register: the lower lane
// (upper == false) keeps the min.
template <class Pol, int E, int M>
__device__ __forceinline__ void lane_stage_up(typename Pol::key_t (&v)[E], bool upper) {
    using K = typename Pol::key_t;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const K p = xor_lane<M>(v[r]);
        K lo = v[r], hi = p;
        Pol::ce(lo, hi);
        v[r] = upper ? hi : lo;
    }
}
// Stages on lane bits LB, LB-1, .., 0.
template <class Pol, int E, int LB>
__device__ __forceinline__ void lane_stages_down(typename Pol::key_t (&v)[E], uint32_t lane) {
    if constexpr (LB >= 0) {
        lane_stage_up<Pol, E, (1 << LB)>(v, (lane & (1u << LB)) != 0);
        lane_stages_down<Pol, E, LB - 1>(v, lane);
    }
}

// ---------------------------------------------------------------- workgroup tile
//
// Lane phases by default: two.  A third one holds more partner values in
// registers (the 4-byte 8192-key tile drops to one wave per SIMD with it);
// the raw-double tiles, one min per key and lane stage, choose per shape.
constexpr int LANE_PHASES = 2;

// The first LP merge phases above the register bits (LP <= 3) run as lane
// stages instead of LDS windows.  In the blocked layout index bit LOGE + k is
// bit k of the thread's logical id; with LP = 3 the logical id is the thread
// id with bits 2 and 3 swapped, so logical bits 0, 1 and 2 are lane xor 1, 2
// (DPP quad_perm) and 8 (DPP row_ror:8) -- VALU moves, no LDS.  The swap stays
// inside every group of 16 lanes, so the lanes that LDS serves together hold
// the same set of positions as without it.
template <class Pol, int P, int E, int LP = LANE_PHASES>
struct BlockBitonic {
    using K = typename Pol::key_t;
    static constexpr int LOGE = ilog2(E);
    static constexpr int LOGP = ilog2(P);
    static constexpr int N = P * E;
    static constexpr int LOGN = LOGE + LOGP;
    static_assert((1 << LOGE) == E && (1 << LOGP) == P, "powers of two");
    static_assert(LOGN >= 2 * LOGE, "window layouts need P >= E");
    static_assert(LP >= 0 && LP <= 3 && P >= 16, "lane phases");

    __device__ static __forceinline__ uint32_t logical(uint32_t t) {
        if constexpr (LP >= 3) return (t & ~12u) | ((t & 4u) << 1) | ((t & 8u) >> 1);
        else return t;
    }
    // 8-byte keys, 32 per thread: one pad slot after every 32 (see phys)
    static constexpr bool PADDED = sizeof(K) == 8 && E == 32;
    static constexpr int LDS_N = PADDED ? N + N / 32 : N;  // LDS slots

    __device__ static __forceinline__ uint32_t idx(uint32_t t, uint32_t r, int w) {
        const uint32_t lm = (1u << w) - 1u;
        return ((t & ~lm) << LOGE) | (r << w) | (t & lm);
    }
    // LDS position of element i, chosen so that, in every layout the network
    // uses (striped, blocked and every window) and for every register, the
    // lanes of a wave touch distinct banks -- both when 32 consecutive lanes
    // of 8-byte keys (64 lanes of 4-byte keys) are served per cycle and when
    // 16 (32) are.  8-byte keys with 32 per thread: one pad slot per 32,
    // i + i/32.  The pad is a sum of the index's bit fields shifted, so in
    // every layout the position of register r of thread t is base(t) + c(r)
    // with c(r) a compile-time constant (an instruction offset): no address
    // arithmetic per key.  Otherwise a GF(2)-linear bijection (each output bit
    // XORed only with higher input bits), found by exhaustive search over XOR
    // taps; for the configurations below every access is conflict-free in
    // both models.
    __device__ static __forceinline__ uint32_t phys(uint32_t i) {
        if constexpr (PADDED) {
            return i + (i >> 5);
        } else if constexpr (sizeof(K) == 4 && E == 32) {
            return i ^ ((i >> 5) & 63u);
        } else if constexpr (sizeof(K) == 8 && E == 16) {
            return i ^ ((i >> 4) & 31u);
        } else if constexpr (sizeof(K) == 8 && E == 8) {
            return i ^ ((i >> 4) & 3u) ^ (((i >> 6) & 1u) * 12u) ^ (((i >> 7) & 1u) << 4);
        } else if constexpr (sizeof(K) == 8 && E == 4) {
            return i ^ (((i >> 4) & 1u) * 5u) ^ (((i >> 5) & 1u) * 10u) ^ (((i >> 6) & 1u) * 16u);
        } else if constexpr (sizeof(K) == 4 && E == 8) {
            return i ^ ((i >> 5) & 1u) ^ (((i >> 6) & 1u) * 10u) ^ (((i >> 7) & 1u) * 20u) ^ (((i >> 8) & 1u) * 32u);
        } else if constexpr (sizeof(K) == 4 && E == 16) {
            return i ^ ((i >> 5) & 7u) ^ (((i >> 8) & 1u) * 24u) ^ (((i >> 9) & 1u) * 32u);
        } else {
            return i ^ ((i >> LOGE) & (E - 1));
        }
    }

    template <int WF, int WT>
    __device__ static __forceinline__ void relayout(K (&v)[E], K* lds, uint32_t t) {
#pragma unroll
        for (int r = 0; r < E; ++r) lds[phys(idx(t, r, WF))] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = lds[phys(idx(t, r, WT))];
        __syncthreads();
    }

    // Stages on index bits B, B-1, .., W while the window [W, W+LOGE) is in
    // the registers; all ascending (descending blocks are flipped).
    template <int W, int B>
    __device__ static __forceinline__ void win_stages(K (&v)[E]) {
        if constexpr (B >= W) {
            reg_stage_up<Pol, E, B - W>(v);
            win_stages<W, B - 1>(v);
        }
    }

    // All stages of one merge phase above the register bits, window by window
    // from the top bit down, then back to the blocked layout for the rest.
    // (Ending the windows in lane stages once the bits left above the
    // registers are lane bits saves a round trip in some phases, but moved the
    // register allocation and measured within 1 %.)
    template <int HI, int WPREV>
    __device__ static __forceinline__ void windows(K (&v)[E], K* lds, uint32_t t) {
        if constexpr (HI >= LOGE) {
            constexpr int NW = (HI - LOGE + 1) > LOGE ? (HI - LOGE + 1) : LOGE;
            relayout<WPREV, NW>(v, lds, t);
            win_stages<NW, HI>(v);
            windows<NW - 1, NW>(v, lds, t);
        } else {
            relayout<WPREV, 0>(v, lds, t);
            win_stages<0, LOGE - 1>(v);
        }
    }

    // Lane stages on logical thread bits KB, KB-1, .., 0.
    template <int KB>
    __device__ static __forceinline__ void lane_phase(K (&v)[E], uint32_t u) {
        if constexpr (KB >= 0) {
            constexpr int M = (LP >= 3 && KB == 2) ? 8 : (1 << KB);
            lane_stage_up<Pol, E, M>(v, ((u >> KB) & 1u) != 0);
            lane_phase<KB - 1>(v, u);
        }
    }

    // The same for raw doubles, one min per key and stage: during the stage
    // on bit KB the keys of the threads whose bit KB is set (the upper side,
    // which keeps the max) are held negated, and max(a, b) = -min(-a, -b), so
    // both sides take min(own, -partner).  Entered with the keys negated for
    // bit KB; between stages only the threads whose side changes flip, and
    // after bit 0 the keys are plain again.  Both operands of a min are on
    // the same side, so no +0.0 meets a -0.0.
    template <int KB>
    __device__ static __forceinline__ void signed_lane_phase(K (&v)[E], uint32_t u) {
        if constexpr (KB >= 0) {
            constexpr int M = (LP >= 3 && KB == 2) ? 8 : (1 << KB);
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = Pol::min_neg(v[r], xor_lane<M>(v[r]));
            const bool here = ((u >> KB) & 1u) != 0;
            const bool next = KB > 0 && ((u >> (KB > 0 ? KB - 1 : 0)) & 1u) != 0;
            flip_all<Pol, E>(v, here != next);
            signed_lane_phase<KB - 1>(v, u);
        }
    }

    // Blocked layout: the stages on index bits HI..LOGE as lane stages
    // (logical thread bits HI-LOGE..0), then the register bits.  `flip` is a
    // pending flip of every register, folded into the lane stages' own.
    template <int HI>
    __device__ static __forceinline__ void lane_then_regs(K (&v)[E], uint32_t u, bool flip) {
        constexpr int KB = HI - LOGE;
        if constexpr (Pol::SIGNED_LANES) {
            flip_all<Pol, E>(v, flip != (((u >> KB) & 1u) != 0));
            signed_lane_phase<KB>(v, u);
        } else {
            flip_all<Pol, E>(v, flip);
            lane_phase<KB>(v, u);
        }
        reg_low_up<Pol, E, LOGE - 1>(v);
    }

    // Merge phase 2^Q: in the blocked layout the direction of element
    // u*E + r (u: logical thread) is bit Q of its index, i.e. bit Q-LOGE of u
    // (ascending for the final phase).  `prev` says whether the registers are
    // flipped for the previous phase.
    template <int Q>
    __device__ static __forceinline__ void merges(K (&v)[E], K* lds, uint32_t u, bool prev) {
        if constexpr (Q <= LOGN) {
            const bool cur = (Q < LOGN) && (((u >> (Q - LOGE)) & 1u) != 0);
            if constexpr (Q - LOGE <= LP) {
                lane_then_regs<Q - 1>(v, u, prev != cur);
            } else {
                flip_all<Pol, E>(v, prev != cur);
                windows<Q - 1, 0>(v, lds, u);
            }
            merges<Q + 1>(v, lds, u, cur);
        }
    }

    // v holds the blocked layout on entry and on exit (thread t holds
    // elements logical(t)*E ..); sorted ascending.
    __device__ static void sort(K (&v)[E], K* lds, uint32_t t) {
        const uint32_t u = logical(t);
        const bool f = (u & 1u) != 0;  // direction of merge size E
        reg_sort<Pol, E>(v, f);
        merges<LOGE + 1>(v, lds, u, f);
    }

    // striped global -> blocked registers, through LDS
    __device__ static __forceinline__ void from_striped(K (&v)[E], K* lds, uint32_t t) {
        const uint32_t u = logical(t);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = lds[phys(u * E + r)];
        __syncthreads();
    }
    __device__ static __forceinline__ void to_striped(const K (&v)[E], K* lds, uint32_t t) {
        const uint32_t u = logical(t);
#pragma unroll
        for (int r = 0; r < E; ++r) lds[phys(u * E + r)] = v[r];
        __syncthreads();
    }
};

// ---------------------------------------------------------------- wave tile
template <class Pol, int E>
struct WaveBitonic {
    using K = typename Pol::key_t;
    static constexpr int LOGE = ilog2(E);
    static constexpr int N = 64 * E;
    static constexpr int LOGN = LOGE + 6;

    template <int Q>
    __device__ static __forceinline__ void merges(K (&v)[E], uint32_t lane, bool prev) {
        if constexpr (Q <= LOGN) {
            const bool cur = (Q < LOGN) && (((lane >> (Q - LOGE)) & 1u) != 0);
            flip_all<Pol, E>(v, prev != cur);
            lane_stages_down<Pol, E, Q - 1 - LOGE>(v, lane);
            reg_low_up<Pol, E, LOGE - 1>(v);
            merges<Q + 1>(v, lane, cur);
        }
    }
    // blocked layout: lane l holds keys l*E .. l*E+E-1; sorted ascending.
    __device__ static void sort(K (&v)[E], uint32_t lane) {
        const bool f = (lane & 1u) != 0;
        reg_sort<Pol, E>(v, f);
        merges<LOGE + 1>(v, lane, f);
    }
};

}  // namespace s512
