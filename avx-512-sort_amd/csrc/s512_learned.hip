// s512_learned.hip -- the learned partitioner's CDF model: its table and the
// GPU training of its network (see s512_learned.hpp).  No reference
// counterpart.
//
//   k_learn_sample   gather ~1% of the keys (runs of 8 at hashed positions), encoded
//   k_tab_range      smallest and largest sampled key
//   k_tab_hist       sampled keys per top cell (LDS histograms), with each
//                    cell's smallest and largest key (repeated values)
//   k_tab_top        cumulative counts: the CDF at every top-cell boundary;
//                    the heaviest cells become nodes (split in CT_FAN)
//   k_tab_nodehist   sampled keys per sub-cell of the newest nodes    } once
//   k_tab_nodes      their CDF values; the heaviest sub-cells become   } per
//                    nodes of the next level                            } level
//   k_tab_train_points  the network's training points (features at 256
//                    evenly spaced CDF values of the table) and its initial
//                    parameters (the piecewise-linear interpolation through
//                    33 of them)
//   k_mlp_train      every Adam step in one workgroup (one point per thread),
//                    on a stream of its own beside the first partition level:
//                    forward + backward, the gradient summed in a fixed order
//                    (deterministic), the update; the final full-batch loss
//   k_tab_mlp        the second level's table: the network's CDF at the same
//                    boundaries, kept within the sample's band, non-decreasing
//   k_tab_local      each root bucket's local table (the network's CDF over the
//                    bucket's key range, renormalised) for the second level
// No step sorts the sample: the table is built from histograms.
// k_mlp_step is the same gradient spread over many workgroups, for the
// gradient check of s512_mlp_loss_grad_f32 on more points than one workgroup.
#include "s512_kernels.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"
#include "s512_learned_host.hpp"
#include <algorithm>
#include <cmath>

namespace s512 {

// Runs of `run` consecutive keys (up to one 64-byte line of 8-byte keys) at
// hashed positions: the gather reads whole lines instead of one key per line.
// The host keeps at least 512 runs per first-level bucket (consecutive keys of
// sorted input are one sample's worth of information, not `run`).
constexpr uint32_t LEARN_RUN_MAX = 8;
template <class T>
__global__ __launch_bounds__(256) void k_learn_sample(const T* a, uint64_t n, uint32_t ns, uint32_t run,
                                                      uint64_t seed, typename KeyTraits<T>::K* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const uint64_t starts = n > run ? n - run + 1 : 1;
    const uint64_t start = __umul64hi(mix64(seed, i / run), starts);
    const uint64_t pos = start + i % run;
    out[i] = KeyTraits<T>::enc(a[pos < n ? pos : n - 1]);
}

struct TrainBufs {
    float* params;  // [MLP_NP]
    float* am;      // Adam first moment
    float* av;      // Adam second moment
    float* part;    // [nblocks][MLP_NP + 1] partial gradients (+ partial loss)
    uint32_t* ctr;  // [2]: finished-block counter, step counter
    float* loss;    // last full-batch loss
};

// Scratch of one workgroup's gradient computation: per-point activations
// and deltas laid out [unit][point] (row stride TRAIN_P + 1: the row sums
// below read one column per lane, on distinct banks).
struct MlpLds {
    float sh1[MLP_H][TRAIN_P + 1];  // h1
    float sh2[MLP_H][TRAIN_P + 1];  // h2 (its ramp is active iff 0 < h2 < 1)
    float sdd[MLP_H][TRAIN_P + 1];  // layer-1 delta (ramp mask applied)
    float sdm[MLP_H][TRAIN_P + 1];  // dy where unit k's layer-2 ramp is active, else 0
    float sdy[TRAIN_P];             // dLoss/dy
    float sx[TRAIN_P];              // feature
    float se2[TRAIN_P];             // squared error
};

// Gradient of sum_i (y(u_i) - t_i)^2 / M over this workgroup's points (one per
// thread; `valid` false for a thread without one): out[i] for parameter i,
// out[MLP_NP] = the workgroup's sum of squared errors.  p: parameters, w2 =
// their squared layer-2 weights (both in LDS).  Every sum over points runs
// over the points in order, so the result does not depend on scheduling.
// Ends with a barrier.  Inlined into its kernels so that the LDS arrays are
// addressed as LDS (a call would pass them as generic pointers).
__device__ __forceinline__ void mlp_block_grad(const float* p, const float* w2, float x, float target, bool valid, uint32_t M,
                               MlpLds& L, float* out) {
    constexpr int H = MLP_H;
    const uint32_t tid = threadIdx.x;
    // forward
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        h1[j] = ramp01(fmaf(p[j] * p[j], x, p[MLP_OFF_B1 + j]));
        L.sh1[j][tid] = h1[j];
    }
    // (the unit loops stay rolled: unrolled, the compiler hoists all H*H
    // weight loads and runs out of registers)
    float y = p[MLP_OFF_B3];
#pragma unroll 2
    for (int k = 0; k < H; ++k) {
        float z = p[MLP_OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < H; ++j) z = fmaf(w2[k * H + j], h1[j], z);
        const float h2 = ramp01(z);
        L.sh2[k][tid] = h2;
        y = fmaf(p[MLP_OFF_P3 + k] * p[MLP_OFF_P3 + k], h2, y);
    }
    const float e = valid ? y - target : 0.0f;
    const float dy = 2.0f * e / (float)M;
    L.sdy[tid] = dy;
    L.sx[tid] = x;
    L.se2[tid] = e * e;
    // backward to layer 1 (per point)
    float d1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) d1[j] = 0.0f;
#pragma unroll 2
    for (int k = 0; k < H; ++k) {
        const float p3 = p[MLP_OFF_P3 + k];
        const float h2 = L.sh2[k][tid];
        const float dm = (h2 > 0.0f && h2 < 1.0f) ? dy : 0.0f;
        L.sdm[k][tid] = dm;
        const float d2 = dm * p3 * p3;
#pragma unroll
        for (int j = 0; j < H; ++j) d1[j] = fmaf(d2, w2[k * H + j], d1[j]);
    }
#pragma unroll
    for (int j = 0; j < H; ++j) L.sdd[j][tid] = (h1[j] > 0.0f && h1[j] < 1.0f) ? d1[j] : 0.0f;
    __syncthreads();
    // sums over points: thread r < 4H + 2 owns one bias / per-unit weight;
    // four accumulators over interleaved points, combined in a fixed order
    if (tid < 4 * H + 2) {
        const uint32_t r = tid;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        auto sum4 = [&](auto term) {
#pragma unroll 4
            for (int i = 0; i < TRAIN_P; i += 4) {
                a0 += term(i);
                a1 += term(i + 1);
                a2 += term(i + 2);
                a3 += term(i + 3);
            }
            return (a0 + a1) + (a2 + a3);
        };
        float sum;
        int dst;
        if (r < H) {                 // p1_j: 2 p1_j sum dd_j x
            sum = 2.0f * p[r] * sum4([&](int i) { return L.sdd[r][i] * L.sx[i]; });
            dst = r;
        } else if (r < 2 * H) {      // b1_j: sum dd_j
            sum = sum4([&](int i) { return L.sdd[r - H][i]; });
            dst = MLP_OFF_B1 + (r - H);
        } else if (r < 3 * H) {      // b2_k: p3_k^2 sum over active points of dy
            const int k = r - 2 * H;
            sum = sum4([&](int i) { return L.sdm[k][i]; });
            sum *= p[MLP_OFF_P3 + k] * p[MLP_OFF_P3 + k];
            dst = MLP_OFF_B2 + k;
        } else if (r < 4 * H) {      // p3_k: 2 p3_k sum dy h2_k
            const int k = r - 3 * H;
            sum = 2.0f * p[MLP_OFF_P3 + k] * sum4([&](int i) { return L.sdy[i] * L.sh2[k][i]; });
            dst = MLP_OFF_P3 + k;
        } else if (r == 4 * H) {     // b3: sum dy
            sum = sum4([&](int i) { return L.sdy[i]; });
            dst = MLP_OFF_B3;
        } else {                     // the workgroup's sum of squared errors
            sum = sum4([&](int i) { return L.se2[i]; });
            dst = MLP_NP;
        }
        out[dst] = sum;
    }
    // layer-2 weights: 2 p2_kj p3_k^2 sum over points active in unit k of dy h1_j,
    // a 2x2 (k, j) tile per thread
    {
        const int k0 = 2 * (tid / (H / 2)), j0 = 2 * (tid % (H / 2));
        float s00 = 0.0f, s01 = 0.0f, s10 = 0.0f, s11 = 0.0f;
#pragma unroll 8
        for (int i = 0; i < TRAIN_P; ++i) {
            const float da = L.sdm[k0][i], db = L.sdm[k0 + 1][i];
            const float xa = L.sh1[j0][i], xb = L.sh1[j0 + 1][i];
            s00 = fmaf(da, xa, s00);
            s01 = fmaf(da, xb, s01);
            s10 = fmaf(db, xa, s10);
            s11 = fmaf(db, xb, s11);
        }
        const float ca = 2.0f * p[MLP_OFF_P3 + k0] * p[MLP_OFF_P3 + k0];
        const float cb = 2.0f * p[MLP_OFF_P3 + k0 + 1] * p[MLP_OFF_P3 + k0 + 1];
        const int q = k0 * H + j0;
        out[MLP_OFF_P2 + q] = ca * p[MLP_OFF_P2 + q] * s00;
        out[MLP_OFF_P2 + q + 1] = ca * p[MLP_OFF_P2 + q + 1] * s01;
        out[MLP_OFF_P2 + q + H] = cb * p[MLP_OFF_P2 + q + H] * s10;
        out[MLP_OFF_P2 + q + H + 1] = cb * p[MLP_OFF_P2 + q + H + 1] * s11;
    }
    __syncthreads();
}

static_assert(TRAIN_P == (MLP_H / 2) * (MLP_H / 2), "one 2x2 layer-2 weight tile per thread");
static_assert(TRAIN_P >= 4 * MLP_H + 2, "one per-unit sum per thread");

// One full-batch step over many workgroups: partial gradients per workgroup,
// the last workgroup to finish sums them in workgroup order (deterministic,
// no spin-waits) and updates the parameters -- or, grad_only, writes the
// summed gradient to gout.
__global__ __launch_bounds__(TRAIN_P) void k_mlp_step(TrainBufs tb, const float* u, const float* tg, uint32_t M,
                                                      float lr, int grad_only, float* gout) {
    __shared__ float p[MLP_NP];
    __shared__ float w2[MLP_H * MLP_H];
    __shared__ MlpLds L;
    __shared__ uint32_t last;
    const uint32_t tid = threadIdx.x;
    for (int i = tid; i < MLP_NP; i += TRAIN_P) p[i] = tb.params[i];
    __syncthreads();
    for (int i = tid; i < MLP_H * MLP_H; i += TRAIN_P) w2[i] = p[MLP_OFF_P2 + i] * p[MLP_OFF_P2 + i];
    __syncthreads();
    const uint32_t pt = blockIdx.x * TRAIN_P + tid;
    const bool valid = pt < M;
    mlp_block_grad(p, w2, valid ? u[pt] : 0.0f, valid ? tg[pt] : 0.0f, valid, M, L,
                   tb.part + (size_t)blockIdx.x * (MLP_NP + 1));
    __threadfence();
    __syncthreads();
    if (tid == 0) last = (atomicAdd(&tb.ctr[0], 1u) == gridDim.x - 1) ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    __threadfence();
    const uint32_t step = tb.ctr[1] + 1;
    const float b1c = 1.0f / (1.0f - powf(0.9f, (float)step));
    const float b2c = 1.0f / (1.0f - powf(0.999f, (float)step));
    for (int i = tid; i <= MLP_NP; i += TRAIN_P) {
        float g = 0.0f;
        for (uint32_t b = 0; b < gridDim.x; ++b) g += tb.part[(size_t)b * (MLP_NP + 1) + i];
        if (i == MLP_NP) {
            *tb.loss = g / (float)M;
            continue;
        }
        if (grad_only) {
            gout[i] = g;
            continue;
        }
        const float m = 0.9f * tb.am[i] + 0.1f * g;
        const float v = 0.999f * tb.av[i] + 0.001f * g * g;
        tb.am[i] = m;
        tb.av[i] = v;
        tb.params[i] -= lr * (m * b1c) / (sqrtf(v * b2c) + 1e-8f);
    }
    if (tid == 0) {
        tb.ctr[0] = 0;
        if (!grad_only) tb.ctr[1] = step;
    }
}

// Every training step in one workgroup (M <= TRAIN_P points, one per thread;
// Adam with beta 0.9 / 0.999, bias-corrected, its moments in registers: thread
// t owns parameters t, t + TRAIN_P, ...).  Writes the trained parameters and
// the full-batch loss at them; with gout, also that loss's gradient (steps =
// 0: the gradient at the given parameters).
__global__ __launch_bounds__(TRAIN_P) void k_mlp_train(float* params, const float* u, const float* tg, uint32_t M,
                                                       uint32_t steps, float lr, float* loss_out, float* gout) {
    constexpr int OWN = (MLP_NP + TRAIN_P - 1) / TRAIN_P;
    __shared__ float p[MLP_NP], g[MLP_NP + 1];
    __shared__ float w2[MLP_H * MLP_H];
    __shared__ MlpLds L;
    const uint32_t tid = threadIdx.x;
    float am[OWN], av[OWN];
#pragma unroll
    for (int q = 0; q < OWN; ++q) {
        const uint32_t i = tid + q * TRAIN_P;
        if (i < MLP_NP) p[i] = params[i];
        am[q] = 0.0f;
        av[q] = 0.0f;
    }
    const bool valid = tid < M;
    const float x = valid ? u[tid] : 0.0f, target = valid ? tg[tid] : 0.0f;
    float b1t = 1.0f, b2t = 1.0f;  // beta1^step, beta2^step
    for (uint32_t step = 0;; ++step) {
        __syncthreads();
        for (int i = tid; i < MLP_H * MLP_H; i += TRAIN_P) w2[i] = p[MLP_OFF_P2 + i] * p[MLP_OFF_P2 + i];
        __syncthreads();
        mlp_block_grad(p, w2, x, target, valid, M, L, g);
        if (step == steps) break;  // the loss (and gradient) at the final parameters
        b1t *= 0.9f;
        b2t *= 0.999f;
        const float b1c = 1.0f / (1.0f - b1t), b2c = 1.0f / (1.0f - b2t);
#pragma unroll
        for (int q = 0; q < OWN; ++q) {
            const uint32_t i = tid + q * TRAIN_P;
            if (i >= MLP_NP) continue;
            const float gi = g[i];
            am[q] = 0.9f * am[q] + 0.1f * gi;
            av[q] = 0.999f * av[q] + 0.001f * gi * gi;
            p[i] -= lr * (am[q] * b1c) / (sqrtf(av[q] * b2c) + 1e-8f);
        }
    }
    for (int i = tid; i < MLP_NP; i += TRAIN_P) {
        params[i] = p[i];
        if (gout) gout[i] = g[i];
    }
    if (tid == 0) *loss_out = g[MLP_NP] / (float)M;
}

__device__ __forceinline__ float mlp_eval(const float* p, float x) {
    constexpr int H = MLP_H;
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) h1[j] = ramp01(fmaf(p[j] * p[j], x, p[MLP_OFF_B1 + j]));
    float y = p[MLP_OFF_B3];
#pragma unroll 1
    for (int k = 0; k < H; ++k) {
        float z = p[MLP_OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < H; ++j) z = fmaf(p[MLP_OFF_P2 + k * H + j] * p[MLP_OFF_P2 + k * H + j], h1[j], z);
        y = fmaf(p[MLP_OFF_P3 + k] * p[MLP_OFF_P3 + k], ramp01(z), y);
    }
    return y;
}

// ---------------------------------------------------------------- CDF table
struct TabScratch {
    uint64_t* geo;      // [2] smallest, largest sampled encoded key
    uint32_t* cnt;      // [CT_TOP] sampled keys per top cell
    uint64_t* cmin;     // [CT_TOP] smallest sampled key of the cell
    uint64_t* cmax;     // [CT_TOP] largest
    uint32_t* ncnt;     // [CT_NODES][CT_FAN] sampled keys per sub-cell of a node (row k - 1)
    uint64_t* nmin;     // [CT_NODES][CT_FAN]
    uint64_t* nmax;     // [CT_NODES][CT_FAN]
    uint32_t* nparent;  // [CT_NODES + 1] word of the entry node k splits
    uint64_t* noff;     // [CT_NODES + 1] offset of node k's range from the origin
    uint32_t* nw;       // [CT_NODES + 1] log2 of node k's width
    uint32_t* ctl;      // [4]: nodes so far, first node of the last round, sampled keys on repeated values
};

// Top cells of width 2^shift, from an origin at a multiple of 2^shift at or
// below kmin, cover [kmin, kmax]: (kmax - origin) >> shift < CT_TOP.  Cells on
// multiples of 2^shift never straddle a binade of fp64 keys while shift <= 52
// (a binade is an aligned range of 2^52 encoded keys), and nodes split them
// on multiples of smaller powers of two, so the density of keys inside a cell
// is that of at most one binade, where linear interpolation is exact for
// locally uniform keys.
struct TabGeo {
    uint64_t origin;
    uint32_t shift;
};
__host__ __device__ inline TabGeo tab_geometry(uint64_t kmin, uint64_t kmax) {
    const uint64_t span = kmax - kmin;
    uint32_t bits = 0;
    while (bits < 64 && (span >> bits) != 0) ++bits;
    uint32_t sh = bits > CT_TOP_LOG ? bits - CT_TOP_LOG : 0u;
    for (;;) {
        const uint64_t org = sh >= 64 ? 0ull : kmin & ~((1ull << sh) - 1ull);
        if (sh >= 63 || ((kmax - org) >> sh) < CT_TOP) return TabGeo{org, sh};
        ++sh;
    }
}

// Fixed-point CDF (CT_CDF_BITS bits) of v sampled keys out of ns.
__device__ __forceinline__ uint32_t cdf_fix(uint64_t v, uint64_t ns) {
    return v >= ns ? CT_CDF_MASK : (uint32_t)((v << CT_CDF_BITS) / ns);
}

__global__ __launch_bounds__(256) void k_tab_init(TabScratch ts) {
    for (uint32_t c = threadIdx.x; c < CT_TOP; c += blockDim.x) {
        ts.cnt[c] = 0;
        ts.cmin[c] = ~0ull;
        ts.cmax[c] = 0;
    }
    for (uint32_t j = threadIdx.x; j < CT_NODES * CT_FAN; j += blockDim.x) {
        ts.ncnt[j] = 0;
        ts.nmin[j] = ~0ull;
        ts.nmax[j] = 0;
    }
    if (threadIdx.x == 0) {
        ts.geo[0] = ~0ull;
        ts.geo[1] = 0;
        for (int i = 0; i < 4; ++i) ts.ctl[i] = 0;
    }
}

template <class K>
__global__ __launch_bounds__(256) void k_tab_range(const K* smp, uint32_t ns, uint64_t* geo) {
    uint64_t lo = ~0ull, hi = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        const uint64_t x = (uint64_t)smp[i];
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint64_t l2 = __shfl_xor(lo, d), h2 = __shfl_xor(hi, d);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    __shared__ uint64_t wl[4], wh[4];  // then one atomic per block: two hot words
    if ((threadIdx.x & 63u) == 0) {
        wl[threadIdx.x >> 6] = lo;
        wh[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = wl[w] < lo ? wl[w] : lo;
            hi = wh[w] > hi ? wh[w] : hi;
        }
        atomicMin((unsigned long long*)&geo[0], (unsigned long long)lo);
        atomicMax((unsigned long long*)&geo[1], (unsigned long long)hi);
    }
}

template <class K>
__global__ __launch_bounds__(256) void k_tab_hist(const K* smp, uint32_t ns, TabScratch ts) {
    __shared__ uint32_t lc[CT_TOP];
    __shared__ unsigned long long lmin[CT_TOP], lmax[CT_TOP];
    for (uint32_t c = threadIdx.x; c < CT_TOP; c += blockDim.x) {
        lc[c] = 0;
        lmin[c] = ~0ull;
        lmax[c] = 0;
    }
    __syncthreads();
    const TabGeo gm = tab_geometry(ts.geo[0], ts.geo[1]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        const uint64_t x = (uint64_t)smp[i];
        const uint64_t c64 = (x - gm.origin) >> gm.shift;
        const uint32_t c = c64 < CT_TOP ? (uint32_t)c64 : CT_TOP - 1;
        atomicAdd(&lc[c], 1u);
        atomicMin(&lmin[c], (unsigned long long)x);
        atomicMax(&lmax[c], (unsigned long long)x);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < CT_TOP; c += blockDim.x) {
        if (!lc[c]) continue;
        atomicAdd(&ts.cnt[c], lc[c]);
        atomicMin((unsigned long long*)&ts.cmin[c], lmin[c]);
        atomicMax((unsigned long long*)&ts.cmax[c], lmax[c]);
    }
}

// Exclusive prefix of one value per thread over a workgroup of 1024 threads
// (16 waves); *total = the sum.  Ends with a barrier.
__device__ __forceinline__ uint32_t block_excl_1024(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if (lane >= (uint32_t)d) incl += y;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < 16; ++w) {
        before += w < wid ? wsum[w] : 0u;
        all += wsum[w];
    }
    *total = all;
    __syncthreads();
    return before + incl - v;
}

// Heaviest-first choice among the candidate cells of a workgroup of 1024
// threads (N candidate counts per thread, 0 = not a candidate): the smallest
// threshold >= lo above which at most `budget` candidates lie.  Every thread
// returns it.
template <int N>
__device__ uint32_t heaviest_threshold(const uint32_t (&cand)[N], uint32_t lo, uint32_t budget, uint32_t* wsum) {
    uint32_t mx = 0;
#pragma unroll
    for (int q = 0; q < N; ++q) mx = cand[q] > mx ? cand[q] : mx;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(mx, d);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = mx;
    __syncthreads();
    uint32_t hi = 0;
    for (int w = 0; w < 16; ++w) hi = wsum[w] > hi ? wsum[w] : hi;
    __syncthreads();
    if (hi < lo) hi = lo;  // nothing above lo: lo itself
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        uint32_t mine = 0, above;
#pragma unroll
        for (int q = 0; q < N; ++q) mine += cand[q] > mid;
        block_excl_1024(mine, wsum, &above);
        if (above <= budget) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// One workgroup of CT_TOP threads (one top cell each): the CDF at the top
// boundaries, the header, and the first level of nodes: the heaviest cells
// over twice the mean mass, at most CT_NODES, whose sampled keys are not all
// one value.  Sampled keys of an unsplit cell whose keys are all one value
// count as keys on repeated values.
__global__ __launch_bounds__(CT_TOP) void k_tab_top(TabScratch ts, uint32_t ns, uint32_t* tab) {
    static_assert(CT_TOP == 1024, "one cell per thread of a 1024-thread workgroup");
    __shared__ uint32_t wsum[16];
    const uint32_t c = threadIdx.x;
    const uint32_t a0 = ts.cnt[c];
    const TabGeo gm = tab_geometry(ts.geo[0], ts.geo[1]);
    uint32_t tot;
    const uint32_t p = block_excl_1024(a0, wsum, &tot);
    const bool one_value = a0 >= 2 && ts.cmin[c] == ts.cmax[c];
    const uint32_t cand[1] = {(!one_value && gm.shift >= CT_FAN_LOG) ? a0 : 0u};
    const uint32_t thr = heaviest_threshold<1>(cand, max(2u * (ns / CT_TOP), 1u), CT_NODES, wsum);
    const uint32_t pick = cand[0] > thr;
    uint32_t npick;
    const uint32_t r = block_excl_1024(pick, wsum, &npick);
    const uint32_t id = pick ? r + 1 : 0u;
    tab[CT_OFF_TOP + c] = (id << CT_CDF_BITS) | cdf_fix(p, ns);
    if (c == CT_TOP - 1) tab[CT_OFF_TOP + CT_TOP] = CT_CDF_MASK;
    if (!pick && one_value) atomicAdd(&ts.ctl[2], a0);
    if (pick) {
        ts.nparent[id] = CT_OFF_TOP + c;
        ts.noff[id] = (uint64_t)c << gm.shift;
        ts.nw[id] = gm.shift;
    }
    // the new nodes' entries start empty (no sub-nodes) until their round
    for (uint32_t j = c; j < npick * (CT_FAN + 1); j += blockDim.x) tab[CT_OFF_NODE + j] = 0u;
    if (c == 0) {
        tab[0] = (uint32_t)gm.origin;
        tab[1] = (uint32_t)(gm.origin >> 32);
        tab[2] = gm.shift;
        tab[3] = npick;
        tab[4] = npick;  // last node of each level: words 4 .. 4 + CT_DEPTH - 1
        for (int w = 5; w < 8; ++w) tab[w] = npick;
        ts.ctl[0] = npick;
        ts.ctl[1] = 1;
    }
}

// Deepest table entry that holds offset d from the origin: its word index,
// the node it belongs to (0: the top), and the bits left below it.
struct TabWalk {
    uint32_t at, node, w;
    uint64_t rem;
};
__device__ __forceinline__ TabWalk tab_walk(const uint32_t* tab, uint64_t d, uint32_t shift) {
    const uint64_t c64 = d >> shift;
    const uint32_t c = c64 < CT_TOP ? (uint32_t)c64 : CT_TOP - 1;
    TabWalk r{CT_OFF_TOP + c, 0u, shift, c64 < CT_TOP ? d - (c64 << shift) : (1ull << shift) - 1ull};
    for (uint32_t lv = 0; lv < CT_DEPTH; ++lv) {
        const uint32_t child = tab[r.at] >> CT_CDF_BITS;
        if (!child) break;
        const uint32_t i = (uint32_t)(r.rem >> (r.w - CT_FAN_LOG));
        r.w -= CT_FAN_LOG;
        r.rem &= (1ull << r.w) - 1ull;
        r.node = child;
        r.at = ct_node(child) + i;
    }
    return r;
}

// Sampled keys per sub-cell of the nodes made in the last round (their
// entries are empty, so the walk ends in them).
template <class K>
__global__ __launch_bounds__(256) void k_tab_nodehist(const K* smp, uint32_t ns, TabScratch ts, const uint32_t* tab) {
    __shared__ uint32_t lc[CT_NODES * CT_FAN];
    __shared__ unsigned long long lmin[CT_NODES * CT_FAN], lmax[CT_NODES * CT_FAN];
    const uint32_t first = ts.ctl[1], last = ts.ctl[0];
    if (first > last) return;  // no new nodes (uniform: the whole grid returns)
    const uint32_t rows = (last - first + 1) * CT_FAN;
    for (uint32_t j = threadIdx.x; j < rows; j += blockDim.x) {
        lc[j] = 0;
        lmin[j] = ~0ull;
        lmax[j] = 0;
    }
    __syncthreads();
    const uint64_t origin = (uint64_t)tab[0] | ((uint64_t)tab[1] << 32);
    const uint32_t shift = tab[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        const uint64_t x = (uint64_t)smp[i];
        const TabWalk r = tab_walk(tab, x - origin, shift);
        if (r.node < first) continue;
        const uint32_t j = (r.node - first) * CT_FAN + (r.at - ct_node(r.node));
        atomicAdd(&lc[j], 1u);
        atomicMin(&lmin[j], (unsigned long long)x);
        atomicMax(&lmax[j], (unsigned long long)x);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < rows; j += blockDim.x) {
        if (!lc[j]) continue;
        const uint32_t g = (first - 1) * CT_FAN + j;
        atomicAdd(&ts.ncnt[g], lc[j]);
        atomicMin((unsigned long long*)&ts.nmin[g], lmin[j]);
        atomicMax((unsigned long long*)&ts.nmax[g], lmax[j]);
    }
}

// One workgroup of 1024 threads: the entries of the nodes made in the last
// round (one wave per node, four sub-cells per lane), then, unless `last`,
// the next level of nodes among their sub-cells (heaviest first, over twice
// the mean top-cell mass, within the node budget, splittable and not all one
// value).  Sampled keys of an unsplit sub-cell whose keys are all one value
// count as keys on repeated values.
__global__ __launch_bounds__(1024) void k_tab_nodes(TabScratch ts, uint32_t ns, uint32_t* tab, uint32_t lvl) {
    const bool last = lvl >= CT_DEPTH;  // the deepest level splits no further
    static_assert(CT_FAN == 4 * 64, "four sub-cells per lane");
    constexpr int PER = (CT_NODES * CT_FAN + 1023) / 1024;
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    const uint32_t first = ts.ctl[1], nlast = ts.ctl[0];
    if (first > nlast) return;
    for (uint32_t k = first + wid; k <= nlast; k += 16) {
        const uint32_t par = ts.nparent[k];
        const uint32_t lo = tab[par] & CT_CDF_MASK, hi = tab[par + 1] & CT_CDF_MASK;
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = ts.ncnt[(k - 1) * CT_FAN + 4 * lane + q];
            sum += v[q];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        uint64_t ex = incl - sum;
        uint32_t* e = tab + ct_node(k);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            e[4 * lane + q] = lo + (uint32_t)(((uint64_t)(hi - lo) * ex) / (total ? total : 1u));
            ex += v[q];
        }
        if (lane == 63) e[CT_FAN] = hi;
    }
    __syncthreads();
    // candidates: sub-cell j of the new nodes (row-major), one count per slot
    const uint32_t rows = (nlast - first + 1) * CT_FAN;
    uint32_t cand[PER];
    bool one[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t j = q * 1024 + t;
        cand[q] = 0;
        one[q] = false;
        if (j < rows) {
            const uint32_t g = (first - 1) * CT_FAN + j;
            const uint32_t k = first + j / CT_FAN;
            const uint32_t a0 = ts.ncnt[g];
            one[q] = a0 >= 2 && ts.nmin[g] == ts.nmax[g];
            if (!last && !one[q] && ts.nw[k] >= 2 * CT_FAN_LOG) cand[q] = a0;
        }
    }
    const uint32_t budget = CT_NODES - nlast;
    const uint32_t thr = heaviest_threshold<PER>(cand, max(2u * (ns / CT_TOP), 1u), budget, wsum);
    uint32_t npick = 0, mine = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) mine += cand[q] > thr;
    const uint32_t r0 = block_excl_1024(mine, wsum, &npick);
    uint32_t r = r0, dup = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t j = q * 1024 + t;
        if (j >= rows) continue;
        const uint32_t k = first + j / CT_FAN, i = j % CT_FAN;
        if (cand[q] > thr) {
            const uint32_t id = nlast + 1 + r++;
            const uint32_t at = ct_node(k) + i;
            tab[at] |= id << CT_CDF_BITS;
            ts.nparent[id] = at;
            ts.nw[id] = ts.nw[k] - CT_FAN_LOG;
            ts.noff[id] = ts.noff[k] + ((uint64_t)i << (ts.nw[k] - CT_FAN_LOG));
        } else if (one[q]) {
            dup += ts.ncnt[(first - 1) * CT_FAN + j];
        }
    }
    if (dup) atomicAdd(&ts.ctl[2], dup);
    for (uint32_t j = t; j < npick * (CT_FAN + 1); j += blockDim.x) tab[ct_node(nlast + 1) + j] = 0u;
    __syncthreads();
    if (t == 0) {
        tab[3] = nlast + npick;
        for (uint32_t w = 4 + lvl; w < 8; ++w) tab[w] = nlast + npick;
        ts.ctl[1] = nlast + 1;
        ts.ctl[0] = nlast + npick;
    }
}

// Word index and fraction where the table's CDF (CT_CDF_BITS) reaches y,
// and the offset of that position from the origin: a binary search in each
// node on the way down (building time only).
struct TabPos {
    uint32_t at;
    double f, off;
};
__device__ TabPos tab_inverse(const uint32_t* tab, uint32_t y) {
    const uint32_t shift = tab[2];
    uint32_t base = CT_OFF_TOP, n = CT_TOP, w = shift;
    double off = 0.0;
    uint32_t at = CT_OFF_TOP;
    for (uint32_t lv = 0; lv <= CT_DEPTH; ++lv) {
        uint32_t lo = 0, hi = n - 1;  // last entry with CDF <= y
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if ((tab[base + mid] & CT_CDF_MASK) <= y) lo = mid;
            else hi = mid - 1;
        }
        at = base + lo;
        const uint32_t cw = lv == 0 ? shift : w - CT_FAN_LOG;  // log2 of this entry's cell width
        off += ldexp((double)lo, (int)cw);
        w = cw;
        const uint32_t child = tab[at] >> CT_CDF_BITS;
        if (!child || lv == CT_DEPTH) break;
        base = ct_node(child);
        n = CT_FAN;
    }
    const uint32_t b0 = tab[at] & CT_CDF_MASK, b1 = tab[at + 1] & CT_CDF_MASK;
    const double f = b1 > b0 ? fmin((double)(y - b0) / (double)(b1 - b0), 1.0) : 0.0;
    return TabPos{at, f, off + f * ldexp(1.0, (int)w)};
}

// One workgroup: the network's M training points (features at evenly spaced
// CDF values of the sample's table) and its initial parameters (the
// piecewise-linear interpolation through 33 of them).
__global__ __launch_bounds__(256) void k_tab_train_points(const TabScratch ts, const uint32_t* tab, uint32_t M, float* u,
                                                          float* tg, float* params) {
    const uint64_t span = ts.geo[1] - ts.geo[0];
    const double inv = span ? 1.0 / (double)span : 0.0;
    const double lead = (double)(ts.geo[0] - ((uint64_t)tab[0] | ((uint64_t)tab[1] << 32)));  // kmin - origin
    auto feature_at = [&](double q) {  // feature of the key where the table's CDF is q in [0, 1]
        const double yq = q * (double)CT_CDF_MASK;
        const TabPos ps = tab_inverse(tab, (uint32_t)yq);
        const double f = (ps.off - lead) * inv;
        return (float)(f < 0.0 ? 0.0 : (f < 1.0 ? f : 1.0));
    };
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        const double q = ((double)i + 0.5) / (double)M;
        u[i] = feature_at(q);
        tg[i] = (float)q;
    }
    __shared__ float kq[MLP_H + 1];
    if (threadIdx.x <= MLP_H) kq[threadIdx.x] = feature_at((double)threadIdx.x / MLP_H);
    __syncthreads();
    if (threadIdx.x < MLP_H) {
        const int j = threadIdx.x;
        const float k0 = kq[j], k1 = kq[j + 1];
        const float w = 1.0f / fmaxf(k1 - k0, 1e-6f);
        params[j] = sqrtf(w);                        // p1
        params[MLP_OFF_B1 + j] = -k0 * w;            // b1
        params[MLP_OFF_B2 + j] = 0.0f;               // b2
        params[MLP_OFF_P3 + j] = sqrtf(1.0f / MLP_H);
        for (int i = 0; i < MLP_H; ++i) params[MLP_OFF_P2 + j * MLP_H + i] = (i == j) ? 1.0f : 0.02f;
    }
    if (threadIdx.x == 0) params[MLP_OFF_B3] = 0.0f;
}

// The second level's table: the first one's cells and nodes, the network's
// CDF at every boundary, each value kept within a quarter of the smaller
// adjacent cell's sampled mass, and within a quarter of a leaf's mass, around
// the sample table's value (neighbouring bands cannot overlap), then made
// non-decreasing inside each node and within the node's bounds.  A wider band
// lets the network's error (its fit is coarser than the sample's resolution)
// bend the CDF inside a root bucket, and the second level's leaves then lose
// their balance: measured with a band of a quarter of the cell mass, 2^28
// keys needed a third level.  One workgroup of CT_TOP threads; nodes level by level
// (a node's id is larger than its parent's).
__device__ __forceinline__ uint32_t wave_max_scan(uint32_t v) {  // inclusive, over the wave's lanes
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d);
        if (lane >= (uint32_t)d) v = o > v ? o : v;
    }
    return v;
}

__global__ __launch_bounds__(CT_TOP) void k_tab_mlp(const uint32_t* tab1, uint32_t* tab2, const float* params,
                                                    const TabScratch ts, uint32_t lf1) {
    __shared__ float p[MLP_NP];
    __shared__ uint32_t wmax[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, wid = t >> 6;
    for (uint32_t i = t; i < MLP_NP; i += blockDim.x) p[i] = params[i];
    for (uint32_t i = t; i < CT_OFF_TOP; i += blockDim.x) tab2[i] = tab1[i];
    __syncthreads();
    const uint64_t span = ts.geo[1] - ts.geo[0];
    const double inv = span ? 1.0 / (double)span : 0.0;
    const double lead = (double)(ts.geo[0] - ((uint64_t)tab1[0] | ((uint64_t)tab1[1] << 32)));  // kmin - origin
    // a quarter of a leaf's mass: the second level splits a root bucket into
    // at most FMAX leaves, so a boundary moved by less than that cannot make
    // any leaf more than half a leaf larger
    const double leaf4 = 0.25 * ldexp((double)CT_CDF_MASK, -(int)(lf1 + FMAX_LOG));
    // network value at offset `off` from the origin, kept within the band
    // around emp (m0, m1: the sampled masses of the cells on either side)
    auto smooth = [&](double off, uint32_t emp, uint32_t m0, uint32_t m1) -> uint32_t {
        const double f = (off - lead) * inv;
        const float y = mlp_eval(p, (float)(f < 0.0 ? 0.0 : (f < 1.0 ? f : 1.0)));
        const double band = fmin(0.25 * (double)(m0 < m1 ? m0 : m1), leaf4);
        double v = (double)y * (double)CT_CDF_MASK;
        v = fmin(fmax(v, (double)emp - band), (double)emp + band);
        v = fmin(fmax(v, 0.0), (double)CT_CDF_MASK);
        return (uint32_t)v;
    };
    // top boundaries, one per thread; entry CT_TOP is fixed
    {
        const uint32_t c = t, shift = tab1[2];
        const uint32_t w0 = tab1[CT_OFF_TOP + c], e = w0 & CT_CDF_MASK;
        const uint32_t m0 = c ? e - (tab1[CT_OFF_TOP + c - 1] & CT_CDF_MASK) : 0u;
        const uint32_t m1 = (tab1[CT_OFF_TOP + c + 1] & CT_CDF_MASK) - e;
        const uint32_t v = c ? smooth(ldexp((double)c, (int)shift), e, m0, m1) : 0u;
        const uint32_t m = wave_max_scan(v);
        if (lane == 63) wmax[wid] = m;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t q = 0; q < wid; ++q) before = wmax[q] > before ? wmax[q] : before;
        tab2[CT_OFF_TOP + c] = (w0 & ~CT_CDF_MASK) | (m > before ? m : before);
        if (c == 0) tab2[CT_OFF_TOP + CT_TOP] = CT_CDF_MASK;
    }
    __syncthreads();
    // nodes, one wave each, four entries per lane, level by level (header
    // word 4 + l - 1: the last node of level l), so a node's parent is final
    // before it
    for (uint32_t lv = 0; lv < CT_DEPTH; ++lv) {
    const uint32_t k_first = lv ? tab1[4 + lv - 1] + 1 : 1u, nn = tab1[4 + lv];
    for (uint32_t k0 = k_first; k0 <= nn; k0 += 16) {
        const uint32_t k = k0 + wid;
        if (k <= nn) {
            const uint32_t* e1 = tab1 + ct_node(k);
            uint32_t* e2 = tab2 + ct_node(k);
            const uint32_t par = ts.nparent[k];
            const uint32_t lo = tab2[par] & CT_CDF_MASK, hi = tab2[par + 1] & CT_CDF_MASK;
            const double o0 = (double)ts.noff[k];
            const uint32_t sw = ts.nw[k] - CT_FAN_LOG;
            uint32_t v[4], run = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = 4 * lane + q;
                const uint32_t e = e1[i] & CT_CDF_MASK;
                const uint32_t m0 = i ? e - (e1[i - 1] & CT_CDF_MASK) : 0u, m1 = (e1[i + 1] & CT_CDF_MASK) - e;
                uint32_t x = i ? smooth(o0 + ldexp((double)i, (int)sw), e, m0, m1) : lo;
                x = x < lo ? lo : (x > hi ? hi : x);
                run = x > run ? x : run;
                v[q] = run;
            }
            const uint32_t m = wave_max_scan(run);
            const uint32_t up = __shfl_up(m, 1);
            const uint32_t before = lane ? up : 0u;
#pragma unroll
            for (int q = 0; q < 4; ++q) e2[4 * lane + q] = (e1[4 * lane + q] & ~CT_CDF_MASK) | (v[q] > before ? v[q] : before);
            if (lane == 0) e2[CT_FAN] = hi;
        }
        __syncthreads();
    }
    }
}

// The second level's local tables, one workgroup per root bucket b: the
// bucket's key range [X_b, X_b+1) runs from where the sample's table crosses
// b / F1 to where it crosses (b + 1) / F1; LT_CELLS cells of width 2^sh from
// a multiple of 2^sh cover it; each boundary holds the network table's CDF
// there, renormalised so that X_b maps to 0 and X_b+1 to 1 (clamped).
__global__ __launch_bounds__(256) void k_tab_local(const uint32_t* tab1, const uint32_t* ntab, uint32_t lf1,
                                                   const uint64_t* geo, uint32_t* loc) {
    __shared__ uint64_t lo_off, hi_off;
    const uint32_t b = blockIdx.x, F1 = 1u << lf1;
    const uint64_t org = (uint64_t)tab1[0] | ((uint64_t)tab1[1] << 32);
    if (threadIdx.x == 0) {
        // the outer buckets end at the sampled extremes (the table's cells
        // reach beyond them; keys past them go to the first or last cell)
        const uint64_t y0 = ((uint64_t)b << CT_CDF_BITS) >> lf1, y1 = ((uint64_t)(b + 1) << CT_CDF_BITS) >> lf1;
        lo_off = b ? (uint64_t)tab_inverse(tab1, (uint32_t)y0).off : geo[0] - org;
        hi_off = b + 1 < F1 ? (uint64_t)tab_inverse(tab1, (uint32_t)y1).off : geo[1] - org;
        if (hi_off < lo_off) hi_off = lo_off;
    }
    __syncthreads();
    const uint64_t span = hi_off - lo_off;
    uint32_t sh = 0;
    while (sh < 56 && ((span >> sh) >= LT_CELLS - 1)) ++sh;  // one cell spare for the aligned origin
    const uint64_t o = lo_off & ~((1ull << sh) - 1ull);
    const uint64_t lim = ~0ull - org;  // offsets saturate at the largest encoded key
    auto at = [&](uint64_t off) { return (uint64_t)table_cdf(ntab, org + (off > lim ? lim : off)); };
    const uint64_t ya = at(lo_off), ye = at(hi_off);
    uint32_t* lt = loc + (size_t)b * LT_WORDS;
    if (threadIdx.x == 0) {
        lt[0] = (uint32_t)(org + o);
        lt[1] = (uint32_t)((org + o) >> 32);
        lt[2] = sh;
        lt[3] = 0;
    }
    for (uint32_t k = threadIdx.x; k <= LT_CELLS; k += blockDim.x) {
        const uint64_t off = o + ((uint64_t)k << sh);
        const uint64_t y = at(off < o ? lim : off);
        uint32_t v;
        if (y <= ya || ye <= ya) v = 0u;
        else if (y >= ye) v = 0xFFFFFFFFu;
        else v = (uint32_t)min<uint64_t>(((y - ya) << 32) / (ye - ya), 0xFFFFFFFFull);
        lt[4 + k] = v;
    }
}

// ---------------------------------------------------------------- host
template <class T>
s512_status build_learned_model(const T* a, size_t n, hipStream_t st, DeviceState& ds, const EngineParams& prm,
                                LearnedModel* out, s512_stats* stats) {
    using K = typename KeyTraits<T>::K;
    // sample_permille of the keys, at least 64 per first-level bucket, at most n
    const uint64_t leaves = (n + prm.target_leaf - 1) / prm.target_leaf;
    uint64_t f1 = 1;
    while (f1 < leaves && f1 < (1ull << prm.max_logf)) f1 <<= 1;
    const uint32_t ns = (uint32_t)std::min<uint64_t>(
        n, std::min<uint64_t>(std::max<uint64_t>((n * prm.sample_permille) / 1000, 64 * f1), 1u << 26));
    const uint32_t M = TRAIN_P;
    const int lf1 = std::min(std::max(1, leaves <= 1 ? 0 : 64 - __builtin_clzll(leaves - 1)), prm.max_logf);
    SizeCounter sc;
    sc.take<K>(ns);
    sc.take<uint64_t>(2);
    sc.take<uint32_t>(CT_TOP);
    sc.take<uint64_t>(CT_TOP);
    sc.take<uint64_t>(CT_TOP);
    sc.take<uint32_t>(CT_NODES * CT_FAN);
    sc.take<uint64_t>(CT_NODES * CT_FAN);
    sc.take<uint64_t>(CT_NODES * CT_FAN);
    sc.take<uint32_t>(CT_NODES + 1);
    sc.take<uint64_t>(CT_NODES + 1);
    sc.take<uint32_t>(CT_NODES + 1);
    sc.take<uint32_t>(4);
    sc.take<uint32_t>(CT_WORDS);
    sc.take<uint32_t>(CT_WORDS);
    sc.take<uint32_t>((size_t)FMAX * LT_WORDS);
    sc.take<float>(M);
    sc.take<float>(M);
    sc.take<float>(MLP_NP);
    sc.take<float>(1);
    s512_status rc;
    if ((rc = ds.learn.reserve(sc.off, st))) return rc;
    Arena ar{(char*)ds.learn.ptr, 0};
    K* smp = ar.take<K>(ns);
    TabScratch ts;
    ts.geo = ar.take<uint64_t>(2);
    ts.cnt = ar.take<uint32_t>(CT_TOP);
    ts.cmin = ar.take<uint64_t>(CT_TOP);
    ts.cmax = ar.take<uint64_t>(CT_TOP);
    ts.ncnt = ar.take<uint32_t>(CT_NODES * CT_FAN);
    ts.nmin = ar.take<uint64_t>(CT_NODES * CT_FAN);
    ts.nmax = ar.take<uint64_t>(CT_NODES * CT_FAN);
    ts.nparent = ar.take<uint32_t>(CT_NODES + 1);
    ts.noff = ar.take<uint64_t>(CT_NODES + 1);
    ts.nw = ar.take<uint32_t>(CT_NODES + 1);
    ts.ctl = ar.take<uint32_t>(4);
    uint32_t* tab1 = ar.take<uint32_t>(CT_WORDS);
    uint32_t* ntab = ar.take<uint32_t>(CT_WORDS);                 // the network's table
    uint32_t* loc = ar.take<uint32_t>((size_t)FMAX * LT_WORDS);   // the root buckets' local tables
    float* u = ar.take<float>(M);
    float* t = ar.take<float>(M);
    float* params = ar.take<float>(MLP_NP);
    float* loss = ar.take<float>(1);
    if (!ds.lstream) {
        // default priority; does not synchronise with the null stream
        S512_HIP(hipStreamCreateWithFlags(&ds.lstream, hipStreamNonBlocking));
        S512_HIP(hipEventCreateWithFlags(&ds.lready, hipEventDisableTiming));
        S512_HIP(hipEventCreateWithFlags(&ds.ltrained, hipEventDisableTiming));
    }
    if ((rc = ds.lpinned.reserve(256, st))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (stats) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st);
    }
    uint64_t* flags = (uint64_t*)((char*)ds.lpinned.ptr + 64);  // kmin, kmax, repeated-value keys
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((ns + 16383) / 16384, 256));
    uint32_t run = LEARN_RUN_MAX;
    while (run > 1 && ns / run < 512ull << lf1) run >>= 1;
    k_learn_sample<T><<<(ns + 255) / 256, 256, 0, st>>>(a, n, ns, run, prm.seed ^ 0x1EA2ull, smp);
    S512_CHECK_LAUNCH();
    k_tab_init<<<1, 256, 0, st>>>(ts);
    S512_CHECK_LAUNCH();
    k_tab_range<K><<<g, 256, 0, st>>>(smp, ns, ts.geo);
    S512_CHECK_LAUNCH();
    k_tab_hist<K><<<g, 256, 0, st>>>(smp, ns, ts);
    S512_CHECK_LAUNCH();
    k_tab_top<<<1, CT_TOP, 0, st>>>(ts, ns, tab1);
    S512_CHECK_LAUNCH();
    for (uint32_t lvl = 1; lvl <= CT_DEPTH; ++lvl) {
        k_tab_nodehist<K><<<g, 256, 0, st>>>(smp, ns, ts, tab1);
        S512_CHECK_LAUNCH();
        k_tab_nodes<<<1, 1024, 0, st>>>(ts, ns, tab1, lvl);
        S512_CHECK_LAUNCH();
    }
    k_tab_train_points<<<1, 256, 0, st>>>(ts, tab1, M, u, t, params);
    S512_CHECK_LAUNCH();
    S512_HIP(hipMemcpyAsync(flags, ts.geo, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(flags + 2, ts.ctl + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (stats) (void)hipEventRecord(e1, st);
    S512_HIP(hipEventRecord(ds.lready, st));
    out->tabs = LearnedModelDev{tab1, loc};
    out->params = params;
    out->flags_host = flags;
    out->ns = ns;
    out->lf1 = (uint32_t)lf1;
    out->training = false;
    out->trained = ds.ltrained;
    out->loss_host = (const float*)ds.lpinned.ptr;
    out->ev0 = e0;
    out->ev1 = e1;
    // the network: trained on the training stream while the caller runs the
    // first level on the sample's table; its table serves the second level
    hipStream_t ls = ds.lstream;
    S512_HIP(hipStreamWaitEvent(ls, ds.lready, 0));
    k_mlp_train<<<1, TRAIN_P, 0, ls>>>(params, u, t, M, prm.train_steps, prm.learn_rate, loss, nullptr);
    S512_CHECK_LAUNCH();
    k_tab_mlp<<<1, CT_TOP, 0, ls>>>(tab1, ntab, params, ts, (uint32_t)lf1);
    S512_CHECK_LAUNCH();
    k_tab_local<<<1u << lf1, 256, 0, ls>>>(tab1, ntab, (uint32_t)lf1, ts.geo, loc);
    S512_CHECK_LAUNCH();
    S512_HIP(hipMemcpyAsync(ds.lpinned.ptr, loss, sizeof(float), hipMemcpyDeviceToHost, ls));
    S512_HIP(hipEventRecord(ds.ltrained, ls));
    out->training = true;
    return S512_OK;
}

template s512_status build_learned_model<double>(const double*, size_t, hipStream_t, DeviceState&,
                                                 const EngineParams&, LearnedModel*, s512_stats*);
template s512_status build_learned_model<int32_t>(const int32_t*, size_t, hipStream_t, DeviceState&,
                                                  const EngineParams&, LearnedModel*, s512_stats*);

}  // namespace s512

using namespace s512;

// Full-batch loss and gradient of the model on the GPU (fp32), for checking
// the training kernel against the oracle's fp64 restatement.
extern "C" s512_status s512_mlp_loss_grad_f32(const float* d_params, const float* d_u, const float* d_t, uint32_t m,
                                              float* d_grad, float* d_loss, void* stream) {
    if (!d_params || !d_u || !d_t || !d_grad || !d_loss) { set_error("null pointer"); return S512_ERR_INVALID; }
    if (m == 0 || m > (1u << 20)) { set_error("m = %u out of range (1..2^20)", m); return S512_ERR_INVALID; }
    hipStream_t st = (hipStream_t)stream;
    int dev;
    s512_status rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mu);
    const uint32_t nblk = (m + TRAIN_P - 1) / TRAIN_P;
    SizeCounter sc;
    sc.take<float>(MLP_NP);
    sc.take<float>((size_t)nblk * (MLP_NP + 1));
    sc.take<uint32_t>(2);
    if ((rc = ds.arena.reserve(sc.off, st))) return rc;
    Arena ar{(char*)ds.arena.ptr, 0};
    if (m <= TRAIN_P) {
        // one workgroup: the training kernel itself, with zero steps
        float* prm = ar.take<float>(MLP_NP);
        S512_HIP(hipMemcpyAsync(prm, d_params, MLP_NP * sizeof(float), hipMemcpyDeviceToDevice, st));
        k_mlp_train<<<1, TRAIN_P, 0, st>>>(prm, d_u, d_t, m, 0, 0.0f, d_loss, d_grad);
        S512_CHECK_LAUNCH();
        S512_HIP(hipStreamSynchronize(st));
        return S512_OK;
    }
    TrainBufs tb;
    tb.params = ar.take<float>(MLP_NP);
    tb.part = ar.take<float>((size_t)nblk * (MLP_NP + 1));
    tb.ctr = ar.take<uint32_t>(2);
    tb.am = tb.av = nullptr;
    tb.loss = d_loss;
    S512_HIP(hipMemcpyAsync(tb.params, d_params, MLP_NP * sizeof(float), hipMemcpyDeviceToDevice, st));
    S512_HIP(hipMemsetAsync(tb.ctr, 0, 2 * sizeof(uint32_t), st));
    k_mlp_step<<<nblk, TRAIN_P, 0, st>>>(tb, d_u, d_t, m, 0.0f, 1, d_grad);
    S512_CHECK_LAUNCH();
    S512_HIP(hipStreamSynchronize(st));
    return S512_OK;
}

extern "C" uint32_t s512_mlp_nparams(void) { return MLP_NP; }
