// s512_learned.hip -- GPU training of the learned partitioner's CDF model and
// its knots (see s512_learned.hpp for the model).  No reference counterpart.
//
//   k_learn_sample   gather ~1% of the keys (hashed positions), encoded
//   (plain sort)     the sample is sorted by this library's own sort engine
//   k_learn_prep     features u and targets (rank+0.5)/ns of M evenly spaced
//                    sample points; parameters initialised so that the net is
//                    the piecewise-linear interpolation through H+1 sample
//                    quantiles
//   k_mlp_train      every Adam step in one workgroup (one point per thread),
//                    on a stream of its own beside the first partition level:
//                    forward + backward, the gradient summed in a fixed order
//                    (deterministic), the update; the final full-batch loss
//   k_knots          knot keys at equal sample mass, the network's CDF there
//                    kept within the sample's band
//   k_knot_scan      running maximum: the knot CDF values made non-decreasing
// k_mlp_step is the same gradient spread over many workgroups, for the
// gradient check of s512_mlp_loss_grad_f32 on more points than one workgroup.
#include "s512_learned.hpp"
#include "s512_kernels.hpp"
#include "s512_runtime.hpp"
#include "s512_engine.hpp"
#include "s512_learned_host.hpp"
#include <algorithm>
#include <cmath>

namespace s512 {

template <class T>
__global__ __launch_bounds__(256) void k_learn_sample(const T* a, uint64_t n, uint32_t ns, uint64_t seed,
                                                      typename KeyTraits<T>::K* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const uint64_t pos = __umul64hi(mix64(seed, i), n);
    out[i] = KeyTraits<T>::enc(a[pos]);
}

// One block.  s: sorted encoded sample of ns keys.  Writes u[M], t[M],
// params[MLP_NP] (initial), and geo = {kmin, kmax} of the sample.
template <class K>
__global__ __launch_bounds__(256) void k_learn_prep(const K* s, uint32_t ns, uint32_t M, float* u, float* t,
                                                    float* params, uint64_t* geo) {
    const uint64_t kmin = (uint64_t)s[0], kmax = (uint64_t)s[ns - 1];
    const float inv = kmax > kmin ? 1.0f / (float)(kmax - kmin) : 0.0f;
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        const uint32_t r = (uint32_t)((((uint64_t)2 * i + 1) * ns) / (2ull * M));
        u[i] = learned_feature(s[r], kmin, inv);
        t[i] = ((float)r + 0.5f) / (float)ns;
    }
    if (threadIdx.x < MLP_H) {
        const int j = threadIdx.x;
        const uint32_t r0 = (uint32_t)(((uint64_t)j * (ns - 1)) / MLP_H);
        const uint32_t r1 = (uint32_t)(((uint64_t)(j + 1) * (ns - 1)) / MLP_H);
        const float k0 = learned_feature(s[r0], kmin, inv), k1 = learned_feature(s[r1], kmin, inv);
        const float w = 1.0f / fmaxf(k1 - k0, 1e-6f);
        params[j] = sqrtf(w);                        // p1
        params[MLP_OFF_B1 + j] = -k0 * w;            // b1
        params[MLP_OFF_B2 + j] = 0.0f;               // b2
        params[MLP_OFF_P3 + j] = sqrtf(1.0f / MLP_H);
        for (int i = 0; i < MLP_H; ++i) params[MLP_OFF_P2 + j * MLP_H + i] = (i == j) ? 1.0f : 0.02f;
    }
    if (threadIdx.x == 0) {
        params[MLP_OFF_B3] = 0.0f;
        geo[0] = kmin;
        geo[1] = kmax;
    }
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

// Point-mass check on the sorted sample: flag = 1 when some fp32 feature value
// is shared by more than `run` consecutive sample keys.  The network sees a
// key only through its feature and its CDF is continuous in it, so it cannot
// place a bucket boundary inside such a value's keys.
template <class K>
__global__ __launch_bounds__(256) void k_point_mass(const K* smp, uint32_t ns, uint32_t run, const uint64_t* geo,
                                                    uint32_t* flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + run >= ns) return;
    const uint64_t kmin = geo[0], kmax = geo[1];
    const float inv = kmax > kmin ? 1.0f / (float)(kmax - kmin) : 0.0f;
    if (learned_feature(smp[i], kmin, inv) == learned_feature(smp[i + run], kmin, inv)) flag[0] = 1u;
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

// Knot k: the sample key of rank r_k = k (ns - 1) / LEARN_KNOTS and the
// network's CDF there, kept within an eighth of the sample mass of the two
// adjacent pieces (a quarter of a knot's mass) around the empirical value
// (r_k + 0.5) / ns.  Neighbouring bands cannot overlap, so the knots stay
// strictly increasing: a flat stretch would place no splitter over a whole
// knot's keys.  params = nullptr: the empirical value itself.  key = nullptr:
// keys already written.
template <class K>
__global__ __launch_bounds__(256) void k_knots(const K* smp, uint32_t ns, const float* params, const uint64_t* geo,
                                               uint64_t* key, float* cdf_raw) {
    __shared__ float p[MLP_NP];
    if (params)
        for (int i = threadIdx.x; i < MLP_NP; i += blockDim.x) p[i] = params[i];
    __syncthreads();
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > LEARN_KNOTS) return;
    const uint64_t kmin = geo[0], kmax = geo[1];
    const float inv = kmax > kmin ? 1.0f / (float)(kmax - kmin) : 0.0f;
    auto rank = [&](uint32_t j) { return (uint32_t)(((uint64_t)j * (ns - 1)) / LEARN_KNOTS); };
    const uint32_t r = rank(k);
    const uint32_t rl = k > 0 ? rank(k - 1) : 0u, rh = k < LEARN_KNOTS ? rank(k + 1) : ns - 1;
    const float emp = ((float)r + 0.5f) / (float)ns;
    const float band = 0.125f * (float)(rh - rl) / (float)ns;
    const float y = params ? mlp_eval(p, learned_feature(smp[r], kmin, inv)) : emp;
    if (key) key[k] = (uint64_t)smp[r];
    cdf_raw[k] = fminf(fmaxf(fminf(fmaxf(y, emp - band), emp + band), 0.0f), 1.0f);
}

// One workgroup of 1024 threads: cdf = running maximum of cdf_raw.
__global__ __launch_bounds__(1024) void k_knot_scan(const float* cdf_raw, float* cdf) {
    constexpr int PER = (LEARN_KNOTS + 1 + 1023) / 1024;
    __shared__ float wmax[16];
    const uint32_t tid = threadIdx.x;
    float v[PER];
    float run = 0.0f;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t k = tid * PER + q;
        run = fmaxf(run, k <= LEARN_KNOTS ? cdf_raw[k] : 0.0f);
        v[q] = run;
    }
    // inclusive max-scan of the per-thread maxima
    float m = run;
    const uint32_t lane = tid & 63u, wid = tid >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float o = __shfl_up(m, d);
        if (lane >= (uint32_t)d) m = fmaxf(m, o);
    }
    if (lane == 63) wmax[wid] = m;
    __syncthreads();
    float before = 0.0f;
    for (uint32_t w = 0; w < wid; ++w) before = fmaxf(before, wmax[w]);
    const float excl = fmaxf(before, __shfl_up(m, 1) * (lane ? 1.0f : 0.0f));
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t k = tid * PER + q;
        if (k <= LEARN_KNOTS) cdf[k] = fmaxf(excl, v[q]);
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
    // a feature value on more than `run` sample keys holds more than
    // prm.imbalance times a first-level bucket's mean mass
    const uint32_t run = (uint32_t)std::max<double>(1.0, std::ceil(prm.imbalance * (double)ns / (double)(1u << lf1)));
    SizeCounter sc;
    sc.take<K>(ns);
    sc.take<float>(M);
    sc.take<float>(M);
    sc.take<float>(MLP_NP);
    sc.take<float>(1);
    sc.take<uint64_t>(2);
    sc.take<uint64_t>(LEARN_KNOTS + 1);
    for (int i = 0; i < 4; ++i) sc.take<float>(LEARN_KNOTS + 1);
    sc.take<uint32_t>(1);
    s512_status rc;
    if ((rc = ds.learn.reserve(sc.off, st))) return rc;
    Arena ar{(char*)ds.learn.ptr, 0};
    K* smp = ar.take<K>(ns);
    float* u = ar.take<float>(M);
    float* t = ar.take<float>(M);
    float* params = ar.take<float>(MLP_NP);
    float* loss = ar.take<float>(1);
    uint64_t* geo = ar.take<uint64_t>(2);
    uint64_t* key = ar.take<uint64_t>(LEARN_KNOTS + 1);
    float* ecdf_raw = ar.take<float>(LEARN_KNOTS + 1);
    float* ecdf = ar.take<float>(LEARN_KNOTS + 1);
    float* cdf_raw = ar.take<float>(LEARN_KNOTS + 1);
    float* cdf = ar.take<float>(LEARN_KNOTS + 1);
    uint32_t* flag = ar.take<uint32_t>(1);
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
    k_learn_sample<T><<<(ns + 255) / 256, 256, 0, st>>>(a, n, ns, prm.seed ^ 0x1EA2ull, smp);
    S512_CHECK_LAUNCH();
    // the sample is sorted by this library's own engine (lock already held),
    // with leaves of up to 8192 keys: for a few million keys one partition
    // level plus a short second one for the buckets above a tile
    EngineParams sp = default_params(sizeof(K) == 8);
    sp.target_leaf = 8192;
    if ((rc = sort_engine_locked<K, false>(smp, nullptr, ns, st, sp, nullptr, false, ds))) return rc;
    k_learn_prep<K><<<1, 256, 0, st>>>(smp, ns, M, u, t, params, geo);
    S512_CHECK_LAUNCH();
    S512_HIP(hipMemsetAsync(flag, 0, sizeof(uint32_t), st));
    k_point_mass<K><<<(ns + 255) / 256, 256, 0, st>>>(smp, ns, run, geo, flag);
    S512_CHECK_LAUNCH();
    // the first level's knots: the sample's own CDF
    k_knots<K><<<(LEARN_KNOTS + 1 + 255) / 256, 256, 0, st>>>(smp, ns, nullptr, geo, key, ecdf_raw);
    S512_CHECK_LAUNCH();
    k_knot_scan<<<1, 1024, 0, st>>>(ecdf_raw, ecdf);
    S512_CHECK_LAUNCH();
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    char* h = (char*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h, geo, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(h + 128, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (stats) (void)hipEventRecord(e1, st);
    S512_HIP(hipEventRecord(ds.lready, st));
    S512_HIP(hipStreamSynchronize(st));
    uint64_t g2[2];
    memcpy(g2, h, sizeof(g2));
    uint32_t pm;
    memcpy(&pm, h + 128, sizeof(pm));
    out->ecdf = LearnedModelDev{key, ecdf, smp, ns, (uint32_t)sizeof(K)};
    out->knots = LearnedModelDev{key, cdf, smp, ns, (uint32_t)sizeof(K)};
    out->params = params;
    out->degenerate = g2[1] <= g2[0];
    out->point_mass = pm != 0;
    out->lf1 = (uint32_t)lf1;
    out->training = false;
    out->trained = ds.ltrained;
    out->loss_host = (const float*)ds.lpinned.ptr;
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        stats->kernel_ms[S512_PHASE_TRAIN] += ms;
        stats->launches[S512_PHASE_TRAIN] += 1;
        stats->keys[S512_PHASE_TRAIN] += ns;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (out->degenerate || out->point_mass) return S512_OK;
    // the network: trained on the training stream while the caller runs the
    // first level; its knots (same keys, its own CDF) serve the second level
    hipStream_t ls = ds.lstream;
    S512_HIP(hipStreamWaitEvent(ls, ds.lready, 0));
    k_mlp_train<<<1, TRAIN_P, 0, ls>>>(params, u, t, M, prm.train_steps, prm.learn_rate, loss, nullptr);
    S512_CHECK_LAUNCH();
    k_knots<K><<<(LEARN_KNOTS + 1 + 255) / 256, 256, 0, ls>>>(smp, ns, params, geo, nullptr, cdf_raw);
    S512_CHECK_LAUNCH();
    k_knot_scan<<<1, 1024, 0, ls>>>(cdf_raw, cdf);
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
