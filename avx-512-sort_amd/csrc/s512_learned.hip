// s512_learned.hip -- GPU training and tabulation of the learned partitioner's
// CDF model (see s512_learned.hpp for the model).  No reference counterpart.
//
//   k_learn_sample   gather ~1% of the keys (hashed positions), encoded
//   (plain sort)     the sample is sorted by this library's own sort engine
//   k_learn_prep     features u and targets (rank+0.5)/M of M evenly spaced
//                    sample points; parameters initialised so that the net is
//                    the piecewise-linear interpolation through H+1 sample
//                    quantiles
//   k_mlp_step       one full-batch Adam step: per point forward + backward,
//                    per-workgroup partial gradients (fixed summation order),
//                    the last workgroup to finish sums them in block order and
//                    updates the parameters (deterministic, no spin-waits)
//   k_tabulate       model on LEARN_G+1 grid points, clamped to [0,1], made
//                    non-decreasing by a running maximum
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

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// grad_only: the last block writes the summed gradient to gout instead of
// updating the parameters.
__global__ __launch_bounds__(TRAIN_P) void k_mlp_step(TrainBufs tb, const float* u, const float* tg, uint32_t M,
                                                      float lr, int grad_only, float* gout) {
    constexpr int H = MLP_H;
    constexpr int NW = TRAIN_P / 64;
    constexpr int NSMALL = 4 * H + 1;  // p1, b1, b2, p3 gradients and b3
    __shared__ float p[MLP_NP];
    __shared__ float w2[H * H];
    __shared__ float sh1[H][TRAIN_P + 1];  // h1 per unit and point
    __shared__ float sd2[H][TRAIN_P + 1];  // layer-2 pre-activation, then its delta
    __shared__ float wred[NW][NSMALL + 1];
    __shared__ uint32_t last;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    for (int i = tid; i < MLP_NP; i += TRAIN_P) p[i] = tb.params[i];
    __syncthreads();
    for (int i = tid; i < H * H; i += TRAIN_P) w2[i] = p[MLP_OFF_P2 + i] * p[MLP_OFF_P2 + i];
    __syncthreads();
    const uint32_t pt = blockIdx.x * TRAIN_P + tid;
    const bool valid = pt < M;
    const float x = valid ? u[pt] : 0.0f, target = valid ? tg[pt] : 0.0f;
    // forward, layer 1 (ramp-derivative masks kept as bits)
    float h1[H];
    uint32_t m1 = 0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const float z = fmaf(p[j] * p[j], x, p[MLP_OFF_B1 + j]);
        h1[j] = ramp01(z);
        m1 |= (z > 0.0f && z < 1.0f) ? (1u << j) : 0u;
        sh1[j][tid] = h1[j];
    }
    // forward, layer 2 and output
    float y = p[MLP_OFF_B3];
    for (int k = 0; k < H; ++k) {
        float z = p[MLP_OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < H; ++j) z = fmaf(w2[k * H + j], h1[j], z);
        sd2[k][tid] = z;
        y = fmaf(p[MLP_OFF_P3 + k] * p[MLP_OFF_P3 + k], ramp01(z), y);
    }
    const float e = valid ? y - target : 0.0f;
    const float dy = 2.0f * e / (float)M;
    // backward, layer 2: delta per unit; reduce the b2 and p3 contributions
    float d1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) d1[j] = 0.0f;
    for (int k = 0; k < H; ++k) {
        const float z = sd2[k][tid];
        const float p3 = p[MLP_OFF_P3 + k];
        const float d2 = (z > 0.0f && z < 1.0f) ? dy * p3 * p3 : 0.0f;
        sd2[k][tid] = d2;
#pragma unroll
        for (int j = 0; j < H; ++j) d1[j] = fmaf(d2, w2[k * H + j], d1[j]);
        float gv = wave_sum(d2);
        if (lane == 0) wred[wid][2 * H + k] = gv;                       // b2
        gv = wave_sum(dy * 2.0f * p3 * ramp01(z));
        if (lane == 0) wred[wid][3 * H + k] = gv;                       // p3
    }
    // backward, layer 1
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const float dd = ((m1 >> j) & 1u) ? d1[j] : 0.0f;
        float gv = wave_sum(dd * 2.0f * p[j] * x);
        if (lane == 0) wred[wid][j] = gv;                               // p1
        gv = wave_sum(dd);
        if (lane == 0) wred[wid][H + j] = gv;                           // b1
    }
    float gv = wave_sum(dy);
    if (lane == 0) wred[wid][4 * H] = gv;                               // b3
    gv = wave_sum(e * e);
    if (lane == 0) wred[wid][NSMALL] = gv;                              // loss
    __syncthreads();
    float* part = tb.part + (size_t)blockIdx.x * (MLP_NP + 1);
    for (int i = tid; i <= NSMALL; i += TRAIN_P) {
        float s = 0.0f;
        for (int w = 0; w < NW; ++w) s += wred[w][i];
        int dst;
        if (i < H) dst = i;                                 // p1
        else if (i < 2 * H) dst = MLP_OFF_B1 + (i - H);     // b1
        else if (i < 3 * H) dst = MLP_OFF_B2 + (i - 2 * H); // b2
        else if (i < 4 * H) dst = MLP_OFF_P3 + (i - 3 * H); // p3
        else if (i == 4 * H) dst = MLP_OFF_B3;              // b3
        else dst = MLP_NP;                                  // loss
        part[dst] = s;
    }
    // layer-2 weights: G[k][j] = sum over points of d2_k * h1_j
    for (int q = tid; q < H * H; q += TRAIN_P) {
        const int k = q / H, j = q % H;
        float s = 0.0f;
        for (int i = 0; i < TRAIN_P; ++i) s = fmaf(sd2[k][i], sh1[j][i], s);
        part[MLP_OFF_P2 + q] = 2.0f * p[MLP_OFF_P2 + q] * s;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) last = (atomicAdd(&tb.ctr[0], 1u) == gridDim.x - 1) ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // last block: sum the partial gradients in block order
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

// Number of sample keys whose feature is < x (sorted encoded sample).
template <class K>
__device__ __forceinline__ uint32_t rank_below(const K* smp, uint32_t ns, uint64_t kmin, float inv, float x) {
    uint32_t lo = 0, hi = ns;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (learned_feature(smp[mid], kmin, inv) < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Empirical CDF of the sample at every grid point: ecdf[g] = rank / ns, and
// the band half-width wgt[g] from the sample keys in the adjacent cells.
template <class K>
__global__ __launch_bounds__(256) void k_calibrate(const K* smp, uint32_t ns, const uint64_t* geo, float* ecdf,
                                                   float* wgt) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > LEARN_G) return;
    const uint64_t kmin = geo[0], kmax = geo[1];
    const float inv = kmax > kmin ? 1.0f / (float)(kmax - kmin) : 0.0f;
    const float x = (float)g / (float)LEARN_G;
    const uint32_t r = g == 0 ? 0u : (g == LEARN_G ? ns : rank_below(smp, ns, kmin, inv, x));
    const float xl = (float)(g > 0 ? g - 1 : 0) / (float)LEARN_G, xh = (float)(g < LEARN_G ? g + 1 : LEARN_G) / (float)LEARN_G;
    const uint32_t rl = rank_below(smp, ns, kmin, inv, xl);
    const uint32_t rh = g + 1 >= LEARN_G ? ns : rank_below(smp, ns, kmin, inv, xh);
    ecdf[g] = (float)r / (float)ns;
    // half-width of the band the network may use around the empirical CDF:
    // half the sample mass of the two adjacent cells
    wgt[g] = 0.5f * (float)(rh - rl) / (float)ns;
}

// One block of 1024 threads: tab[g] = model(g / G) kept inside the sample's
// band around its empirical CDF, then a running maximum so the table is
// non-decreasing.
__global__ __launch_bounds__(1024) void k_tabulate(const float* params, const float* ecdf, const float* wgt,
                                                   float* tab) {
    constexpr int H = MLP_H;
    constexpr int PER = (LEARN_G + 1 + 1023) / 1024;
    __shared__ float p[MLP_NP];
    __shared__ float wmax[16];
    const uint32_t tid = threadIdx.x;
    for (int i = tid; i < MLP_NP; i += 1024) p[i] = params[i];
    __syncthreads();
    float v[PER];
    float run = 0.0f;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t g = tid * PER + q;
        float c = 0.0f;
        if (g <= LEARN_G) {
            const float x = (float)g / (float)LEARN_G;
            float h1[H];
#pragma unroll
            for (int j = 0; j < H; ++j) h1[j] = ramp01(fmaf(p[j] * p[j], x, p[MLP_OFF_B1 + j]));
            float y = p[MLP_OFF_B3];
            for (int k = 0; k < H; ++k) {
                float z = p[MLP_OFF_B2 + k];
#pragma unroll
                for (int j = 0; j < H; ++j) z = fmaf(p[MLP_OFF_P2 + k * H + j] * p[MLP_OFF_P2 + k * H + j], h1[j], z);
                y = fmaf(p[MLP_OFF_P3 + k] * p[MLP_OFF_P3 + k], ramp01(z), y);
            }
            y = fminf(fmaxf(y, 0.0f), 1.0f);
            // the network may move a grid value only within the sample's
            // resolution there (wgt = its half-width); where the sample has no
            // key nearby (wgt = 0) the empirical value stands
            c = fminf(fmaxf(y, ecdf[g] - wgt[g]), ecdf[g] + wgt[g]);
            c = fminf(fmaxf(c, 0.0f), 1.0f);
        }
        run = fmaxf(run, c);
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
        const uint32_t g = tid * PER + q;
        if (g <= LEARN_G) tab[g] = fmaxf(excl, v[q]);
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
    const uint32_t M = 1024, nblk = M / TRAIN_P;
    SizeCounter sc;
    sc.take<K>(ns);
    sc.take<float>(M);
    sc.take<float>(M);
    sc.take<float>(MLP_NP);
    sc.take<float>(MLP_NP);
    sc.take<float>(MLP_NP);
    sc.take<float>((size_t)nblk * (MLP_NP + 1));
    sc.take<uint32_t>(2);
    sc.take<float>(1);
    sc.take<uint64_t>(2);
    sc.take<float>(LEARN_G + 1);
    sc.take<float>(LEARN_G + 1);
    sc.take<float>(LEARN_G + 1);
    s512_status rc;
    if ((rc = ds.learn.reserve(sc.off, st))) return rc;
    Arena ar{(char*)ds.learn.ptr, 0};
    K* smp = ar.take<K>(ns);
    float* u = ar.take<float>(M);
    float* t = ar.take<float>(M);
    TrainBufs tb;
    tb.params = ar.take<float>(MLP_NP);
    tb.am = ar.take<float>(MLP_NP);
    tb.av = ar.take<float>(MLP_NP);
    tb.part = ar.take<float>((size_t)nblk * (MLP_NP + 1));
    tb.ctr = ar.take<uint32_t>(2);
    tb.loss = ar.take<float>(1);
    uint64_t* geo = ar.take<uint64_t>(2);
    float* tab = ar.take<float>(LEARN_G + 1);
    float* ecdf = ar.take<float>(LEARN_G + 1);
    float* wgt = ar.take<float>(LEARN_G + 1);
    S512_HIP(hipMemsetAsync(tb.am, 0, MLP_NP * sizeof(float), st));
    S512_HIP(hipMemsetAsync(tb.av, 0, MLP_NP * sizeof(float), st));
    S512_HIP(hipMemsetAsync(tb.ctr, 0, 2 * sizeof(uint32_t), st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (stats) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, st);
    }
    k_learn_sample<T><<<(ns + 255) / 256, 256, 0, st>>>(a, n, ns, prm.seed ^ 0x1EA2ull, smp);
    S512_CHECK_LAUNCH();
    // the sample is sorted by this library's own engine (lock already held)
    EngineParams sp = default_params(sizeof(K) == 8);
    if ((rc = sort_engine_locked<K, false>(smp, nullptr, ns, st, sp, nullptr, false, ds))) return rc;
    k_learn_prep<K><<<1, 256, 0, st>>>(smp, ns, M, u, t, tb.params, geo);
    S512_CHECK_LAUNCH();
    for (uint32_t s = 0; s < prm.train_steps; ++s) {
        k_mlp_step<<<nblk, TRAIN_P, 0, st>>>(tb, u, t, M, prm.learn_rate, 0, nullptr);
        S512_CHECK_LAUNCH();
    }
    k_calibrate<K><<<(LEARN_G + 1 + 255) / 256, 256, 0, st>>>(smp, ns, geo, ecdf, wgt);
    S512_CHECK_LAUNCH();
    k_tabulate<<<1, 1024, 0, st>>>(tb.params, ecdf, wgt, tab);
    S512_CHECK_LAUNCH();
    if ((rc = ds.pinned2.reserve(4096, st))) return rc;
    char* h = (char*)ds.pinned2.ptr;
    S512_HIP(hipMemcpyAsync(h, geo, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    S512_HIP(hipMemcpyAsync(h + 64, tb.loss, sizeof(float), hipMemcpyDeviceToHost, st));
    if (stats) (void)hipEventRecord(e1, st);
    S512_HIP(hipStreamSynchronize(st));
    uint64_t g2[2];
    memcpy(g2, h, sizeof(g2));
    float loss;
    memcpy(&loss, h + 64, sizeof(float));
    out->table.tab = tab;
    out->table.kmin = g2[0];
    out->table.inv_range = g2[1] > g2[0] ? 1.0f / (float)(g2[1] - g2[0]) : 0.0f;
    out->degenerate = g2[1] <= g2[0];
    out->loss = loss;
    out->params = tb.params;
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        stats->kernel_ms[S512_PHASE_TRAIN] += ms;
        stats->launches[S512_PHASE_TRAIN] += prm.train_steps;
        stats->keys[S512_PHASE_TRAIN] += ns;
        stats->model_loss = loss;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
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
