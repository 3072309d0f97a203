/*
 * s512_oracle.c -- CPU restatement of the reference sort path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, calls or
 * loads this file; it is the checker that tests/, the smoke entry and the
 * CPU-baseline leg of the benchmark compare the HIP path against.
 *
 * What it restates (reference = berenger-eu/avx-512-sort, header-only C++):
 *   - Sort512::Sort / CoreSort recursion, median-of-3 pivot, pivot swap
 *     (sort512.hpp:5916-5963), with an explicit stack instead of recursion so
 *     that the quadratic inputs cannot overflow the C stack.  Visiting order
 *     is the reference's: right part first, then left part (5955-5956).
 *   - Partition512 contract (sort512.hpp:5695-5910): inclusive range
 *     [left,right], keys `<= pivot` first, returns left + #{a[i] <= pivot}.
 *     The reference's scalar path CoreScalarPartition (5695-5711) is restated
 *     exactly; the vector path produces a different arrangement inside each
 *     side, which the reference does not specify (only the return value and
 *     the partition property are tested, sort512test.cpp:41-60, 1451-1483).
 *   - SmallSort16V (sort512.hpp:2255-2633, 5312-5688): pad the array to
 *     ceil(len/lanes) vectors with DBL_MAX / INT_MAX, run a bitonic network,
 *     store the first len results.  Restated as a plain bitonic sorting
 *     network over the padded power-of-two length.  This reproduces the
 *     reference's padding side effects (an +inf key comes back as DBL_MAX
 *     when len % 8 != 0).
 *   - Sort512kv (sort512kv.hpp:5235-6135): same recursion carrying a value
 *     array; in the networks a pair moves iff its key changes, i.e. a
 *     compare-exchange swaps the pair iff key_i > key_j (sort512kv.hpp:55-56);
 *     pads are (INT_MAX, INT_MAX) (sort512kv.hpp:5244-5247).
 *   - ParallelInplace::FindMedian co-rank split (parallelInplace.hpp:77-134)
 *     and the merge it feeds (std::inplace_merge, parallelInplace.hpp:190).
 *   - Synthetic input generators (SplitMix64 counter PRNG, Box-Muller,
 *     POSIX drand48 as used by the reference tests, sort512test.cpp:77-82).
 *   - fp64 forward/backward of the 1-32-32-1 monotone MLP CDF model used by
 *     the learned partitioner (no reference counterpart; clamped-ramp
 *     activation, squared weights).
 *
 * Build: oracle/Makefile -> oracle/libs512oracle.so (gcc -O2, no AVX needed).
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <limits.h>

/* ------------------------------------------------------------------------ */
/* Scalar partition: sort512.hpp:5695-5711 (CoreScalarPartition).            */
/* ------------------------------------------------------------------------ */
#define DEFINE_SCALAR_PARTITION(NAME, T)                                      \
    static size_t NAME(T *a, size_t left, size_t right, T pivot)             \
    {                                                                         \
        /* left/right inclusive; right may be left-1 (empty range) */         \
        size_t l = left;                                                      \
        while (l <= right && a[l] <= pivot) ++l;                              \
        for (size_t i = l; i <= right; ++i) {                                 \
            if (a[i] <= pivot) { T t = a[i]; a[i] = a[l]; a[l] = t; ++l; }    \
        }                                                                     \
        return l;                                                             \
    }
DEFINE_SCALAR_PARTITION(part_scalar_i32, int32_t)
DEFINE_SCALAR_PARTITION(part_scalar_f64, double)

/* kv: sort512kv.hpp:5924-5941, values follow their keys */
static size_t part_scalar_kv(int32_t *k, int32_t *v, size_t left, size_t right, int32_t pivot)
{
    size_t l = left;
    while (l <= right && k[l] <= pivot) ++l;
    for (size_t i = l; i <= right; ++i) {
        if (k[i] <= pivot) {
            int32_t t = k[i]; k[i] = k[l]; k[l] = t;
            t = v[i]; v[i] = v[l]; v[l] = t;
            ++l;
        }
    }
    return l;
}

/* Public Partition512 contract (sort512.hpp:5729, 5821; sort512kv.hpp:5958).
 * right < left (empty) is accepted and returns left. */
size_t orc_partition_i32(int32_t *a, size_t left, size_t right, int32_t pivot)
{
    if (right + 1 <= left) return left;
    return part_scalar_i32(a, left, right, pivot);
}
size_t orc_partition_f64(double *a, size_t left, size_t right, double pivot)
{
    if (right + 1 <= left) return left;
    return part_scalar_f64(a, left, right, pivot);
}
size_t orc_partition_kv(int32_t *k, int32_t *v, size_t left, size_t right, int32_t pivot)
{
    if (right + 1 <= left) return left;
    return part_scalar_kv(k, v, left, right, pivot);
}

/* ------------------------------------------------------------------------ */
/* SmallSort16V restatement: pad to whole vectors with the type maximum,     */
/* bitonic network over the next power of two, keep the first len.          */
/* sort512.hpp:2255-2633 (f64, 8 lanes, <=128), 5312-5688 (i32, 16 lanes,    */
/* <=256); sort512kv.hpp:5235-5914 (kv, 16 lanes, <=256).                    */
/* ------------------------------------------------------------------------ */
static size_t next_pow2(size_t x) { size_t p = 1; while (p < x) p <<= 1; return p; }

#define DEFINE_BITONIC(NAME, T)                                               \
    static void NAME(T *x, size_t n) /* n power of two, ascending */          \
    {                                                                         \
        for (size_t k = 2; k <= n; k <<= 1)                                   \
            for (size_t j = k >> 1; j > 0; j >>= 1)                           \
                for (size_t i = 0; i < n; ++i) {                              \
                    size_t p = i ^ j;                                         \
                    if (p <= i) continue;                                     \
                    int up = (i & k) == 0;                                    \
                    T lo = x[i], hi = x[p];                                   \
                    if ((lo > hi) == up) { x[i] = hi; x[p] = lo; }            \
                }                                                             \
    }
DEFINE_BITONIC(bitonic_i32, int32_t)
DEFINE_BITONIC(bitonic_f64, double)

static void bitonic_kv(int32_t *k, int32_t *v, size_t n)
{
    /* a pair moves iff the key compare says so; equal keys never exchange
       (sort512kv.hpp:55-56: value taken from the partner iff key changed) */
    for (size_t kk = 2; kk <= n; kk <<= 1)
        for (size_t j = kk >> 1; j > 0; j >>= 1)
            for (size_t i = 0; i < n; ++i) {
                size_t p = i ^ j;
                if (p <= i) continue;
                int up = (i & kk) == 0;
                int swap = up ? (k[i] > k[p]) : (k[i] < k[p]);
                if (swap) {
                    int32_t t = k[i]; k[i] = k[p]; k[p] = t;
                    t = v[i]; v[i] = v[p]; v[p] = t;
                }
            }
}

int orc_smallsort_i32(int32_t *a, size_t len)
{
    if (len == 0 || len > 256) return -1;
    int32_t buf[256];
    size_t nv = (len + 15) / 16, padded = nv * 16, n = next_pow2(padded);
    for (size_t i = 0; i < n; ++i) buf[i] = i < len ? a[i] : INT_MAX;
    bitonic_i32(buf, n);
    memcpy(a, buf, len * sizeof(int32_t));
    return 0;
}
int orc_smallsort_f64(double *a, size_t len)
{
    if (len == 0 || len > 128) return -1;
    double buf[128];
    size_t nv = (len + 7) / 8, padded = nv * 8, n = next_pow2(padded);
    for (size_t i = 0; i < n; ++i) buf[i] = i < len ? a[i] : DBL_MAX;
    bitonic_f64(buf, n);
    memcpy(a, buf, len * sizeof(double));
    return 0;
}
int orc_smallsort_kv(int32_t *k, int32_t *v, size_t len)
{
    if (len == 0 || len > 256) return -1;
    int32_t kb[256], vb[256];
    size_t nv = (len + 15) / 16, padded = nv * 16, n = next_pow2(padded);
    for (size_t i = 0; i < n; ++i) {
        kb[i] = i < len ? k[i] : INT_MAX;
        vb[i] = i < len ? v[i] : INT_MAX;
    }
    bitonic_kv(kb, vb, n);
    memcpy(k, kb, len * sizeof(int32_t));
    memcpy(v, vb, len * sizeof(int32_t));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CoreSort: sort512.hpp:5916-5963 (and kv 6086-6135).                       */
/* ------------------------------------------------------------------------ */
typedef struct { size_t l, r; } range_t;

#define DEFINE_QSORT(NAME, T, LIMIT, SMALL, PART)                             \
    static size_t NAME##_pivot(const T *a, size_t l, size_t r)               \
    {                                                                         \
        size_t m = ((r - l) / 2) + l;                                         \
        if (a[l] <= a[m] && a[m] <= a[r]) return m;                           \
        else if (a[m] <= a[l] && a[l] <= a[r]) return l;                      \
        return r;                                                             \
    }                                                                         \
    int NAME(T *a, size_t n)                                                  \
    {                                                                         \
        if (n <= 1) return 0; /* the reference underflows size-1 at n==0 */   \
        size_t cap = 64, top = 0;                                             \
        range_t *st = (range_t *)malloc(cap * sizeof(range_t));               \
        if (!st) return -1;                                                   \
        st[top++] = (range_t){0, n - 1};                                      \
        while (top) {                                                         \
            range_t rg = st[--top];                                           \
            size_t l = rg.l, r = rg.r;                                        \
            if (r - l < (LIMIT)) { SMALL(a + l, r - l + 1); continue; }       \
            size_t part = l;                                                  \
            if (r - l > 1) {                                                  \
                size_t p = NAME##_pivot(a, l, r);                             \
                T t = a[p]; a[p] = a[r]; a[r] = t;                            \
                part = PART(a, l, r - 1, a[r]);                               \
                t = a[part]; a[part] = a[r]; a[r] = t;                        \
            }                                                                 \
            if (top + 2 > cap) {                                              \
                cap *= 2;                                                     \
                range_t *ns = (range_t *)realloc(st, cap * sizeof(range_t));  \
                if (!ns) { free(st); return -1; }                             \
                st = ns;                                                      \
            }                                                                 \
            /* reference order: right first (5955), then left (5956);       \
               a stack pops last-pushed first, so push left then right */    \
            if (part && l < part - 1) st[top++] = (range_t){l, part - 1};     \
            if (part + 1 < r) st[top++] = (range_t){part + 1, r};             \
        }                                                                     \
        free(st);                                                             \
        return 0;                                                             \
    }

static void small_i32(int32_t *a, size_t len) { orc_smallsort_i32(a, len); }
static void small_f64(double *a, size_t len) { orc_smallsort_f64(a, len); }
/* SortLimite = 16*64/sizeof(T): 256 for int, 128 for double (sort512.hpp:5949) */
DEFINE_QSORT(orc_sort_i32, int32_t, 256, small_i32, orc_partition_i32)
DEFINE_QSORT(orc_sort_f64, double, 128, small_f64, orc_partition_f64)

/* Segmented sort: the caller's loop of Sort512::Sort (sort512.hpp:5960) over
 * the segments a[off[i] .. off[i+1]); returns -1 on decreasing offsets. */
#define DEFINE_SEGSORT(NAME, T, SORT)                                         \
    int NAME(T *a, const uint64_t *off, size_t nseg)                          \
    {                                                                         \
        for (size_t i = 0; i < nseg; ++i)                                     \
            if (off[i + 1] < off[i]) return -1;                               \
        for (size_t i = 0; i < nseg; ++i)                                     \
            if (SORT(a + off[i], (size_t)(off[i + 1] - off[i]))) return -1;   \
        return 0;                                                             \
    }
DEFINE_SEGSORT(orc_segsort_i32, int32_t, orc_sort_i32)
DEFINE_SEGSORT(orc_segsort_f64, double, orc_sort_f64)

static size_t kv_pivot(const int32_t *a, size_t l, size_t r)
{
    size_t m = ((r - l) / 2) + l;
    if (a[l] <= a[m] && a[m] <= a[r]) return m;
    else if (a[m] <= a[l] && a[l] <= a[r]) return l;
    return r;
}
/* sort512kv.hpp:6099-6135: pivot swaps also move the values (6103-6107) */
int orc_sort_kv(int32_t *k, int32_t *v, size_t n)
{
    if (n <= 1) return 0;
    size_t cap = 64, top = 0;
    range_t *st = (range_t *)malloc(cap * sizeof(range_t));
    if (!st) return -1;
    st[top++] = (range_t){0, n - 1};
    while (top) {
        range_t rg = st[--top];
        size_t l = rg.l, r = rg.r;
        if (r - l < 256) { orc_smallsort_kv(k + l, v + l, r - l + 1); continue; }
        size_t part = l;
        if (r - l > 1) {
            size_t p = kv_pivot(k, l, r);
            int32_t t = k[p]; k[p] = k[r]; k[r] = t;
            t = v[p]; v[p] = v[r]; v[r] = t;
            part = orc_partition_kv(k, v, l, r - 1, k[r]);
            t = k[part]; k[part] = k[r]; k[r] = t;
            t = v[part]; v[part] = v[r]; v[r] = t;
        }
        if (top + 2 > cap) {
            cap *= 2;
            range_t *ns = (range_t *)realloc(st, cap * sizeof(range_t));
            if (!ns) { free(st); return -1; }
            st = ns;
        }
        if (part && l < part - 1) st[top++] = (range_t){l, part - 1};
        if (part + 1 < r) st[top++] = (range_t){part + 1, r};
    }
    free(st);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Closed-form checkers (ascending order is unique): O(n log n) sorts that   */
/* never go quadratic, for sizes where the restated quicksort would.         */
/* ------------------------------------------------------------------------ */
static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}
static int cmp_f64(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}
static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}
void orc_qsort_i32(int32_t *a, size_t n) { qsort(a, n, sizeof(int32_t), cmp_i32); }
void orc_qsort_f64(double *a, size_t n) { qsort(a, n, sizeof(double), cmp_f64); }
/* pairs sorted lexicographically by (key, value) as one u64 each */
void orc_pair_multiset_kv(const int32_t *k, const int32_t *v, size_t n, uint64_t *out)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = ((uint64_t)((uint32_t)k[i] ^ 0x80000000u) << 32) | (uint32_t)v[i];
    qsort(out, n, sizeof(uint64_t), cmp_u64);
}

/* ------------------------------------------------------------------------ */
/* Merge: ParallelInplace::FindMedian (parallelInplace.hpp:77-134) restated, */
/* and the merge of two sorted runs it splits.                               */
/* ------------------------------------------------------------------------ */
#define DEFINE_FINDMEDIAN(NAME, T)                                            \
    void NAME(const T *a, long center, long size, long *mA, long *mB)        \
    {                                                                         \
        if (center == 0 || center == size || a[center - 1] <= a[center]) {    \
            *mA = center; *mB = 0; return;                                    \
        }                                                                     \
        long ls = 0, ll = center, lp = (ll - ls) / 2 + ls;                    \
        long rs = center, rl = size, rp = (rl - rs) / 2 + rs;                 \
        while (ls < ll && rs < rl && !(a[lp] == a[rp])) {                     \
            long A0 = lp, A1 = center - lp, B0 = rp - center, B1 = size - rp; \
            if (a[lp] < a[rp]) {                                              \
                if (A0 + B0 < A1 + B1) { ls = lp + 1; lp = (ll - ls) / 2 + ls; } \
                else { rl = rp; rp = (rl - rs) / 2 + rs; }                    \
            } else {                                                          \
                if (A0 + B0 < A1 + B1) { rs = rp + 1; rp = (rl - rs) / 2 + rs; } \
                else { ll = lp; lp = (ll - ls) / 2 + ls; }                    \
            }                                                                 \
        }                                                                     \
        *mA = lp; *mB = rp - center;                                          \
    }
DEFINE_FINDMEDIAN(orc_findmedian_i32, int32_t)
DEFINE_FINDMEDIAN(orc_findmedian_f64, double)

/* Merge the sorted runs a[0,center) and a[center,size) in place (through a
 * temporary, as std::inplace_merge does when it can allocate). Stable. */
#define DEFINE_MERGE(NAME, T)                                                 \
    int NAME(T *a, size_t center, size_t size)                               \
    {                                                                         \
        if (center == 0 || center >= size) return 0;                          \
        T *tmp = (T *)malloc(size * sizeof(T));                               \
        if (!tmp) return -1;                                                  \
        size_t i = 0, j = center, o = 0;                                      \
        while (i < center && j < size) tmp[o++] = (a[j] < a[i]) ? a[j++] : a[i++]; \
        while (i < center) tmp[o++] = a[i++];                                 \
        while (j < size) tmp[o++] = a[j++];                                   \
        memcpy(a, tmp, size * sizeof(T));                                     \
        free(tmp);                                                            \
        return 0;                                                             \
    }
DEFINE_MERGE(orc_merge_i32, int32_t)
DEFINE_MERGE(orc_merge_f64, double)

/* Merge-path co-rank: number of elements taken from A among the first diag
 * outputs of the stable merge of A (len na) and B (len nb).  Brute-force
 * definition: the unique i with A[i-1] <= B[diag-i] and B[diag-i-1] < A[i]. */
#define DEFINE_CORANK(NAME, T)                                                \
    size_t NAME(const T *A, size_t na, const T *B, size_t nb, size_t diag)   \
    {                                                                         \
        size_t lo = diag > nb ? diag - nb : 0, hi = diag < na ? diag : na;    \
        while (lo < hi) {                                                     \
            size_t i = (lo + hi) / 2, j = diag - i - 1;                       \
            if (B[j] < A[i]) hi = i; else lo = i + 1;                         \
        }                                                                     \
        return lo;                                                            \
    }
DEFINE_CORANK(orc_corank_i32, int32_t)
DEFINE_CORANK(orc_corank_f64, double)

/* ------------------------------------------------------------------------ */
/* Generators.                                                               */
/* ------------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea, Flood 2014) as a counter-based hash of (seed,i). */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double u01(uint64_t seed, uint64_t i) { return (double)(orc_splitmix64(seed, i) >> 11) * 0x1.0p-53; }

void orc_gen_i32_full(int32_t *out, size_t n, uint64_t seed, uint64_t start)
{
    for (size_t i = 0; i < n; ++i) out[i] = (int32_t)(uint32_t)(orc_splitmix64(seed, start + i) >> 32);
}
void orc_gen_f64_uniform(double *out, size_t n, uint64_t seed, uint64_t start)
{
    for (size_t i = 0; i < n; ++i) out[i] = u01(seed, start + i);
}
/* Box-Muller, one normal per index from two uniforms (u1 guarded away from 0). */
void orc_gen_f64_normal(double *out, size_t n, uint64_t seed, uint64_t start)
{
    for (size_t i = 0; i < n; ++i) {
        uint64_t g = start + i;
        double u1 = u01(seed, 2 * g), u2 = u01(seed, 2 * g + 1);
        if (u1 < 0x1.0p-53) u1 = 0x1.0p-53;
        out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}
void orc_gen_f64_distinct(double *out, size_t n, uint64_t seed, uint64_t start, uint64_t ndistinct)
{
    for (size_t i = 0; i < n; ++i) out[i] = (double)(orc_splitmix64(seed, start + i) % ndistinct);
}

/* POSIX drand48: X_{k+1} = (a X_k + c) mod 2^48, a = 0x5DEECE66D, c = 0xB,
 * srand48(s) sets X = (s << 16) | 0x330E; result = X / 2^48. */
static uint64_t d48_state = 0x330Eull;
void orc_srand48(long s) { d48_state = (((uint64_t)(uint32_t)s) << 16) | 0x330Eull; }
double orc_drand48(void)
{
    d48_state = (0x5DEECE66Dull * d48_state + 0xBull) & ((1ull << 48) - 1);
    return (double)d48_state * 0x1.0p-48;
}
/* createRandVec (sort512test.cpp:77-82): NumType(drand48()*size) */
void orc_rand_vec_i32(int32_t *a, size_t n)
{
    for (size_t i = 0; i < n; ++i) a[i] = (int32_t)(orc_drand48() * (double)n);
}
void orc_rand_vec_f64(double *a, size_t n)
{
    for (size_t i = 0; i < n; ++i) a[i] = orc_drand48() * (double)n;
}

/* ------------------------------------------------------------------------ */
/* Learned partitioner model, fp64 restatement (no reference counterpart).   */
/* Monotone 1-H-H-1 MLP with the clamped ramp r(z) = min(max(z,0),1):        */
/*   h1 = r(w1*x + b1), h2 = r(W2 h1 + b2), y = w3 . h2 + b3                 */
/* with w1, W2, w3 >= 0 obtained as squares of the free parameters (w = p*p),*/
/* so y is non-decreasing in x; the bounded ramp (unlike a plain ReLU, which */
/* with non-negative weights only yields convex functions) lets the net fit  */
/* S-shaped CDFs.  r'(z) is taken as 1 on (0,1) and 0 elsewhere.            */
/* Parameter vector layout (H = hidden width):                               */
/*   p1[H], b1[H], p2[H*H] (row j = unit j of layer 2), b2[H], p3[H], b3[1]  */
/* Loss over a batch: mean over i of (y(x_i) - t_i)^2.                       */
/* ------------------------------------------------------------------------ */
static double ramp(double z) { return z <= 0 ? 0 : (z >= 1 ? 1 : z); }
static double dramp(double z) { return (z > 0 && z < 1) ? 1.0 : 0.0; }

size_t orc_mlp_nparams(int H) { return (size_t)H * 3 + (size_t)H * H + (size_t)H + 1; }

double orc_mlp_forward(const double *p, int H, double x)
{
    const double *p1 = p, *b1 = p + H, *p2 = p + 2 * H, *b2 = p2 + H * H, *p3 = b2 + H, *b3 = p3 + H;
    double h1[256], y = b3[0];
    for (int j = 0; j < H; ++j) h1[j] = ramp(p1[j] * p1[j] * x + b1[j]);
    for (int j = 0; j < H; ++j) {
        double z = b2[j];
        for (int i = 0; i < H; ++i) z += p2[j * H + i] * p2[j * H + i] * h1[i];
        y += p3[j] * p3[j] * ramp(z);
    }
    return y;
}

/* Returns the loss; writes d loss / d p into grad (same layout). */
double orc_mlp_loss_grad(const double *p, int H, const double *x, const double *t, size_t n, double *grad)
{
    const double *p1 = p, *b1 = p + H, *p2 = p + 2 * H, *b2 = p2 + H * H, *p3 = b2 + H, *b3 = p3 + H;
    size_t np = orc_mlp_nparams(H);
    memset(grad, 0, np * sizeof(double));
    double *g1 = grad, *gb1 = grad + H, *g2 = grad + 2 * H, *gb2 = g2 + H * H, *g3 = gb2 + H, *gb3 = g3 + H;
    double loss = 0;
    double z1[256], h1[256], z2[256], h2[256], d2[256], d1[256];
    for (size_t s = 0; s < n; ++s) {
        double xv = x[s];
        for (int j = 0; j < H; ++j) { z1[j] = p1[j] * p1[j] * xv + b1[j]; h1[j] = ramp(z1[j]); }
        double y = b3[0];
        for (int j = 0; j < H; ++j) {
            double z = b2[j];
            for (int i = 0; i < H; ++i) z += p2[j * H + i] * p2[j * H + i] * h1[i];
            z2[j] = z; h2[j] = ramp(z);
            y += p3[j] * p3[j] * h2[j];
        }
        double e = y - t[s];
        loss += e * e;
        double dy = 2.0 * e / (double)n;
        gb3[0] += dy;
        for (int j = 0; j < H; ++j) {
            g3[j] += dy * 2.0 * p3[j] * h2[j];
            d2[j] = dy * p3[j] * p3[j] * dramp(z2[j]);
            gb2[j] += d2[j];
        }
        for (int i = 0; i < H; ++i) d1[i] = 0;
        for (int j = 0; j < H; ++j) {
            if (d2[j] == 0.0) continue;
            for (int i = 0; i < H; ++i) {
                double w = p2[j * H + i];
                g2[j * H + i] += d2[j] * 2.0 * w * h1[i];
                d1[i] += d2[j] * w * w;
            }
        }
        for (int i = 0; i < H; ++i) {
            double d = d1[i] * dramp(z1[i]);
            gb1[i] += d;
            g1[i] += d * 2.0 * p1[i] * xv;
        }
    }
    return loss / (double)n;
}
