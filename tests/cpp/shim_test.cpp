// shim_test.cpp -- the reference's own test pattern (sort512test.cpp: drand48
// inputs, sortedness, a multiset Checker, the kv value rule key*100+1) run
// through the drop-in headers include/sort512.hpp and include/sort512kv.hpp.
// Exit status 0 = pass.  Built by tests/cpp/Makefile, run by a GPU test.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sort512.hpp"
#include "sort512kv.hpp"

static int failures = 0;
#define EXPECT(c, ...)                                 \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
            ++failures;                                \
        }                                              \
    } while (0)

template <class T>
static void rand_vec(std::vector<T>& v) {
    for (auto& x : v) x = T(drand48() * double(v.size()));
}
template <class T>
static bool sorted(const std::vector<T>& v) {
    return std::is_sorted(v.begin(), v.end());
}
template <class T>
static bool same_multiset(std::vector<T> a, std::vector<T> b) {
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    return a == b;
}

template <class T>
static void test_sorts() {
    srand48(0);
    for (size_t n = 1; n <= (1 << 20); n *= 2) {
        std::vector<T> a(n);
        rand_vec(a);
        auto b = a, c = a, d = a;
        Sort512::Sort<T, size_t>(b.data(), n);
        EXPECT(sorted(b) && same_multiset(a, b), "Sort n=%zu", n);
        Sort512::SortOmpPartition<T, size_t>(c.data(), n);
        EXPECT(c == b, "SortOmpPartition n=%zu", n);
        Sort512::SortOmpParMerge<T, size_t>(d.data(), n);
        EXPECT(d == b, "SortOmpParMerge n=%zu", n);
        std::vector<T> e = a;
        Sort512::SortOmpMerge<T, int>(e.data(), int(n));
        EXPECT(e == b, "SortOmpMerge n=%zu", n);
        e = a;
        Sort512::SortOmpMergeDeps<T, long>(e.data(), long(n));
        EXPECT(e == b, "SortOmpMergeDeps n=%zu", n);
    }
}

template <class T>
static void test_partition() {
    srand48(0);
    for (size_t n = 1; n <= 1000; ++n) {
        std::vector<T> a(n);
        rand_vec(a);
        auto b = a;
        const T pivot = T(n / 2);
        size_t lim = Sort512::Partition512<size_t>(b.data(), 0, n - 1, pivot);
        size_t want = 0;
        for (auto x : a) want += x <= pivot;
        EXPECT(lim == want, "Partition512 n=%zu", n);
        for (size_t i = 0; i < n; ++i) EXPECT((i < lim) == (b[i] <= pivot), "partitioned n=%zu i=%zu", n, i);
        EXPECT(same_multiset(a, b), "multiset n=%zu", n);
    }
}

template <class T>
static void test_small() {
    srand48(0);
    const size_t maxn = 16 * 64 / sizeof(T);
    for (size_t n = 1; n <= maxn; ++n) {
        std::vector<T> a(n);
        rand_vec(a);
        auto b = a;
        Sort512::SmallSort16V(b.data(), n);
        EXPECT(sorted(b) && same_multiset(a, b), "SmallSort16V n=%zu", n);
    }
}

static void test_kv() {
    srand48(0);
    for (size_t n = 1; n <= (1 << 18); n *= 2) {
        std::vector<int> k(n), v(n);
        rand_vec(k);
        for (size_t i = 0; i < n; ++i) v[i] = k[i] * 100 + 1;
        auto k0 = k;
        Sort512kv::Sort<int, size_t>(k.data(), v.data(), n);
        EXPECT(sorted(k) && same_multiset(k0, k), "kv Sort n=%zu", n);
        for (size_t i = 0; i < n; ++i) EXPECT(v[i] == k[i] * 100 + 1, "kv pair n=%zu i=%zu", n, i);
        size_t lim = Sort512kv::Partition512<size_t>(k0.data(), v.data(), 0, n - 1, int(n / 2));
        (void)lim;
    }
    for (size_t n = 1; n <= 256; ++n) {
        std::vector<int> k(n), v(n);
        rand_vec(k);
        for (size_t i = 0; i < n; ++i) v[i] = k[i] * 100 + 1;
        Sort512kv::SmallSort16V(k.data(), v.data(), n);
        EXPECT(sorted(k), "kv SmallSort16V n=%zu", n);
        for (size_t i = 0; i < n; ++i) EXPECT(v[i] == k[i] * 100 + 1, "kv small pair n=%zu", n);
    }
}

static void test_merge() {
    for (int na : {1, 100, 5000, 70000})
        for (int nb : {1, 333, 90001}) {
            std::vector<double> a(na + nb);
            for (int i = 0; i < na; ++i) a[i] = double((i * 7) % 1000);
            for (int i = 0; i < nb; ++i) a[na + i] = double((i * 13) % 1000);
            std::sort(a.begin(), a.begin() + na);
            std::sort(a.begin() + na, a.end());
            auto want = a;
            std::inplace_merge(want.begin(), want.begin() + na, want.end());
            ParallelInplace::parallelMergeInPlace(a.data(), na + nb, na);
            EXPECT(a == want, "merge %d+%d", na, nb);
        }
}

// The reference's seven-argument form, called the way its SortOmpParMerge
// calls it (sort512.hpp:6104-6174): T threads sort T chunks, then at each
// level every thread of a group of 2^level calls the merge with the group's
// first thread, which then publishes that the group's run is done; a group's
// first thread first waits for the other half's flag.
static void test_merge_team() {
#ifdef _OPENMP
    const int T = 4;
    for (int n : {7, 1000, 100003, 2000000}) {
        std::vector<int> a(n);
        rand_vec(a);
        std::vector<int> want = a;
        std::sort(want.begin(), want.end());
        ParallelInplace::WorkingInterval<int> intervals[T] = {};
        int barrier[T] = {0};
        int done[3][T] = {{0}};
        const int chunk = (n + T - 1) / T;
#pragma omp parallel num_threads(T)
        {
            const int me = omp_get_thread_num();
            const int first = std::min(n, chunk * me), last = std::min(n, chunk * (me + 1));
            std::sort(a.begin() + first, a.begin() + last);
#pragma omp atomic write
            done[0][me] = 1;
            for (int level = 1; (1 << level) <= T; ++level) {
                const int lead = (me >> level) << level;
                const int half = (me >> (level - 1)) << (level - 1);
                const int other = (half >> (level - 1)) + (lead == half ? 1 : -1);
                for (;;) {
                    int d;
#pragma omp atomic read
                    d = done[level - 1][other];
                    if (d) break;
                }
                const int lo = std::min(n, lead * chunk);
                const int mid = std::min(n, lo + (1 << (level - 1)) * chunk);
                const int hi = std::min(n, lo + (1 << level) * chunk);
                ParallelInplace::parallelMergeInPlace(&a[lo], hi - lo, mid - lo, 1L << level, (long)lead, intervals,
                                                      barrier);
                if (me == lead) {
#pragma omp atomic write
                    done[level][me >> level] = 1;
                }
            }
        }
        EXPECT(a == want, "team merge n=%d", n);
    }
#endif
}

int main() {
    test_merge_team();
    test_sorts<int>();
    test_sorts<double>();
    test_partition<int>();
    test_partition<double>();
    test_small<int>();
    test_small<double>();
    test_kv();
    test_merge();
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
