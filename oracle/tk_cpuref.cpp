/*
 * tk_cpuref.cpp -- multi-threaded CPU restatement of the normalization job's
 * table passes (count + per-read coverage statistics), the CPU baseline
 * bench.py times on the host cores.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and tests/ load it;
 * it is checked against tk_oracle.c (tests/test_oracle.py).  Same semantics
 * as the oracle's or_count + or_read_stats (jellyfish count -m k -C,
 * Trinity:2612-2619; fastaToKmerCoverageStats --DS with missing k-mers as 1,
 * util/insilico_read_normalization.pl:837-845), written the way a CPU tool
 * would: one OpenMP pass over the reads into a lock-free open-addressing
 * table (CAS on the key, atomic add on the count), then one pass of lookups
 * per read with an nth_element median.  N-containing windows count as
 * missing; the median of an even number of values is the upper one.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

static inline uint64_t cr_mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static inline int cr_code(unsigned char c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return -1;
    }
}

typedef struct {
    uint64_t* keys;   /* key + 1 (0 = empty) */
    uint32_t* counts;
    uint64_t mask;
} cr_table;

static void cr_add(cr_table* t, uint64_t key) {
    const uint64_t tagged = key + 1;
    uint64_t h = cr_mix64(key) & t->mask;
    for (;;) {
        uint64_t cur = __atomic_load_n(&t->keys[h], __ATOMIC_RELAXED);
        if (cur == 0) {
            uint64_t expect = 0;
            if (__atomic_compare_exchange_n(&t->keys[h], &expect, tagged, 0, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED))
                cur = tagged;
            else
                cur = expect;
        }
        if (cur == tagged) {
            __atomic_fetch_add(&t->counts[h], 1u, __ATOMIC_RELAXED);
            return;
        }
        h = (h + 1) & t->mask;
    }
}

static uint32_t cr_get(const cr_table* t, uint64_t key) {
    const uint64_t tagged = key + 1;
    uint64_t h = cr_mix64(key) & t->mask;
    for (;;) {
        const uint64_t cur = t->keys[h];
        if (cur == tagged) return t->counts[h];
        if (cur == 0) return 0;
        h = (h + 1) & t->mask;
    }
}

/* Calls f(ctx, valid, key) for every window of read [s, s+len). */
#define CR_WINDOWS(s, len, k, canonical, BODY)                                  \
    do {                                                                        \
        const uint64_t kmask_ = (k) == 32 ? ~0ull : ((1ull << (2 * (k))) - 1); \
        uint64_t f_ = 0, r_ = 0;                                                \
        int run_ = 0;                                                           \
        for (uint64_t i_ = 0; i_ < (len); i_++) {                               \
            const int c_ = cr_code((unsigned char)(s)[i_]);                     \
            if (c_ < 0) {                                                       \
                run_ = 0;                                                       \
            } else {                                                            \
                f_ = ((f_ << 2) | (uint64_t)c_) & kmask_;                       \
                r_ = (r_ >> 2) | ((uint64_t)(3 - c_) << (2 * (k) - 2));         \
                run_++;                                                         \
            }                                                                   \
            if (i_ + 1 >= (uint64_t)(k)) {                                      \
                const int valid = run_ >= (k);                                  \
                const uint64_t key = (canonical) && r_ < f_ ? r_ : f_;          \
                BODY;                                                           \
            }                                                                   \
        }                                                                       \
    } while (0)

/*
 * Count the k-mers of reads [0, n) (offsets off[n+1] into seq) with `threads`
 * OpenMP threads, then the per-read median/nwin/sum of window values
 * (count if >= min_count, else 1).  Writes wall-clock seconds of each pass
 * and the number of distinct keys; with `digest` set, also an
 * order-independent digest of the (key, count) table, the sum mod 2^64 of
 * mix64(key ^ count * 0x9E3779B97F4A7C15) over every key (a table check that
 * needs no sorted dump), and with `histo` set the counts' histogram in
 * jellyfish histo's default bins (10,001 bins: counts 1..10000, then > 10000).
 * Returns 0, or -1 when memory runs out.
 */
extern "C" int cr_count_stats(const char* seq, const uint64_t* off, uint64_t n, int k, int canonical,
                   uint32_t min_count, int threads, uint64_t expected_distinct, uint32_t* median,
                   uint32_t* nwin, uint64_t* sum, double* t_count, double* t_stats,
                   uint64_t* n_distinct, uint64_t* digest, uint64_t* histo) {
    uint64_t cap = 1024;
    while (cap < 2 * expected_distinct) cap <<= 1;
    cr_table t;
    t.keys = (uint64_t*)calloc(cap, sizeof(uint64_t));
    t.counts = (uint32_t*)calloc(cap, sizeof(uint32_t));
    t.mask = cap - 1;
    if (!t.keys || !t.counts) {
        free(t.keys);
        free(t.counts);
        return -1;
    }
    omp_set_num_threads(threads);
    double a = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 256)
    for (uint64_t r = 0; r < n; r++) {
        const char* s = seq + off[r];
        const uint64_t len = off[r + 1] - off[r];
        CR_WINDOWS(s, len, k, canonical, if (valid) cr_add(&t, key));
    }
    double b = omp_get_wtime();
    uint64_t nd = 0;
#pragma omp parallel for reduction(+ : nd)
    for (uint64_t i = 0; i < cap; i++) nd += t.keys[i] != 0;
    double c = omp_get_wtime();
    if (digest) {
        uint64_t dg = 0;
#pragma omp parallel for reduction(+ : dg)
        for (uint64_t i = 0; i < cap; i++)
            if (t.keys[i]) dg += cr_mix64((t.keys[i] - 1) ^ ((uint64_t)t.counts[i] * 0x9E3779B97F4A7C15ULL));
        *digest = dg;
    }
    if (histo) {
        memset(histo, 0, 10001 * sizeof(uint64_t));
#pragma omp parallel
        {
            uint64_t* h = (uint64_t*)calloc(10001, sizeof(uint64_t));
#pragma omp for
            for (uint64_t i = 0; i < cap; i++)
                if (t.keys[i]) h[t.counts[i] > 10000 ? 10000 : t.counts[i] - 1]++;
#pragma omp critical
            for (int b = 0; b < 10001; b++) histo[b] += h[b];
            free(h);
        }
    }
    c = omp_get_wtime();
#pragma omp parallel
    {
        uint32_t* vals = NULL;
        uint64_t vcap = 0;
#pragma omp for schedule(dynamic, 256)
        for (uint64_t r = 0; r < n; r++) {
            const char* s = seq + off[r];
            const uint64_t len = off[r + 1] - off[r];
            const uint64_t nw = len >= (uint64_t)k ? len - (uint64_t)k + 1 : 0;
            if (nw > vcap) {
                free(vals);
                vcap = nw;
                vals = (uint32_t*)malloc(vcap * sizeof(uint32_t));
            }
            uint32_t m = 0;
            uint64_t sm = 0;
            CR_WINDOWS(s, len, k, canonical, {
                uint32_t v = 1;
                if (valid) {
                    const uint32_t cnt = cr_get(&t, key);
                    v = cnt >= min_count && cnt ? cnt : 1;
                }
                vals[m++] = v;
                sm += v;
            });
            nwin[r] = m;
            sum[r] = sm;
            if (m) std::nth_element(vals, vals + m / 2, vals + m);
            median[r] = m ? vals[m / 2] : 0;
        }
        free(vals);
    }
    double d = omp_get_wtime();
    *t_count = b - a;
    *t_stats = d - c;
    *n_distinct = nd;
    free(t.keys);
    free(t.counts);
    return 0;
}
