/*
 * tk_oracle_iworm.c -- CPU restatement of Inchworm's greedy contig
 * construction (TEST INFRASTRUCTURE ONLY; see tk_oracle.h).
 *
 * Trinity runs `inchworm --kmers <dump> --run_inchworm -K 25 -L 25 --DS ...`
 * (Trinity:2654-2711) over the k-mer catalog and reads its contigs as
 * ">a<N>;<cov> ..." records (filter_iworm_by_min_length_or_cov.pl:22-26,
 * scaffold_iworm_contigs.pl:67 requires /^a\d+;\d+/).  Inchworm's own source
 * is an empty submodule in the reference tree; what the tree states about the
 * method (Changelog.txt:500-514,1049-1061: seeds taken by abundance,
 * extension by the most abundant overlapping k-mer, used k-mers removed,
 * DS mode storing one of kmer/revcomp) is restated here, and the conventions
 * it leaves open are fixed as documented in tk_oracle.h (seed tie order, base
 * tie order, cov rounding).  Plain C99, one thread, binary search over the
 * sorted dictionary.
 */
#include <stdlib.h>
#include <string.h>

#include "tk_oracle.h"

static const uint64_t *g_keys;
static const uint32_t *g_counts;

static int seed_cmp(const void *a, const void *b) {
    const uint64_t i = *(const uint64_t *)a, j = *(const uint64_t *)b;
    if (g_counts[i] != g_counts[j]) return g_counts[i] > g_counts[j] ? -1 : 1;
    if (g_keys[i] != g_keys[j]) return g_keys[i] < g_keys[j] ? -1 : 1;
    return 0;
}

/* index of key in the sorted dictionary, or -1 */
static int64_t find(const uint64_t *keys, uint64_t n, uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && keys[lo] == key) ? (int64_t)lo : -1;
}

static int composition_keep(uint64_t key, int k, const uint8_t *keep) {
    int n[4] = {0, 0, 0, 0};
    for (int i = 0; i < k; i++) { n[key & 3]++; key >>= 2; }
    const int K1 = k + 1;
    return keep[(n[0] * K1 + n[1]) * K1 + n[2]] != 0;
}

/* The dictionary index of the oriented k-mer x (canonical form in DS mode). */
static int64_t lookup(const uint64_t *keys, const uint32_t *counts, uint64_t n, uint64_t x, int k, int ds) {
    uint64_t key = x;
    if (ds) {
        uint64_t rc = or_revcomp(x, k);
        if (rc < key) key = rc;
    }
    int64_t i = find(keys, n, key);
    if (i >= 0 && counts[i] == 0) return -1;
    return i;
}

/* One greedy extension step from oriented k-mer x to the right; returns the
 * chosen base (0..3) and its dictionary index, or -1 when no candidate is
 * present and unused. */
static int step_right(const uint64_t *keys, const uint32_t *counts, uint64_t n, const uint8_t *used,
                      uint64_t x, int k, int ds, int64_t *idx_out) {
    const uint64_t mask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    int best = -1;
    uint32_t best_c = 0;
    int64_t best_i = -1;
    for (int b = 0; b < 4; b++) {
        const uint64_t y = ((x << 2) | (uint64_t)b) & mask;
        const int64_t i = lookup(keys, counts, n, y, k, ds);
        if (i < 0 || used[i]) continue;
        if (best < 0 || counts[i] > best_c) {
            best = b;
            best_c = counts[i];
            best_i = i;
        }
    }
    *idx_out = best_i;
    return best;
}

/* The same step to the left of oriented k-mer y: candidates b + y[0..k-2]. */
static int step_left(const uint64_t *keys, const uint32_t *counts, uint64_t n, const uint8_t *used,
                     uint64_t y, int k, int ds, int64_t *idx_out) {
    int best = -1;
    uint32_t best_c = 0;
    int64_t best_i = -1;
    for (int b = 0; b < 4; b++) {
        const uint64_t z = ((uint64_t)b << (2 * (k - 1))) | (y >> 2);
        const int64_t i = lookup(keys, counts, n, z, k, ds);
        if (i < 0 || used[i]) continue;
        if (best < 0 || counts[i] > best_c) {
            best = b;
            best_c = counts[i];
            best_i = i;
        }
    }
    *idx_out = best_i;
    return best;
}

int64_t or_inchworm(const uint64_t *keys, const uint32_t *counts, uint64_t n, int k, int ds,
                    uint32_t min_seed_cov, const uint8_t *seed_keep, char *seq_out, uint64_t seq_cap,
                    uint64_t *off_out, uint64_t *sum_out, uint64_t *seed_out, uint64_t contig_cap,
                    uint64_t *n_contigs) {
    static const char B[4] = {'A', 'C', 'G', 'T'};
    uint64_t *order = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint8_t *used = (uint8_t *)calloc(n ? n : 1, 1);
    /* bases of the current contig: left part (reversed) and right part */
    char *left = (char *)malloc(n + (uint64_t)k + 1);
    char *right = (char *)malloc(n + (uint64_t)k + 1);
    if (!order || !used || !left || !right) {
        free(order); free(used); free(left); free(right);
        return -1;
    }
    uint64_t ns = 0;
    for (uint64_t i = 0; i < n; i++)
        if (counts[i] >= min_seed_cov && counts[i] > 0 && (!seed_keep || composition_keep(keys[i], k, seed_keep)))
            order[ns++] = i;
    g_keys = keys;
    g_counts = counts;
    qsort(order, ns, sizeof(uint64_t), seed_cmp);
    uint64_t nc = 0, len_total = 0;
    int64_t rc = 0;
    if (contig_cap) off_out[0] = 0;
    for (uint64_t s = 0; s < ns; s++) {
        const uint64_t si = order[s];
        if (used[si]) continue;
        used[si] = 1;
        uint64_t sum = counts[si];
        uint64_t nl = 0, nr = 0;
        /* right extension of the seed's forward string */
        uint64_t x = keys[si];
        for (;;) {
            int64_t i;
            const int b = step_right(keys, counts, n, used, x, k, ds, &i);
            if (b < 0) break;
            used[i] = 1;
            sum += counts[i];
            right[nr++] = B[b];
            x = ((x << 2) | (uint64_t)b) & (k == 32 ? ~0ull : ((1ull << (2 * k)) - 1));
        }
        /* left extension from the seed's first k-mer */
        x = keys[si];
        for (;;) {
            int64_t i;
            const int b = step_left(keys, counts, n, used, x, k, ds, &i);
            if (b < 0) break;
            used[i] = 1;
            sum += counts[i];
            left[nl++] = B[b];
            x = ((uint64_t)b << (2 * (k - 1))) | (x >> 2);
        }
        const uint64_t L = nl + (uint64_t)k + nr;
        if (nc < contig_cap && rc == 0 && len_total + L <= seq_cap) {
            char *o = seq_out + len_total;
            for (uint64_t j = 0; j < nl; j++) o[j] = left[nl - 1 - j];
            for (int j = 0; j < k; j++) o[nl + j] = B[(keys[si] >> (2 * (k - 1 - j))) & 3];
            memcpy(o + nl + k, right, nr);
            off_out[nc + 1] = len_total + L;
            sum_out[nc] = sum;
            seed_out[nc] = keys[si];
        } else {
            rc = -2;
        }
        len_total += L;
        nc++;
    }
    *n_contigs = nc;
    free(order); free(used); free(left); free(right);
    return rc < 0 ? -(int64_t)len_total - 1 : (int64_t)len_total;
}

/* ---- counts of a fixed k-mer set (GraphFromFasta weld k-mers, QuantifyGraph nodes) ---- */

typedef unsigned __int128 u128;

static const u128 *g_kkeys;
static int kkey_cmp(const void *a, const void *b) {
    const u128 x = g_kkeys[*(const uint64_t *)a], y = g_kkeys[*(const uint64_t *)b];
    return x < y ? -1 : x > y ? 1 : 0;
}

int or_kset_count(const uint64_t *keys, uint64_t n, const char *seq, const uint64_t *off, uint64_t n_reads, int kk,
                  int canonical, const uint32_t *tags, uint64_t *counts) {
    u128 *kv = (u128 *)malloc((n ? n : 1) * sizeof(u128));
    uint64_t *ord = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!kv || !ord) { free(kv); free(ord); return -1; }
    for (uint64_t i = 0; i < n; i++) {
        kv[i] = ((u128)keys[2 * i + 1] << 64) | keys[2 * i];
        ord[i] = i;
        counts[i] = 0;
    }
    g_kkeys = kv;
    qsort(ord, n, sizeof(uint64_t), kkey_cmp);
    for (uint64_t i = 1; i < n; i++)
        if (kv[ord[i]] == kv[ord[i - 1]]) { free(kv); free(ord); return -2; }
    const u128 mask = kk == 64 ? ~(u128)0 : (((u128)1 << (2 * kk)) - 1);
    for (uint64_t r = 0; r < n_reads; r++) {
        u128 f = 0, rc = 0;
        int run = 0;
        for (uint64_t p = off[r]; p < off[r + 1]; p++) {
            const int c = or_base_code((unsigned char)seq[p]);
            if (c < 0) {
                run = 0;
                f = (f << 2) & mask;
                rc = rc >> 2 | ((u128)3 << (2 * kk - 2));
                continue;
            }
            run++;
            f = ((f << 2) | (u128)c) & mask;
            rc = (rc >> 2) | ((u128)(3 - c) << (2 * kk - 2));
            if (run < kk) continue;
            u128 key = (canonical && rc < f) ? rc : f;
            if (tags) key = ((u128)tags[r] << 64) | (uint64_t)key;
            uint64_t lo = 0, hi = n;
            while (lo < hi) {
                const uint64_t mid = lo + (hi - lo) / 2;
                if (kv[ord[mid]] < key) lo = mid + 1;
                else hi = mid;
            }
            if (lo < n && kv[ord[lo]] == key) counts[ord[lo]]++;
        }
    }
    free(kv);
    free(ord);
    return 0;
}
