/*
 * tk_oracle.c -- CPU restatement of Trinity's k-mer hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Reference behaviour restated here (file:line in trinityrnaseq v2.15.2):
 *   counting      jellyfish count -m 25 [--canonical]      Trinity:2612-2619
 *                                                          util/insilico_read_normalization.pl:617-622
 *   dump / histo  jellyfish dump -L m ; jellyfish histo    Trinity:2625,2630 ; :636,641
 *   load + prune  inchworm --kmers ... --min_any_entropy   Trinity:2654,2676-2711,1050
 *   read stats    fastaToKmerCoverageStats --DS            util/insilico_read_normalization.pl:837-845
 *                 missing k-mer => 1, unsigned counts      Changelog.txt:536-537
 *   pair merge    nbkc_merge_left_right_stats.pl:80-145
 *   selection     nbkc_normalize.pl:67-127 (srand 12345, rand(1) <= max_cov/median)
 *   R2T vote      ReadsToTranscripts -p P                  Trinity:2238-2250 ; Changelog.txt:467,871-873
 *
 * Parity pins (tests/test_oracle.py): brute-force string-based Python
 * counters/voters on small inputs, closed forms (canon(rc(w)) = canon(w),
 * sum of counts = valid windows, histogram sums), Perl's own rand() stream
 * (tests/golden/perl_rand_12345.txt), golden vectors the reference's
 * nbkc_merge_left_right_stats.pl / nbkc_normalize.pl wrote
 * (tests/golden/nbkc/), the reference's kmer_histo witness files, and
 * textbook mean / population stdev for the stats row.  The reference's
 * sample reads (tests/golden/sample_data/) are run through the oracle and
 * the GPU path side by side; they check the GPU, not the oracle.
 * Conventions the reference tree does not fix (median of an even-length
 * vector, N-containing windows, stdev form, R2T tie rule and pct format)
 * are parameters; their default settings are parity unpinned.
 */
#include "tk_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int or_base_code(int ch) {
    switch (ch) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return -1;
    }
}

uint64_t or_revcomp(uint64_t v, int k) {
    uint64_t r = 0;
    for (int i = 0; i < k; i++) {
        r = (r << 2) | (3u - (v & 3u));
        v >>= 2;
    }
    return r;
}

uint64_t or_nwin(uint64_t len, int k) {
    return len >= (uint64_t)k ? len - (uint64_t)k + 1 : 0;
}

/* Rolling window walk shared by every pass: calls back for each window with
 * valid flag and forward value. */
typedef void (*win_fn)(void* ctx, uint64_t read, uint64_t pos, int valid,
                       uint64_t fwd);

static void walk_windows(const char* seq, uint64_t start, uint64_t len, int k,
                         uint64_t read, win_fn fn, void* ctx) {
    uint64_t mask = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
    uint64_t fwd = 0;
    int run = 0; /* consecutive valid bases ending here */
    for (uint64_t i = 0; i < len; i++) {
        int c = or_base_code((unsigned char)seq[start + i]);
        if (c < 0) {
            run = 0;
            fwd = 0;
        } else {
            fwd = ((fwd << 2) | (uint64_t)c) & mask;
            run++;
        }
        if (i + 1 >= (uint64_t)k) {
            uint64_t pos = i + 1 - (uint64_t)k;
            fn(ctx, read, pos, run >= k, fwd);
        }
    }
}

/* ---------- LSD radix sort of u64 keys with a u32 payload ---------- */
static void radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t n, int key_bits) {
    if (n < 2) return;
    uint64_t* tk = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint32_t* tv = vals ? (uint32_t*)malloc(n * sizeof(uint32_t)) : NULL;
    uint64_t* cnt = (uint64_t*)malloc(65536 * sizeof(uint64_t));
    for (int shift = 0; shift < key_bits; shift += 16) {
        memset(cnt, 0, 65536 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[(keys[i] >> shift) & 0xFFFF]++;
        uint64_t s = 0;
        for (int d = 0; d < 65536; d++) { uint64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t d = (keys[i] >> shift) & 0xFFFF;
            uint64_t p = cnt[d]++;
            tk[p] = keys[i];
            if (vals) tv[p] = vals[i];
        }
        memcpy(keys, tk, n * sizeof(uint64_t));
        if (vals) memcpy(vals, tv, n * sizeof(uint32_t));
    }
    free(tk); free(tv); free(cnt);
}

/* ---------- counting ---------- */
typedef struct { uint64_t* keys; uint64_t n; int k; int canonical; } collect_ctx;

static void collect_fn(void* vctx, uint64_t read, uint64_t pos, int valid, uint64_t fwd) {
    (void)read; (void)pos;
    collect_ctx* c = (collect_ctx*)vctx;
    if (!valid) return;
    uint64_t key = fwd;
    if (c->canonical) {
        uint64_t rc = or_revcomp(fwd, c->k);
        if (rc < key) key = rc;
    }
    c->keys[c->n++] = key;
}

void or_window_totals(const char* seq, const uint64_t* off, uint64_t n_reads,
                      int k, uint64_t* total, uint64_t* valid) {
    uint64_t t = 0, v = 0;
    for (uint64_t r = 0; r < n_reads; r++) {
        uint64_t len = off[r + 1] - off[r];
        t += or_nwin(len, k);
        int run = 0;
        for (uint64_t i = 0; i < len; i++) {
            int c = or_base_code((unsigned char)seq[off[r] + i]);
            run = (c < 0) ? 0 : run + 1;
            if (i + 1 >= (uint64_t)k && run >= k) v++;
        }
    }
    *total = t;
    *valid = v;
}

int64_t or_count(const char* seq, const uint64_t* off, uint64_t n_reads, int k,
                 int canonical, uint64_t* keys_out, uint32_t* counts_out,
                 uint64_t cap) {
    uint64_t total = 0;
    for (uint64_t r = 0; r < n_reads; r++) total += or_nwin(off[r + 1] - off[r], k);
    collect_ctx c;
    c.keys = (uint64_t*)malloc((total ? total : 1) * sizeof(uint64_t));
    c.n = 0; c.k = k; c.canonical = canonical;
    for (uint64_t r = 0; r < n_reads; r++)
        walk_windows(seq, off[r], off[r + 1] - off[r], k, r, collect_fn, &c);
    radix_sort_u64(c.keys, NULL, c.n, 2 * k);
    uint64_t nd = 0;
    for (uint64_t i = 0; i < c.n; i++)
        if (i == 0 || c.keys[i] != c.keys[i - 1]) nd++;
    if (nd > cap) { free(c.keys); return -(int64_t)nd; }
    uint64_t j = 0;
    for (uint64_t i = 0; i < c.n; i++) {
        if (i == 0 || c.keys[i] != c.keys[i - 1]) {
            keys_out[j] = c.keys[i];
            counts_out[j] = 1;
            j++;
        } else {
            counts_out[j - 1]++;
        }
    }
    free(c.keys);
    return (int64_t)nd;
}

/* ---------- histo ---------- */
/* jellyfish histo: base = low>1 ? (inc>=low ? 1 : low-inc) : 1; ceil = high+inc;
 * nb = (ceil + inc - base)/inc; bin label i -> base + i*inc. */
uint64_t or_histo_base(uint64_t low, uint64_t inc) {
    return low > 1 ? (inc >= low ? 1 : low - inc) : 1;
}
uint64_t or_histo_nbins(uint64_t low, uint64_t high, uint64_t inc) {
    uint64_t base = or_histo_base(low, inc);
    uint64_t ceil = high + inc;
    return (ceil + inc - base) / inc;
}
void or_histo(const uint32_t* counts, uint64_t n, uint64_t low, uint64_t high,
              uint64_t inc, uint64_t* bins) {
    uint64_t base = or_histo_base(low, inc);
    uint64_t ceil = high + inc;
    uint64_t nb = or_histo_nbins(low, high, inc);
    memset(bins, 0, nb * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v = counts[i];
        if (v < base) bins[0]++;
        else if (v >= ceil) bins[nb - 1]++;
        else bins[(v - base) / inc]++;
    }
}

/* ---------- entropy + prune ---------- */
double or_entropy_counts(int nA, int nC, int nG, int nT, int k) {
    int cs[4] = {nA, nC, nG, nT};
    double h = 0.0;
    for (int i = 0; i < 4; i++) {
        if (cs[i] <= 0) continue;
        double p = (double)cs[i] / (double)k;
        h -= p * log2(p);
    }
    return h;
}

void or_entropy_keep_table(int k, double min_entropy, uint8_t* keep) {
    int K1 = k + 1;
    for (int a = 0; a <= k; a++)
        for (int c = 0; c <= k; c++)
            for (int g = 0; g <= k; g++) {
                int t = k - a - c - g;
                uint8_t v = 0;
                if (t >= 0) v = or_entropy_counts(a, c, g, t, k) >= min_entropy;
                keep[(a * K1 + c) * K1 + g] = v;
            }
}

static void composition(uint64_t key, int k, int* a, int* c, int* g) {
    int n[4] = {0, 0, 0, 0};
    for (int i = 0; i < k; i++) { n[key & 3]++; key >>= 2; }
    *a = n[0]; *c = n[1]; *g = n[2];
}

uint64_t or_prune(const uint64_t* keys, const uint32_t* counts, uint64_t n,
                  int k, uint32_t min_count, double min_entropy,
                  uint64_t* keys_out, uint32_t* counts_out) {
    uint64_t j = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (counts[i] < min_count) continue;
        int a, c, g;
        composition(keys[i], k, &a, &c, &g);
        double h = or_entropy_counts(a, c, g, k - a - c - g, k);
        if (h < min_entropy) continue;
        keys_out[j] = keys[i];
        counts_out[j] = counts[i];
        j++;
    }
    return j;
}

/* ---------- per-read stats ---------- */
static int64_t bsearch_key(const uint64_t* keys, uint64_t n, uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < n && keys[lo] == key) ? (int64_t)lo : -1;
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

typedef struct {
    int k, canonical;
    const uint64_t* keys; const uint32_t* counts; uint64_t n;
    uint32_t min_count, missing;
    int nwin_policy;
    uint32_t* vals; uint64_t nv;
} stats_ctx;

static void stats_fn(void* vctx, uint64_t read, uint64_t pos, int valid, uint64_t fwd) {
    (void)read; (void)pos;
    stats_ctx* s = (stats_ctx*)vctx;
    if (!valid) {
        if (s->nwin_policy == OR_NWIN_AS_MISSING) s->vals[s->nv++] = s->missing;
        return;
    }
    uint64_t key = fwd;
    if (s->canonical) {
        uint64_t rc = or_revcomp(fwd, s->k);
        if (rc < key) key = rc;
    }
    int64_t ix = bsearch_key(s->keys, s->n, key);
    uint32_t v = s->missing;
    if (ix >= 0 && s->counts[ix] >= s->min_count) v = s->counts[ix];
    s->vals[s->nv++] = v;
}

void or_read_stats(const char* seq, const uint64_t* off, uint64_t n_reads,
                   int k, int canonical, const uint64_t* dict_keys,
                   const uint32_t* dict_counts, uint64_t n_dict,
                   uint32_t min_count, uint32_t missing_value,
                   int nwin_policy, int median_policy, or_read_stat* out) {
    uint64_t maxlen = 0;
    for (uint64_t r = 0; r < n_reads; r++)
        if (off[r + 1] - off[r] > maxlen) maxlen = off[r + 1] - off[r];
    stats_ctx s;
    s.k = k; s.canonical = canonical; s.keys = dict_keys; s.counts = dict_counts;
    s.n = n_dict; s.min_count = min_count; s.missing = missing_value;
    s.nwin_policy = nwin_policy;
    s.vals = (uint32_t*)malloc((maxlen + 1) * sizeof(uint32_t));
    for (uint64_t r = 0; r < n_reads; r++) {
        s.nv = 0;
        walk_windows(seq, off[r], off[r + 1] - off[r], k, r, stats_fn, &s);
        or_read_stat* o = &out[r];
        memset(o, 0, sizeof(*o));
        o->nwin = (uint32_t)s.nv;
        if (s.nv == 0) continue;
        unsigned __int128 sq = 0;
        uint64_t sum = 0;
        for (uint64_t i = 0; i < s.nv; i++) {
            sum += s.vals[i];
            sq += (unsigned __int128)s.vals[i] * s.vals[i];
        }
        o->sum = sum;
        o->sumsq_lo = (uint64_t)sq;
        o->sumsq_hi = (uint64_t)(sq >> 64);
        qsort(s.vals, s.nv, sizeof(uint32_t), cmp_u32);
        uint64_t n = s.nv;
        switch (median_policy) {
            case OR_MEDIAN_LOWER: o->median = s.vals[(n - 1) / 2]; break;
            case OR_MEDIAN_MEAN_FLOOR:
                o->median = (n & 1) ? s.vals[n / 2]
                                    : (uint32_t)(((uint64_t)s.vals[n / 2 - 1] + s.vals[n / 2]) / 2);
                break;
            default: o->median = s.vals[n / 2]; break;
        }
    }
    free(s.vals);
}

/* mean = sum/n ; population stdev = sqrt((n*sumsq - sum^2)) / n ; both as
 * doubles from the exact integers, printed %g (ostream default precision 6). */
int or_format_stats_row(const char* acc, const or_read_stat* st, char* buf, size_t bufsz) {
    double mean = 0.0, sd = 0.0, pct = 0.0;
    if (st->nwin > 0) {
        unsigned __int128 sq = ((unsigned __int128)st->sumsq_hi << 64) | st->sumsq_lo;
        unsigned __int128 n = st->nwin;
        unsigned __int128 num = n * sq - (unsigned __int128)st->sum * st->sum;
        mean = (double)st->sum / (double)st->nwin;
        sd = sqrt((double)num) / (double)st->nwin;
        pct = mean > 0 ? sd / mean * 100.0 : 0.0;
    }
    return snprintf(buf, bufsz, "%s\t%u\t%g\t%g\t%g\n", acc, st->median, mean, sd, pct);
}

/* ---------- drand48 as Perl's rand() ---------- */
#define DR_A 0x5DEECE66Dull
#define DR_C 0xBull
#define DR_M ((1ull << 48) - 1)

uint64_t or_drand48_seed(uint32_t seed) {
    return (((uint64_t)seed << 16) | 0x330Eull) & DR_M;
}
double or_drand48_next(uint64_t* state) {
    *state = (DR_A * *state + DR_C) & DR_M;
    return ldexp((double)*state, -48);
}

double or_perl_avg1(double a, double b) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.1f", (a + b) / 2);
    return strtod(buf, NULL);
}

uint64_t or_select_pairs(const double* l_med, const double* r_med,
                         const double* l_mean, const double* r_mean,
                         const double* l_sd, const double* r_sd, uint64_t n,
                         int max_cov, int min_cov, int max_cv,
                         uint8_t* selected) {
    uint64_t st = or_drand48_seed(12345);
    uint64_t nsel = 0;
    for (uint64_t i = 0; i < n; i++) {
        selected[i] = 0;
        double med = or_perl_avg1(l_med[i], r_med[i]);
        if (med < (double)min_cov) continue;
        double sd = or_perl_avg1(l_sd[i], r_sd[i]);
        double u = or_perl_avg1(l_mean[i], r_mean[i]);
        if (u <= 0) continue;
        double cv = sd / u;
        if (cv > (double)max_cv) continue;
        double r = or_drand48_next(&st);
        if (r <= (double)max_cov / med) { selected[i] = 1; nsel++; }
    }
    return nsel;
}

/* ---------- ReadsToTranscripts ---------- */
typedef struct { uint64_t* keys; uint32_t* comps; uint64_t n; uint32_t comp; int k, canonical; } r2tb_ctx;

static void r2tb_fn(void* vctx, uint64_t read, uint64_t pos, int valid, uint64_t fwd) {
    (void)read; (void)pos;
    r2tb_ctx* c = (r2tb_ctx*)vctx;
    if (!valid) return;
    uint64_t key = fwd;
    if (c->canonical) {
        uint64_t rc = or_revcomp(fwd, c->k);
        if (rc < key) key = rc;
    }
    c->keys[c->n] = key;
    c->comps[c->n] = c->comp;
    c->n++;
}

int64_t or_r2t_build(const char* seq, const uint64_t* off, const uint32_t* comp,
                     uint64_t n_bundles, int k, int canonical,
                     uint64_t* keys_out, uint32_t* comp_out, uint64_t cap) {
    uint64_t total = 0;
    for (uint64_t r = 0; r < n_bundles; r++) total += or_nwin(off[r + 1] - off[r], k);
    r2tb_ctx c;
    c.keys = (uint64_t*)malloc((total ? total : 1) * sizeof(uint64_t));
    c.comps = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
    c.n = 0; c.k = k; c.canonical = canonical;
    for (uint64_t r = 0; r < n_bundles; r++) {
        c.comp = comp[r];
        walk_windows(seq, off[r], off[r + 1] - off[r], k, r, r2tb_fn, &c);
    }
    radix_sort_u64(c.keys, c.comps, c.n, 2 * k);
    uint64_t nd = 0;
    for (uint64_t i = 0; i < c.n; i++)
        if (i == 0 || c.keys[i] != c.keys[i - 1]) nd++;
    if (nd > cap) { free(c.keys); free(c.comps); return -(int64_t)nd; }
    uint64_t j = 0;
    for (uint64_t i = 0; i < c.n; i++) {
        if (i == 0 || c.keys[i] != c.keys[i - 1]) {
            keys_out[j] = c.keys[i];
            comp_out[j] = c.comps[i];
            j++;
        } else if (c.comps[i] < comp_out[j - 1]) {
            comp_out[j - 1] = c.comps[i];
        }
    }
    free(c.keys); free(c.comps);
    return (int64_t)nd;
}

typedef struct {
    int k, canonical; const uint64_t* keys; const uint32_t* comps; uint64_t n;
    uint32_t* hit; uint64_t nh;
} r2tv_ctx;

static void r2tv_fn(void* vctx, uint64_t read, uint64_t pos, int valid, uint64_t fwd) {
    (void)read; (void)pos;
    r2tv_ctx* c = (r2tv_ctx*)vctx;
    if (!valid) return;
    uint64_t key = fwd;
    if (c->canonical) {
        uint64_t rc = or_revcomp(fwd, c->k);
        if (rc < key) key = rc;
    }
    int64_t ix = bsearch_key(c->keys, c->n, key);
    if (ix >= 0) c->hit[c->nh++] = c->comps[ix];
}

void or_r2t_vote_reads(const char* seq, const uint64_t* off, uint64_t n_reads,
                       int k, int canonical, const uint64_t* keys,
                       const uint32_t* comps, uint64_t n_keys,
                       or_r2t_vote* out) {
    uint64_t maxlen = 0;
    for (uint64_t r = 0; r < n_reads; r++)
        if (off[r + 1] - off[r] > maxlen) maxlen = off[r + 1] - off[r];
    r2tv_ctx c;
    c.k = k; c.canonical = canonical; c.keys = keys; c.comps = comps; c.n = n_keys;
    c.hit = (uint32_t*)malloc((maxlen + 1) * sizeof(uint32_t));
    for (uint64_t r = 0; r < n_reads; r++) {
        c.nh = 0;
        uint64_t len = off[r + 1] - off[r];
        walk_windows(seq, off[r], len, k, r, r2tv_fn, &c);
        out[r].comp = 0xFFFFFFFFu;
        out[r].hits = 0;
        out[r].nwin = (uint32_t)or_nwin(len, k);
        if (c.nh == 0) continue;
        qsort(c.hit, c.nh, sizeof(uint32_t), cmp_u32);
        uint32_t best = c.hit[0], bestn = 0;
        uint64_t i = 0;
        while (i < c.nh) {
            uint64_t j = i;
            while (j < c.nh && c.hit[j] == c.hit[i]) j++;
            if ((uint32_t)(j - i) > bestn) { bestn = (uint32_t)(j - i); best = c.hit[i]; }
            i = j;
        }
        out[r].comp = best;
        out[r].hits = bestn;
    }
    free(c.hit);
}

int or_r2t_passes(const or_r2t_vote* v, uint32_t min_pct) {
    if (v->hits == 0 || v->nwin == 0) return 0;
    return (uint64_t)100 * v->hits >= (uint64_t)min_pct * v->nwin;
}
