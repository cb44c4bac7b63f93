/*
 * tk_oracle.h -- CPU restatement of Trinity's k-mer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the parity tests, the
 * smoke check and the bench's cpu_baseline leg compare the HIP library
 * against.  Nothing in the product (trinityrnaseq_amd/) links, loads or
 * calls it.
 *
 * Every function is plain, single-threaded C99 over host arrays and cites
 * the reference call site / consumer whose behaviour it restates.  The tools
 * themselves (jellyfish 2.x, Inchworm's fastaToKmerCoverageStats and
 * inchworm, Chrysalis' ReadsToTranscripts, GraphFromFasta, QuantifyGraph)
 * are absent from the reference tree, so conventions the reference tree
 * cannot fix are explicit parameters (see the *_policy arguments) or stated
 * at the function.  Pins (tests/test_oracle.py): string brute force for
 * counting, histo, prune, stats, R2T, Inchworm and the k-mer set counts;
 * Perl's own rand() and golden vectors written by the reference's Perl
 * scripts for the selection; the reference's histogram files for histo's
 * default shape; textbook mean / population stdev for the stats row.
 * Parity unpinned (no reference binary or test fixes them): the stats
 * tool's columns and number format, even-length median and N-window
 * conventions; R2T's pct format and tie rule; or_inchworm's seed and base
 * tie orders, seed thresholds and the cov rounding of its callers; the weld
 * rule GraphFromFasta applies on top of or_kset_count.
 */
#ifndef TK_ORACLE_H
#define TK_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2-bit code: A/a=0 C/c=1 G/g=2 T/t=3, anything else -1 (breaks windows). */
int or_base_code(int ch);
uint64_t or_revcomp(uint64_t v, int k);

/* Number of k-windows of a read of length len (0 if len < k). */
uint64_t or_nwin(uint64_t len, int k);

/*
 * Canonical (or forward) k-mer counting -- `jellyfish count -m K [-C]`
 * (Trinity:2612-2619, util/insilico_read_normalization.pl:617-622).
 * Writes distinct keys in ascending order with their counts.  Returns the
 * number of distinct keys, or -(needed) if cap is too small.
 */
int64_t or_count(const char* seq, const uint64_t* off, uint64_t n_reads, int k,
                 int canonical, uint64_t* keys_out, uint32_t* counts_out,
                 uint64_t cap);

/* Total windows and N-free windows of a read set. */
void or_window_totals(const char* seq, const uint64_t* off, uint64_t n_reads,
                      int k, uint64_t* total, uint64_t* valid);

/*
 * `jellyfish histo -l low -h high -i inc` bins (Trinity:2630).
 * nbins = (high + inc + inc - base) / inc with base as in jellyfish;
 * counts below base go to bin 0, counts >= high+inc to the last bin.
 */
uint64_t or_histo_nbins(uint64_t low, uint64_t high, uint64_t inc);
uint64_t or_histo_base(uint64_t low, uint64_t inc);
void or_histo(const uint32_t* counts, uint64_t n, uint64_t low, uint64_t high,
              uint64_t inc, uint64_t* bins);

/* Base-composition Shannon entropy (bits) of a k-mer with the given counts. */
double or_entropy_counts(int nA, int nC, int nG, int nT, int k);
/* keep[(a*(k+1)+c)*(k+1)+g] = entropy >= min_entropy (t = k-a-c-g). */
void or_entropy_keep_table(int k, double min_entropy, uint8_t* keep);

/*
 * Inchworm dictionary load + prune (Trinity:2654,2676-2711,1050): drop
 * count < min_count, drop entropy < min_entropy.  Input sorted; output
 * stays sorted.  Returns kept count.
 */
uint64_t or_prune(const uint64_t* keys, const uint32_t* counts, uint64_t n,
                  int k, uint32_t min_count, double min_entropy,
                  uint64_t* keys_out, uint32_t* counts_out);

/* Conventions of fastaToKmerCoverageStats that the reference tree leaves open. */
enum { OR_MEDIAN_UPPER = 0, OR_MEDIAN_LOWER = 1, OR_MEDIAN_MEAN_FLOOR = 2 };
enum { OR_NWIN_AS_MISSING = 0, OR_NWIN_SKIP = 1 };

typedef struct {
    uint32_t median;
    uint32_t nwin;       /* windows that contributed a value */
    uint64_t sum;
    uint64_t sumsq_lo;   /* 128-bit sum of squares */
    uint64_t sumsq_hi;
} or_read_stat;

/*
 * fastaToKmerCoverageStats per-read coverage
 * (util/insilico_read_normalization.pl:837-845; Changelog.txt:536-537:
 * counts unsigned, missing k-mers count as 1).  dict_keys sorted ascending.
 * A window's value is dict count if present and >= min_count, else
 * missing_value.  Windows with a non-ACGT base follow nwin_policy.
 */
void or_read_stats(const char* seq, const uint64_t* off, uint64_t n_reads,
                   int k, int canonical, const uint64_t* dict_keys,
                   const uint32_t* dict_counts, uint64_t n_dict,
                   uint32_t min_count, uint32_t missing_value,
                   int nwin_policy, int median_policy, or_read_stat* out);

/* Format one stats row "acc\tmedian\tmean\tstdev\tpct_dev\n"; returns length. */
int or_format_stats_row(const char* acc, const or_read_stat* st, char* buf,
                        size_t bufsz);

/* ---- Perl rand(): drand48 (nbkc_normalize.pl:69,111) ---- */
uint64_t or_drand48_seed(uint32_t seed);              /* srand(seed) state */
double or_drand48_next(uint64_t* state);              /* rand(1) */

/*
 * nbkc_merge_left_right_stats.pl --sorted + nbkc_normalize.pl over paired
 * numeric stats that are already in sorted-accession order and already
 * paired (pair i = left i, right i).  The medians/means/stdevs given are the
 * values as printed by the stats tool (doubles parsed from its text).
 * selected[i] = 1 when pair i is printed by nbkc_normalize.pl.
 * Returns the number selected.
 */
uint64_t or_select_pairs(const double* l_med, const double* r_med,
                         const double* l_mean, const double* r_mean,
                         const double* l_sd, const double* r_sd, uint64_t n,
                         int max_cov, int min_cov, int max_cv,
                         uint8_t* selected);

/* sprintf("%.1f", (a+b)/2) parsed back, as Perl numifies the merged column. */
double or_perl_avg1(double a, double b);

/* ---- ReadsToTranscripts (Trinity:2238-2250) ---- */
/*
 * Bundle k-mer table: every valid window of every bundle sequence
 * (">s_<comp>" records, iworm contigs joined by 'X' -- add_LR_reads_to_iworm_bundle.pl:26,34-42)
 * maps key -> smallest component id holding it.  Sorted by key.
 * Returns number of distinct keys or -(needed).
 */
int64_t or_r2t_build(const char* seq, const uint64_t* off, const uint32_t* comp,
                     uint64_t n_bundles, int k, int canonical,
                     uint64_t* keys_out, uint32_t* comp_out, uint64_t cap);

typedef struct {
    uint32_t comp;   /* best component, 0xFFFFFFFF if none */
    uint32_t hits;   /* windows voting for it */
    uint32_t nwin;   /* all windows of the read (len-k+1) */
} or_r2t_vote;

/* Per-read vote: most-hit component, ties -> smallest id. */
void or_r2t_vote_reads(const char* seq, const uint64_t* off, uint64_t n_reads,
                       int k, int canonical, const uint64_t* keys,
                       const uint32_t* comps, uint64_t n_keys,
                       or_r2t_vote* out);

/* pct >= min_pct test done exactly in integers: 100*hits >= min_pct*nwin. */
int or_r2t_passes(const or_r2t_vote* v, uint32_t min_pct);

/* ---- Inchworm greedy contig construction (tk_oracle_iworm.c) ---- */
/*
 * `inchworm --kmers <dump> --run_inchworm -K k -L L [--DS]` (Trinity:2654-2711)
 * over a dictionary (keys ascending, canonical in DS mode; count 0 = pruned).
 * Seeds: keys with count >= min_seed_cov (and seed_keep[composition], the
 * seed-entropy table, when given), taken in order of count descending, then
 * key ascending; a seed already used by an earlier contig is skipped.
 * A contig starts as the seed's forward string (the key), then extends to the
 * right: from its last k-mer x the candidates are x[1..k-1]+b, b = A,C,G,T; a
 * candidate is available when it is in the dictionary with count > 0 and not
 * used (DS: looked up and marked by canonical form); the one with the largest
 * count is appended (ties: smallest base), marked used, and the walk goes on
 * until no candidate is available.  Then the same to the left of the seed
 * (candidates b+y[0..k-2], ties to the smallest base).  Every contig is
 * returned in seed order, with the sum of its k-mers' counts and its seed key
 * (the CLI reports those of length >= L as ">a<N>;<floor(sum/nkmers)>").
 * seq_out gets the contigs' bases back to back, off_out[i..i+1] bounds
 * contig i.  Returns the total length, or -(total+1) when a cap was short
 * (n_contigs is still set).
 */
int64_t or_inchworm(const uint64_t* keys, const uint32_t* counts, uint64_t n, int k, int ds,
                    uint32_t min_seed_cov, const uint8_t* seed_keep, char* seq_out, uint64_t seq_cap,
                    uint64_t* off_out, uint64_t* sum_out, uint64_t* seed_out, uint64_t contig_cap,
                    uint64_t* n_contigs);

/*
 * Counts of a fixed set of keys in the reads' windows (tk_oracle_iworm.c):
 * keys[2i] = lo, keys[2i+1] = hi of key i.  Without tags a key is the
 * 2kk-bit value of a kk-mer (kk <= 64); with tags (one per read, kk <= 32)
 * the key of a read's window is (hi = tags[read], lo = k-mer).  Windows with
 * a non-ACGT base do not count; canonical keys are min(k-mer, revcomp).
 * Returns 0, -2 when the keys are not distinct.
 */
int or_kset_count(const uint64_t* keys, uint64_t n, const char* seq, const uint64_t* off, uint64_t n_reads,
                  int kk, int canonical, const uint32_t* tags, uint64_t* counts);

#ifdef __cplusplus
}
#endif
#endif
