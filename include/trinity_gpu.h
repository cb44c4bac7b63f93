/*
 * trinity_gpu.h -- C ABI of the MI355X (gfx950) k-mer hot path of Trinity.
 *
 * The reference boundary is process-level: the Perl driver runs the tools
 * below with system("bash","-c",cmd) (Trinity:2910-2926) and reads their
 * files.  This ABI is what those tools' inner loops become; the CLI shims in
 * trinityrnaseq_amd/csrc/tools/ keep the reference command lines and file
 * formats and call only these entry points.
 *
 *   reference tool / call site                             -> entry points
 *   jellyfish count -m 25 -s S [--canonical] reads.fa
 *        Trinity:2612-2619, insilico_read_normalization.pl:617-622
 *                                                -> tk_reads_from_host, tk_estimate_distinct,
 *                                                   tk_table_create, tk_count
 *   jellyfish dump -L m db   (Trinity:2625, :641) -> tk_table_dump_size, tk_table_dump
 *   jellyfish histo db       (Trinity:2630, :636) -> tk_table_histo
 *   inchworm --kmers dump --DS --min_any_entropy E [--minKmerCount m]
 *        Trinity:2654,2669-2711,1050             -> tk_table_load, tk_table_prune
 *   inchworm --run_inchworm -L L (greedy contigs, Trinity:2654-2711)
 *                                                -> tk_inchworm, tk_contigs_*
 *   fastaToKmerCoverageStats --reads --kmers --kmer_size --DS
 *        insilico_read_normalization.pl:837-845  -> tk_read_cov_stats
 *   nbkc_merge_left_right_stats.pl + nbkc_normalize.pl
 *        nbkc_normalize.pl:67-127, nbkc_merge_left_right_stats.pl:80-145
 *                                                -> tk_select_pairs
 *   ReadsToTranscripts -i reads -f bundles -p P [-strand]
 *        Trinity:2238-2250                       -> tk_ctable_build, tk_r2t_vote
 *   GraphFromFasta -k 24 -kk 48 (weld support, Trinity:2180)
 *   QuantifyGraph -k 24 (per-component node weights, Trinity:2352)
 *                                                -> tk_kset_create, tk_kset_count
 *
 * Conventions
 *   - Every function returns TK_OK (0) or a negative tk_status; the message
 *     of the last failure on the calling thread is tk_last_error().
 *   - The caller owns every host buffer.  The library owns device objects
 *     (tk_reads, tk_table, tk_ctable) until their *_free call.
 *   - A tk_ctx is one device + one HIP stream; it is not re-entrant.  Use one
 *     context per host thread; objects may be shared by contexts of the
 *     same device once the producing call has returned.
 *   - Functions with a _d suffix take/return DEVICE pointers and are
 *     asynchronous on the context's stream (call tk_sync before reading).
 *   - k-mer keys are 2-bit packed, first base most significant, A=0 C=1
 *     G=2 T=3 (so numeric order = lexicographic order).  Canonical key =
 *     min(kmer, reverse complement).  1 <= k <= 31.
 */
#ifndef TRINITY_GPU_H
#define TRINITY_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TK_ABI_VERSION 1

typedef enum {
    TK_OK = 0,
    TK_ERR_ARG = -1,    /* invalid argument or unsupported size            */
    TK_ERR_HIP = -2,    /* HIP runtime error                               */
    TK_ERR_OOM = -3,    /* device allocation failed                        */
    TK_ERR_FULL = -4,   /* hash table capacity exhausted (grow and retry)  */
    TK_ERR_STATE = -5,  /* object in the wrong state for this call         */
    TK_ERR_CAP = -6     /* caller's output buffer too small                */
} tk_status;

typedef struct tk_ctx tk_ctx;
typedef struct tk_reads tk_reads;
typedef struct tk_table tk_table;
typedef struct tk_ctable tk_ctable;

/* ---------------- context ---------------- */
const char* tk_last_error(void);
int tk_abi_version(void);
int tk_init(int device, tk_ctx** out);
int tk_ctx_free(tk_ctx* ctx);
int tk_sync(tk_ctx* ctx);
/* Per-kernel HIP-event timing on the context's stream. */
int tk_timing_enable(tk_ctx* ctx, int on);
int tk_timing_reset(tk_ctx* ctx);
/* Sum of event-measured durations and launch count of one kernel name. */
int tk_timing_get(tk_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);
/* Device memory currently held by objects of this process (bytes). */
int tk_device_bytes_in_use(uint64_t* bytes);
/* Words (distinct keys per merge, with their occurrences) that the last
   tk_count on this context streamed through its region-resident passes; 0
   when it counted with table atomics.  Diagnostic, for the traffic figures of
   the count (no counterpart in the reference). */
int tk_count_words(tk_ctx* ctx, uint64_t* words);

/* Keep the device blocks that the calling thread frees (tables, read sets,
   buffers) on a free list of at most limit_bytes and hand them to its next
   allocations of a similar size, instead of hipFree + hipMalloc per job:
   hipFree waits for the whole device, and a large hipMalloc after large
   frees can stall for seconds.  A reused block is not cleared (tables track
   their own clearing).  limit_bytes = 0 releases the list and turns reuse
   off.  The thread must issue all its work through one context: stream
   order is what makes a reused block safe without a wait.  Server workers
   do this by themselves.  No counterpart in the reference (its tools run
   one process per job). */
int tk_thread_cache(uint64_t limit_bytes);

/* Plain device buffers (inputs kept resident in HBM, outputs of _d calls). */
int tk_buffer_alloc(tk_ctx* ctx, uint64_t bytes, void** d_ptr);
int tk_buffer_free(tk_ctx* ctx, void* d_ptr);
int tk_memcpy_h2d(tk_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
int tk_memcpy_d2h(tk_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
/* Device-to-device copy on the context stream (also a known-size traffic
 * reference when reading HBM byte counters). */
int tk_memcpy_d2d(tk_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);
/* Page-locked host memory: reads uploaded from it (tk_reads_from_host,
 * tk_server_submit) move by DMA without a staging copy. */
int tk_host_alloc(uint64_t bytes, void** h_ptr);
int tk_host_free(void* h_ptr);

/* ---------------- reads (device-resident, 2-bit packed) ---------------- */
/*
 * seq: concatenated read sequences (ASCII, no separators); offsets[n+1]:
 * byte offset of each read in seq.  Bases A/C/G/T in either case are
 * valid; any other byte (N, X, ...) is invalid and no window spans it.
 * Replaces the per-read parse loops of the tools above (reads FASTA as
 * produced by Trinity:2746-2880 / Fasta_reader.pm:45-73).
 */
int tk_reads_from_host(tk_ctx* ctx, const char* seq, const uint64_t* offsets,
                       uint64_t n_reads, tk_reads** out);
/* Same from device-resident buffers (d_seq needs >= offsets[n] bytes). */
int tk_reads_from_device_d(tk_ctx* ctx, const char* d_seq, const uint64_t* d_offsets,
                           uint64_t n_reads, tk_reads** out);
int tk_reads_info(const tk_reads* r, uint64_t* n_reads, uint64_t* n_bases,
                  uint32_t* max_len);
/* Windows of length k over all reads: total (len-k+1 summed). */
int tk_reads_windows(const tk_reads* r, int k, uint64_t* total);
int tk_reads_free(tk_reads* r);

/* ---------------- k-mer -> count table ---------------- */
/*
 * HyperLogLog estimate of the distinct (canonical) k-mers of a read set,
 * used to size tables (jellyfish's -s hint, Trinity:2598-2606) and to admit
 * jobs by device footprint.
 */
int tk_estimate_distinct(tk_ctx* ctx, const tk_reads* reads, int k, int canonical,
                         double* estimate);
/*
 * Create an empty table for k-mers of length k that holds at least
 * expected_distinct keys at load factor <= max_load (0 < max_load <= 0.9).
 */
int tk_table_create(tk_ctx* ctx, int k, int canonical, uint64_t expected_distinct,
                    double max_load, tk_table** out);
int tk_table_info(const tk_table* t, int* k, int* canonical, uint64_t* capacity,
                  uint64_t* device_bytes);
/* Count every N-free window of reads (jellyfish count).  Accumulates.  Large
   counts (>= 2^32 windows) group their merged keys by table region and count
   each region in LDS, keeping two word streams as context scratch (freed by
   tk_ctx_free); others add merged keys to the table with atomics. */
int tk_count(tk_ctx* ctx, tk_table* t, const tk_reads* reads);
/* Number of keys whose count lies in [min_count, max_count]. */
int tk_table_dump_size(tk_ctx* ctx, tk_table* t, uint32_t min_count, uint32_t max_count,
                       uint64_t* n);
/* Keys with count in [min_count, max_count], ascending key order (jellyfish dump -L -U). */
int tk_table_dump(tk_ctx* ctx, tk_table* t, uint32_t min_count, uint32_t max_count,
                  uint64_t* keys, uint32_t* counts, uint64_t cap, uint64_t* n_out);
/*
 * jellyfish histo -l low -h high -i inc: bins[i] counts keys with count in
 * [base+i*inc, base+(i+1)*inc), base = low>1 ? (inc>=low ? 1 : low-inc) : 1;
 * bin 0 also takes counts < base, the last bin counts >= high+inc.
 * nbins must equal (high + 2*inc - base) / inc.
 */
int tk_table_histo(tk_ctx* ctx, tk_table* t, uint64_t low, uint64_t high, uint64_t inc,
                   uint64_t* bins, uint64_t nbins);
/* Insert (key,count) pairs, e.g. a parsed jellyfish dump (inchworm --kmers).
 * Keys must already be canonical for a canonical table; counts add up. */
int tk_table_load(tk_ctx* ctx, tk_table* t, const uint64_t* keys, const uint32_t* counts,
                  uint64_t n);
/*
 * Inchworm prune: remove keys with count < min_count or whose base
 * composition is rejected by keep_table ((k+1)^3 bytes, index
 * (nA*(k+1)+nC)*(k+1)+nG, nonzero = keep).  NULL keep_table keeps all.
 */
int tk_table_prune(tk_ctx* ctx, tk_table* t, uint32_t min_count, const uint8_t* keep_table,
                   uint64_t* n_kept);
/* Look up keys (already canonical for a canonical table); 0 = absent. */
int tk_table_lookup(tk_ctx* ctx, tk_table* t, const uint64_t* keys, uint64_t n,
                    uint32_t* counts);
int tk_table_free(tk_table* t);

/* ---------------- counts of a fixed k-mer set ---------------- */
/*
 * Occurrences in the reads of a given set of keys, each 128 bits as two
 * words (lo, hi), hi < 2^62, all distinct (TK_ERR_ARG otherwise):
 *  - plain kk-mers, kk <= 63 (tags NULL): the key is the 2kk-bit k-mer
 *    value, hi = its bits 64..127; GraphFromFasta's weld k-mers
 *    (-k 24 -kk 48, Trinity:2180) counted in the reads;
 *  - tagged k-mers, kk <= 32 (tags = one u32 per read): the key of a read's
 *    window is (hi = tags[read], lo = k-mer); QuantifyGraph's node (k = 24)
 *    and edge (k = 25) weights of each component from the reads
 *    ReadsToTranscripts gave it (Trinity:2352).
 * A window counts when all its bases are ACGT; canonical = min(k-mer,
 * reverse complement) in numeric order.  counts[i] = windows equal to key i.
 */
typedef struct tk_kset tk_kset;
int tk_kset_create(tk_ctx* ctx, const uint64_t* keys, uint64_t n_keys, tk_kset** out);
int tk_kset_count(tk_ctx* ctx, tk_kset* s, const tk_reads* reads, int kk, int canonical,
                  const uint32_t* tags, uint64_t* counts);
int tk_kset_free(tk_kset* s);

/* ---------------- Inchworm contigs ---------------- */
/*
 * Greedy contig construction over a (pruned) table, the assembly step of
 * `inchworm --kmers <dump> --run_inchworm -K k -L L [--DS]`
 * (Trinity:2654-2711; contigs read as ">a<N>;<cov>" by
 * filter_iworm_by_min_length_or_cov.pl:22-26).  Seeds: keys with count >=
 * min_seed_cov whose composition seed_keep accepts ((k+1)^3 bytes as in
 * tk_table_prune; NULL accepts all), in order of count descending, then key
 * ascending; a seed used by an earlier contig is skipped.  Each contig
 * starts as its seed's forward string, extends to the right by the unused
 * candidate k-mer (last k-1 bases + b) of largest count, ties to the
 * smallest base, until none is left, then to the left the same way
 * (b + first k-1 bases); lookups and use are by canonical form on a
 * canonical table.  Contigs come in seed order, each with the sum of its
 * k-mers' counts and its seed key; the table is not modified.  Tables of
 * 2^32 or more slots are refused.
 */
typedef struct tk_contigs tk_contigs;
int tk_inchworm(tk_ctx* ctx, tk_table* t, uint32_t min_seed_cov, const uint8_t* seed_keep,
                tk_contigs** out);
/* Number of contigs, total bases, and the rounds / walks the device ran. */
int tk_contigs_info(const tk_contigs* c, uint64_t* n_contigs, uint64_t* total_len, uint64_t* rounds,
                    uint64_t* walks);
/* seq: total_len bytes (no separators); offsets: n+1; sums, seed_keys: n. */
int tk_contigs_get(const tk_contigs* c, char* seq, uint64_t* offsets, uint64_t* sums, uint64_t* seed_keys);
int tk_contigs_free(tk_contigs* c);

/* ---------------- per-read k-mer coverage (fastaToKmerCoverageStats) ---------------- */
enum { TK_MEDIAN_UPPER = 0, TK_MEDIAN_LOWER = 1, TK_MEDIAN_MEAN_FLOOR = 2 };
enum { TK_NWIN_AS_MISSING = 0, TK_NWIN_SKIP = 1 };

typedef struct {
    uint32_t min_count;      /* dictionary entries below this count as missing (dump -L)  */
    uint32_t missing_value;  /* value of a missing k-mer (1: Changelog.txt:537)           */
    int32_t nwin_policy;     /* TK_NWIN_*: windows holding a non-ACGT base                */
    int32_t median_policy;   /* TK_MEDIAN_*: median of an even number of values           */
} tk_stats_opts;

typedef struct {
    uint32_t median;
    uint32_t nwin;           /* values that entered the statistics */
    uint64_t sum;
    uint64_t sumsq_lo;       /* 128-bit sum of squares */
    uint64_t sumsq_hi;
} tk_read_stat;

/* One tk_read_stat per read, in read order. */
int tk_read_cov_stats(tk_ctx* ctx, tk_table* t, const tk_reads* reads,
                      const tk_stats_opts* opts, tk_read_stat* out);
int tk_read_cov_stats_d(tk_ctx* ctx, tk_table* t, const tk_reads* reads,
                        const tk_stats_opts* opts, tk_read_stat* d_out);

/* ---------------- pair merge + normalization selection ---------------- */
/*
 * nbkc_merge_left_right_stats.pl --sorted followed by nbkc_normalize.pl
 * (srand(12345); keep pair when rand(1) <= max_cov/median) for pairs that
 * are already in sorted-accession order with left i paired to right i.
 * Requires stats computed with missing_value >= 1 and
 * max_cv >= 1.2*sqrt(max nwin) + 1 (then the coefficient-of-variation and
 * zero-mean tests depend on nwin only and the selection is exact); a read
 * with windows whose mean is below 1, or a drawn pair whose merged median
 * is 0, fails with TK_ERR_ARG.  d_selected[i] = 1 when pair i is kept.
 */
int tk_select_pairs_d(tk_ctx* ctx, const tk_read_stat* d_left, const tk_read_stat* d_right,
                      uint64_t n_pairs, int max_cov, int min_cov, int max_cv,
                      uint32_t max_nwin, uint8_t* d_selected, uint64_t* n_selected);
int tk_select_pairs(tk_ctx* ctx, const tk_read_stat* left, const tk_read_stat* right,
                    uint64_t n_pairs, int max_cov, int min_cov, int max_cv,
                    uint8_t* selected, uint64_t* n_selected);
/*
 * tk_select_pairs_d without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as tk_select_pairs_d, for at most 2^24 pairs (more fail
 * with TK_ERR_ARG); its working memory (flags, per-pair draw counts and
 * ranks, scan storage: 132 MB) is a static device array that each GPU holds
 * until the process ends.  All calls on one GPU share that array: calls that
 * overlap on one GPU (other threads, other contexts) race.
 */
int tk_select_pairs_d_noalloc(tk_ctx* ctx, const tk_read_stat* d_left, const tk_read_stat* d_right,
                              uint64_t n_pairs, int max_cov, int min_cov, int max_cv,
                              uint32_t max_nwin, uint8_t* d_selected, uint64_t* n_selected);

/* ---------------- ReadsToTranscripts ---------------- */
/*
 * k-mer -> component table from bundle sequences (">s_<comp>" records of
 * iworm contigs joined by 'X'; add_LR_reads_to_iworm_bundle.pl:26,34-42).
 * comp_ids[i] is the component of bundle i.  A k-mer present in several
 * components maps to the smallest id.
 */
int tk_ctable_build(tk_ctx* ctx, int k, int canonical, const tk_reads* bundles,
                    const uint32_t* comp_ids, tk_ctable** out);
int tk_ctable_info(const tk_ctable* c, uint64_t* n_keys, uint64_t* capacity);

typedef struct {
    uint32_t comp;   /* most-voted component (ties: smallest id); 0xFFFFFFFF = none */
    uint32_t hits;   /* windows voting for comp */
    uint32_t nwin;   /* all windows of the read, len-k+1 */
} tk_r2t_vote;

int tk_r2t_vote_reads(tk_ctx* ctx, const tk_ctable* c, const tk_reads* reads,
                      tk_r2t_vote* out);
int tk_r2t_vote_reads_d(tk_ctx* ctx, const tk_ctable* c, const tk_reads* reads,
                        tk_r2t_vote* d_out);
int tk_ctable_free(tk_ctable* c);

/* ---------------- batched components (Trinity phase 2) ---------------- */
/*
 * Phase 2 runs one Trinity per Chrysalis component through ParaFly
 * (Trinity:2510-2529,3652; write_partitioned_trinity_cmds.pl:79-93), each
 * counting its reads (inchworm --reads, Trinity:2669-2672) and assigning
 * them to its contigs (ReadsToTranscripts -p 50, Trinity:1044-1046,2241).
 * A batch holds many components: reads[read_comp_off[c], read_comp_off[c+1])
 * and contigs[contig_comp_off[c], contig_comp_off[c+1]) belong to component
 * c.  One run counts every component's reads in its own table, scores each
 * read against its component (tk_read_cov_stats semantics) and votes it
 * against its component's contigs (tk_r2t_vote semantics; comp = index of
 * the contig within the component).  The reads/contigs objects must outlive
 * the batch.
 */
typedef struct tk_batch tk_batch;
typedef struct {
    double window_load;   /* table slots >= component windows / window_load (0: 0.7) */
    int32_t stats_waves;  /* wavefronts per block of the per-read kernels (0: 4)      */
    int32_t use_graph;    /* replay the run as one captured hipGraph                  */
} tk_batch_opts;

int tk_batch_create(tk_ctx* ctx, int k, int canonical, const tk_reads* reads,
                    const uint64_t* read_comp_off, const tk_reads* contigs,
                    const uint64_t* contig_comp_off, uint64_t n_comp,
                    const tk_batch_opts* opts, tk_batch** out);
/* d_stats / d_votes: device arrays, one entry per read (NULL skips that part). */
int tk_batch_run_d(tk_ctx* ctx, tk_batch* b, const tk_stats_opts* opts,
                   tk_read_stat* d_stats, tk_r2t_vote* d_votes);
int tk_batch_run(tk_ctx* ctx, tk_batch* b, const tk_stats_opts* opts,
                 tk_read_stat* stats, tk_r2t_vote* votes);
/* Time the window_load x stats_waves candidates on this batch, keep the fastest. */
int tk_batch_autotune(tk_ctx* ctx, tk_batch* b, const tk_stats_opts* opts,
                      tk_read_stat* d_stats, tk_r2t_vote* d_votes, double* best_ms);
int tk_batch_info(const tk_batch* b, uint64_t* device_bytes, double* window_load,
                  int* stats_waves);
int tk_batch_free(tk_batch* b);

/* ---------------- resident normalization service ---------------- */
/*
 * insilico_read_normalization.pl runs one sample per process
 * (Trinity:3421-3446): jellyfish count, fastaToKmerCoverageStats on both
 * ends, pair merge + selection (util/insilico_read_normalization.pl:617-645,
 * 837-845,951-976).  The service runs that job for many samples at once on
 * one GPU: `slots` workers with their own streams overlap their kernels; a
 * job is admitted when its device footprint fits hbm_budget_bytes (packed
 * reads first, then the count table sized from its distinct-k-mer estimate);
 * the table stays on the device from count to selection.
 */
typedef struct tk_server tk_server;
typedef struct {
    int32_t device;
    int32_t slots;             /* concurrent jobs (1..16)                      */
    int32_t k;                 /* 25                                            */
    int32_t min_count;         /* dump -L: 2 for normalization                  */
    int32_t max_cov;           /* nbkc_normalize.pl --max_cov                   */
    int32_t max_cv;            /* nbkc_normalize.pl --max_CV                    */
    double max_load;           /* count-table load factor                       */
    uint64_t hbm_budget_bytes; /* admission budget (0: 64 GiB)                 */
} tk_server_opts;

typedef struct {
    int32_t status;            /* TK_OK or the failing call's status            */
    uint64_t n_selected;
    double distinct_estimate;
    double wait_s;             /* submit -> admitted                            */
    double run_s;              /* admitted -> finished                          */
    double latency_s;          /* submit -> finished                            */
    char message[256];
} tk_job_result;

int tk_server_create(const tk_server_opts* opts, tk_server** out);
/* seq/offsets: both.fa layout, left read i and right read n_pairs+i form pair
 * i (sorted accession order).  selected (n_pairs bytes, may be NULL) gets the
 * pairs nbkc_normalize.pl keeps.  Buffers must stay valid until tk_server_wait. */
int tk_server_submit(tk_server* s, const char* seq, const uint64_t* offsets, uint64_t n_pairs,
                     uint8_t* selected, uint64_t* job_id);
/*
 * The job of insilico_read_normalization.pl in general form
 * (util/insilico_read_normalization.pl:270-464): every read of seq/offsets
 * (the both.fa or single.fa order) is counted (canonical unless
 * strand-specific: :617-622) and gets its coverage statistics
 * (fastaToKmerCoverageStats: :837-845); then one or two selections.  Unit u
 * of a selection is the pair (left[u], right[u]) of read indices, in the
 * merged-row order of nbkc_merge_left_right_stats.pl --sorted
 * (--pairs_together: :951-987), or the single read left[u] when right is
 * NULL (nbkc_normalize.pl on one stats file: single reads, and each side of
 * paired reads normalized separately: :917-944).  Parameters are per job, so
 * one server takes jobs of any settings; its opts give min_count only.  The
 * job struct and every buffer it names stay valid until tk_server_wait;
 * n_selected is set when the job ends.
 */
typedef struct {
    const uint64_t* left;
    const uint64_t* right;     /* NULL: single-read units                      */
    uint64_t n_units;
    uint8_t* selected;         /* n_units bytes: 1 = kept                      */
    uint64_t n_selected;       /* out                                          */
} tk_server_selection;

typedef struct {
    const char* seq;
    const uint64_t* offsets;   /* n_reads + 1                                  */
    uint64_t n_reads;
    int32_t k;                 /* --KMER_SIZE (<= 31)                          */
    int32_t canonical;         /* 0 with --SS_lib_type                         */
    int32_t max_cov;           /* nbkc_normalize.pl --max_cov                  */
    int32_t min_cov;           /* --min_cov                                    */
    int32_t max_cv;            /* --max_CV                                     */
    int32_t n_selections;      /* 1 or 2                                       */
    tk_server_selection sel[2];
} tk_server_job;
int tk_server_submit_job(tk_server* s, tk_server_job* job, uint64_t* job_id);
int tk_server_wait(tk_server* s, uint64_t job_id, tk_job_result* result);
int tk_server_free(tk_server* s);

#ifdef __cplusplus
}
#endif
#endif /* TRINITY_GPU_H */
