// tk_server.cpp -- resident normalization service on one GPU.
//
// The reference normalizes one sample per insilico_read_normalization.pl
// process (Trinity:3421-3446; util/insilico_read_normalization.pl:270-464):
// jellyfish count (:617-633), fastaToKmerCoverageStats on left and right
// (:837-845), nbkc_merge_left_right_stats.pl + nbkc_normalize.pl (:951-976).
// The server keeps that job's data on the device from count to selection
// and runs many samples at once:
//   * `slots` worker threads, each with its own context (HIP stream), so the
//     kernels of different samples overlap on the GPU;
//   * admission by device footprint in two stages: the packed reads (known
//     from the input size) before upload, the count table once its HyperLogLog
//     estimate is known; a job waits until its bytes fit the budget;
//   * results: selected pair flags (host), distinct estimate, per-stage times;
//   * each worker keeps the device blocks its jobs free (up to budget/(2*slots)
//     bytes) and reuses them, so steady-state jobs make no hipMalloc/hipFree
//     calls, which would otherwise wait for every worker's kernels.
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/trinity_gpu.h"

namespace tk {
// Per-thread device block cache (tk_lib.hip): a worker frees and reallocates
// the same few block sizes job after job without hipFree/hipMalloc.
void cache_begin(size_t limit_bytes);
void cache_end();
// Device bytes this thread's region-resident counts may use for their word
// streams (tk_lib.hip); a count whose streams do not fit uses table atomics.
void resident_limit(size_t bytes);
// d_out[i] = d_stats[d_idx[i]] on the context's stream (tk_lib.hip).
int gather_stats(tk_ctx* c, const tk_read_stat* d_stats, uint64_t n_reads, const uint64_t* d_idx, uint64_t n,
                 tk_read_stat* d_out);
}  // namespace tk

namespace {

using Clock = std::chrono::steady_clock;

double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

struct Job {
  uint64_t id = 0;
  const char* seq = nullptr;
  const uint64_t* offsets = nullptr;
  uint64_t n_reads = 0;
  int k = 25, canonical = 1, max_cov = 0, min_cov = 1, max_cv = 0;
  // tk_server_submit: pair i = reads i and n_pairs + i (both.fa layout)
  uint64_t n_pairs = 0;
  uint8_t* selected = nullptr;      // host, n_pairs
  tk_server_job* general = nullptr; // tk_server_submit_job: the caller's units
  Clock::time_point t_submit;
  tk_job_result res{};
  bool done = false;
};

}  // namespace

struct tk_server {
  tk_server_opts opts{};
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_queue, cv_mem, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;
  uint64_t next_id = 1;
  uint64_t budget = 0, in_use = 0;
  int holders = 0;        // jobs holding a stage-1 reservation
  int stage2_waiting = 0; // of those, jobs waiting for their table reservation
  bool stop = false;
  int init_error = 0;
  std::string init_msg;
};

namespace {

// Wait until `bytes` fit the budget, then reserve them.  Stage 1 (reads)
// admits a job; stage 2 (its table) is reserved while the job holds stage 1.
// A job larger than the whole budget runs alone; when every job holding
// memory is waiting for its stage-2 bytes, one of them goes ahead over
// budget, so the holders can never wait on each other forever.
void reserve(tk_server* s, uint64_t bytes, bool stage2) {
  std::unique_lock<std::mutex> lk(s->mu);
  if (stage2) s->stage2_waiting++;
  s->cv_mem.wait(lk, [&] {
    return s->stop || s->in_use == 0 || s->in_use + bytes <= s->budget ||
           (stage2 && s->stage2_waiting == s->holders);
  });
  if (stage2) s->stage2_waiting--;
  else s->holders++;
  s->in_use += bytes;
}

void release(tk_server* s, uint64_t bytes, bool job_done) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->in_use -= bytes;
    if (job_done) s->holders--;
  }
  s->cv_mem.notify_all();
}

// One selection on device stats: units gathered into merged-row order (when
// given as indices), then nbkc_normalize.pl's draws.
int run_selection(tk_ctx* ctx, const Job& j, const tk_read_stat* d_stats, uint32_t max_nwin,
                  const uint64_t* left, const uint64_t* right, uint64_t n, uint8_t* selected, uint64_t* nsel) {
  *nsel = 0;
  if (!n) return TK_OK;
  int rc = TK_OK;
  void *d_idx = nullptr, *d_rows = nullptr, *d_sel = nullptr;
  do {
    const tk_read_stat* dl = d_stats;
    const tk_read_stat* dr = d_stats + j.n_pairs;
    if (left) {
      if ((rc = tk_buffer_alloc(ctx, n * 8, &d_idx))) break;
      if ((rc = tk_buffer_alloc(ctx, (right ? 2 : 1) * n * sizeof(tk_read_stat), &d_rows))) break;
      tk_read_stat* rows = (tk_read_stat*)d_rows;
      if ((rc = tk_memcpy_h2d(ctx, d_idx, left, n * 8))) break;
      if ((rc = tk::gather_stats(ctx, d_stats, j.n_reads, (const uint64_t*)d_idx, n, rows))) break;
      dl = dr = rows;
      if (right) {
        // the copy waits for the gather above on the context's stream
        if ((rc = tk_memcpy_h2d(ctx, d_idx, right, n * 8))) break;
        if ((rc = tk::gather_stats(ctx, d_stats, j.n_reads, (const uint64_t*)d_idx, n, rows + n))) break;
        dr = rows + n;
      }
    }
    if ((rc = tk_buffer_alloc(ctx, n + 64, &d_sel))) break;
    if ((rc = tk_select_pairs_d(ctx, dl, dr, n, j.max_cov, j.min_cov, j.max_cv, max_nwin, (uint8_t*)d_sel, nsel)))
      break;
    if (selected && (rc = tk_memcpy_d2h(ctx, selected, d_sel, n))) break;
  } while (0);
  tk_buffer_free(ctx, d_sel);
  tk_buffer_free(ctx, d_rows);
  tk_buffer_free(ctx, d_idx);
  return rc;
}

int run_job(tk_server* s, tk_ctx* ctx, Job& j) {
  const tk_server_opts& o = s->opts;
  const uint64_t n_reads = j.n_reads;
  const uint64_t n_bases = j.offsets[n_reads];
  uint64_t unit_bytes = j.n_pairs;
  if (j.general)
    for (int q = 0; q < j.general->n_selections; q++)
      unit_bytes += j.general->sel[q].n_units * (8 + 2 * sizeof(tk_read_stat) + 1) + 64;
  // stage 1: reads (ASCII staging + packed codes + stats output) and the
  // selections' rows
  const uint64_t read_bytes = n_bases + n_bases * 3 / 8 + n_reads * (16 + sizeof(tk_read_stat)) + unit_bytes;
  const auto t0 = Clock::now();
  reserve(s, read_bytes, false);
  const auto t1 = Clock::now();
  j.res.wait_s = secs(j.t_submit, t1);
  int rc = TK_OK;
  tk_reads* reads = nullptr;
  tk_table* table = nullptr;
  void* d_stats = nullptr;
  uint64_t table_bytes = 0;
  do {
    if ((rc = tk_reads_from_host(ctx, j.seq, j.offsets, n_reads, &reads))) break;
    double est = 0;
    if ((rc = tk_estimate_distinct(ctx, reads, j.k, j.canonical, &est))) break;
    j.res.distinct_estimate = est;
    const uint64_t expected = (uint64_t)(est * 1.05) + 4096;
    // stage 2: the table (8-byte slots at load <= max_load; the layout
    // rounds up by at most 2x) plus its side table
    table_bytes = (uint64_t)(2.0 * (double)expected / o.max_load * 8.0) + (1ull << 20);
    reserve(s, table_bytes, true);
    if ((rc = tk_table_create(ctx, j.k, j.canonical, expected, o.max_load, &table))) break;
    if ((rc = tk_count(ctx, table, reads))) break;
    if ((rc = tk_buffer_alloc(ctx, n_reads * sizeof(tk_read_stat) + 64, &d_stats))) break;
    tk_stats_opts so{(uint32_t)o.min_count, 1, TK_NWIN_AS_MISSING, TK_MEDIAN_UPPER};
    if ((rc = tk_read_cov_stats_d(ctx, table, reads, &so, (tk_read_stat*)d_stats))) break;
    // the table is no longer needed: its blocks go back before the selection
    tk_table_free(table);
    table = nullptr;
    uint32_t max_len = 0;
    tk_reads_info(reads, nullptr, nullptr, &max_len);
    const uint32_t max_nwin = max_len >= (uint32_t)j.k ? max_len - j.k + 1 : 0;
    const tk_read_stat* ds = (const tk_read_stat*)d_stats;
    uint64_t nsel = 0;
    if (!j.general) {
      if ((rc = run_selection(ctx, j, ds, max_nwin, nullptr, nullptr, j.n_pairs, j.selected, &nsel))) break;
      j.res.n_selected = nsel;
    } else {
      for (int q = 0; q < j.general->n_selections && !rc; q++) {
        tk_server_selection& sel = j.general->sel[q];
        rc = run_selection(ctx, j, ds, max_nwin, sel.left, sel.right, sel.n_units, sel.selected, &nsel);
        sel.n_selected = nsel;
        j.res.n_selected += nsel;
      }
    }
  } while (0);
  tk_buffer_free(ctx, d_stats);
  tk_table_free(table);
  tk_reads_free(reads);
  if (table_bytes) release(s, table_bytes, false);
  release(s, read_bytes, true);
  const auto t2 = Clock::now();
  j.res.run_s = secs(t1, t2);
  j.res.latency_s = secs(j.t_submit, t2);
  (void)t0;
  return rc;
}

void worker(tk_server* s, int idx) {
  // Every object of a job is created and freed on this thread and its
  // context's stream, which is what the block cache requires.
  tk::cache_begin(s->budget / (2 * (uint64_t)s->opts.slots));
  // jobs run side by side under the HBM budget: count with table atomics
  // rather than holding region-resident word streams per worker
  tk::resident_limit(1);
  tk_ctx* ctx = nullptr;
  int rc = tk_init(s->opts.device, &ctx);
  if (rc) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->init_error = rc;
    s->init_msg = tk_last_error();
  }
  (void)idx;
  while (true) {
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_queue.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) break;  // stopping
      j = s->queue.front();
      s->queue.pop_front();
    }
    if (!ctx) {
      j->res.status = rc;
    } else {
      j->res.status = run_job(s, ctx, *j);
      if (j->res.status) {
        const char* m = tk_last_error();
        strncpy(j->res.message, m, sizeof(j->res.message) - 1);
      }
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      j->done = true;
    }
    s->cv_done.notify_all();
  }
  if (ctx) tk_ctx_free(ctx);
  tk::cache_end();
}

int enqueue(tk_server* s, std::shared_ptr<Job> j, uint64_t* job_id) {
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return TK_ERR_STATE;
    j->id = s->next_id++;
    s->jobs[j->id] = j;
    s->queue.push_back(j);
  }
  s->cv_queue.notify_one();
  *job_id = j->id;
  return TK_OK;
}

}  // namespace

extern "C" {

int tk_server_create(const tk_server_opts* o, tk_server** out) {
  if (!o || !out) return -1;
  *out = nullptr;
  if (o->slots < 1 || o->slots > 16 || o->k < 1 || o->k > 31 || !(o->max_load > 0 && o->max_load <= 0.9) ||
      o->max_cov <= 0)
    return TK_ERR_ARG;
  tk_server* s = new tk_server();
  s->opts = *o;
  s->budget = o->hbm_budget_bytes ? o->hbm_budget_bytes : (64ull << 30);
  for (int i = 0; i < o->slots; i++) s->workers.emplace_back(worker, s, i);
  *out = s;
  return TK_OK;
}

int tk_server_submit(tk_server* s, const char* seq, const uint64_t* offsets, uint64_t n_pairs,
                     uint8_t* selected, uint64_t* job_id) {
  if (!s || !offsets || !job_id || (n_pairs && !seq)) return TK_ERR_ARG;
  auto j = std::make_shared<Job>();
  j->seq = seq;
  j->offsets = offsets;
  j->n_reads = 2 * n_pairs;
  j->n_pairs = n_pairs;
  j->selected = selected;
  j->k = s->opts.k;
  j->canonical = 1;
  j->max_cov = s->opts.max_cov;
  j->min_cov = 1;
  j->max_cv = s->opts.max_cv;
  return enqueue(s, j, job_id);
}

int tk_server_submit_job(tk_server* s, tk_server_job* g, uint64_t* job_id) {
  if (!s || !g || !job_id || !g->offsets || (g->n_reads && !g->seq)) return TK_ERR_ARG;
  if (g->k < 1 || g->k > 31 || g->max_cov <= 0 || g->n_selections < 1 || g->n_selections > 2) return TK_ERR_ARG;
  for (int q = 0; q < g->n_selections; q++) {
    const tk_server_selection& sel = g->sel[q];
    if (sel.n_units && !sel.left) return TK_ERR_ARG;
    for (uint64_t u = 0; u < sel.n_units; u++)
      if (sel.left[u] >= g->n_reads || (sel.right && sel.right[u] >= g->n_reads)) return TK_ERR_ARG;
  }
  auto j = std::make_shared<Job>();
  j->seq = g->seq;
  j->offsets = g->offsets;
  j->n_reads = g->n_reads;
  j->k = g->k;
  j->canonical = g->canonical ? 1 : 0;
  j->max_cov = g->max_cov;
  j->min_cov = g->min_cov;
  j->max_cv = g->max_cv;
  j->general = g;
  return enqueue(s, j, job_id);
}

int tk_server_wait(tk_server* s, uint64_t job_id, tk_job_result* res) {
  if (!s || !res) return TK_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return TK_ERR_ARG;
  std::shared_ptr<Job> j = it->second;
  s->cv_done.wait(lk, [&] { return j->done; });
  *res = j->res;
  s->jobs.erase(job_id);
  return TK_OK;
}

int tk_server_free(tk_server* s) {
  if (!s) return TK_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_queue.notify_all();
  s->cv_mem.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
  return TK_OK;
}

}  // extern "C"
