// pk_server.cpp -- batched and served forms of the counting / cutoff /
// macro-node path (pk_server_* in include/pakhip.h).
//
// n_slots contexts on one device; slot i is driven by host thread i on the
// context's own HIP stream, so jobs in different slots overlap on the GPU
// (their host synchronisations, small launches and tails fill each other's
// gaps).  Jobs are admitted FIFO by estimated HBM footprint.  A finished
// job's table, kept table and nodes are copied into one device allocation
// owned by the job (stream-ordered allocator when the device supports memory
// pools, so steady-state serving does not synchronise the device) and stay
// resident until pk_server_release.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pakhip.h"

namespace {

struct Job {
  pk_job spec{};
  uint64_t ticket = 0;
  int state = 0;  // 0 queued, 1 running, 2 done
  int readers = 0;  // pk_server_download calls copying the results now
  pk_job_result res{};
  std::vector<uint64_t> hist;
  void* mem = nullptr;
  uint64_t mem_bytes = 0;
};

inline uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Working-set estimate of a job: FASTA copy (host input), line index and read
// table, two instance arrays and the per-instance counts, the table.  The
// instance count is bounded by the byte count.
uint64_t estimate_bytes(const pk_job& j) {
  const uint64_t W = j.k <= 32 ? 8 : 16;
  const uint64_t n = j.nbytes;
  uint64_t b = n * (2 * W + 4) + n * (W + 4) + n * 2 + (64ull << 20);
  if (!j.fasta_on_device) b += n;
  return b;
}

}  // namespace

struct pk_server {
  int device = 0;
  int n_slots = 0;
  uint64_t budget = 0;
  bool pools = false;
  std::vector<pk_ctx*> ctx;
  std::vector<std::thread> workers;
  hipStream_t free_stream = nullptr;
  std::mutex mu;
  std::condition_variable cv_q, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, Job*> jobs;
  uint64_t next_ticket = 1;
  uint64_t running_est = 0, resident = 0;
  int running = 0;
  bool stop = false;
  std::string err;
  std::vector<pk_tune_entry> tune;  // ascending max_bytes
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  bool admissible() const {
    if (queue.empty()) return false;
    if (running == 0) return true;
    return running_est + resident + queue.front()->res.est_bytes <= budget;
  }
  // err is written under mu (callers hold it, or use fail_unlocked)
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
  int fail_unlocked(int code, const std::string& m) {
    std::lock_guard<std::mutex> lk(mu);
    return fail(code, m);
  }
  void worker(int slot);
  void run(int slot, Job* j);
  void run_guarded(int slot, Job* j);
};

void pk_server::run(int slot, Job* j) {
  pk_ctx* c = ctx[slot];
  pk_job_result& r = j->res;
  auto set_err = [&](int rc) {
    r.status = rc;
    std::snprintf(r.message, sizeof(r.message), "%s", pk_last_error(c));
  };
  pk_tuning tu{};
  {
    std::lock_guard<std::mutex> lk(mu);
    const int words = j->spec.k <= 32 ? 1 : 2;
    for (const auto& e : tune)
      if (j->spec.nbytes <= e.max_bytes && (e.key_words == 0 || e.key_words == words)) {
        tu = e.tuning;
        break;
      }
  }
  pk_ctx_set_tuning(c, &tu);
  pk_table t{};
  int rc = pk_count(c, j->spec.fasta, j->spec.nbytes, j->spec.k, j->spec.fasta_on_device, &t);
  if (rc) {
    set_err(rc);
    return;
  }
  // the context's timing describes this job only once its count succeeded
  pk_timing tm{};
  pk_last_timing(c, &tm);
  r.n_reads = tm.n_reads;
  r.n_instances = tm.n_instances;
  pk_table kt{};
  pk_mn_table mt{};
  int crc = PK_OK;
  if (j->spec.t >= 0) {
    const int tt = j->spec.t == 0 ? 21 : j->spec.t;
    j->hist.assign((size_t)tt, 0);
    uint64_t nk = 0;
    int mb = 0;
    crc = pk_cutoff(c, j->spec.t, j->hist.data(), &mb, &nk);
    r.min_bucket = mb;
    if (crc == PK_OK) {
      pk_kept_table(c, &kt);
      if (j->spec.want_nodes) crc = pk_macro_nodes(c, &mt);
    }
    if (crc) set_err(crc);
    if (crc && crc != PK_EEMPTY) return;
  }
  // one allocation: table keys | counts | kept keys | counts | nodes
  const uint64_t w = (uint64_t)t.words * 8;
  const uint64_t o1 = 0, o2 = align256(o1 + t.n * w), o3 = align256(o2 + t.n * 4),
                 o4 = align256(o3 + kt.n * w), o5 = align256(o4 + kt.n * 4),
                 end = align256(o5 + mt.n * sizeof(pk_mnode));
  hipStream_t st = (hipStream_t)pk_ctx_stream(c);
  // the job's kernels finish before its result block is allocated
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    r.status = PK_EHIP;
    std::snprintf(r.message, sizeof(r.message), "job failed: %s", hipGetErrorString(e));
    return;
  }
  void* mem = nullptr;
  e = pools ? hipMallocAsync(&mem, end ? end : 256, st) : hipMalloc(&mem, end ? end : 256);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    r.status = PK_ENOMEM;
    std::snprintf(r.message, sizeof(r.message), "result allocation failed: %s", hipGetErrorString(e));
    return;
  }
  char* m = (char*)mem;
  auto cp = [&](uint64_t off, const void* src, uint64_t bytes) {
    if (bytes && e == hipSuccess) e = hipMemcpyAsync(m + off, src, bytes, hipMemcpyDeviceToDevice, st);
  };
  cp(o1, t.keys, t.n * w);
  cp(o2, t.counts, t.n * 4);
  cp(o3, kt.keys, kt.n * w);
  cp(o4, kt.counts, kt.n * 4);
  cp(o5, mt.nodes, mt.n * sizeof(pk_mnode));
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    r.status = PK_EHIP;
    std::snprintf(r.message, sizeof(r.message), "result copy failed: %s", hipGetErrorString(e));
  }
  j->mem = mem;
  j->mem_bytes = end;
  r.table = pk_table{t.n, t.words, t.k, (const uint64_t*)(m + o1), (const int32_t*)(m + o2)};
  r.kept = pk_table{kt.n, t.words, t.k, (const uint64_t*)(m + o3), (const int32_t*)(m + o4)};
  r.nodes = pk_mn_table{mt.n, t.k, (const pk_mnode*)(m + o5)};
}

// run() with host allocation failures (the histogram vector) turned into the
// job's PK_ENOMEM: nothing may throw out of a worker thread.
void pk_server::run_guarded(int slot, Job* j) {
  try {
    run(slot, j);
  } catch (const std::bad_alloc&) {
    j->res.status = PK_ENOMEM;
    std::snprintf(j->res.message, sizeof(j->res.message), "host allocation failed");
  } catch (...) {
    j->res.status = PK_EINVAL;
    std::snprintf(j->res.message, sizeof(j->res.message), "unexpected host exception");
  }
}

void pk_server::worker(int slot) {
  (void)hipSetDevice(device);
  const uint64_t slot_keep = budget / (uint64_t)(n_slots > 0 ? n_slots : 1);
  for (;;) {
    Job* j;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_q.wait(lk, [&] { return stop || admissible(); });
      if (stop) return;
      j = queue.front();
      queue.pop_front();
      j->state = 1;
      running++;
      running_est += j->res.est_bytes;
      j->res.t_start = now();
    }
    run_guarded(slot, j);
    // a slot that grew past its share gives its buffers back
    if (pk_ctx_device_bytes(ctx[slot]) > slot_keep) (void)pk_ctx_release_buffers(ctx[slot]);
    {
      std::lock_guard<std::mutex> lk(mu);
      running--;
      running_est -= j->res.est_bytes;
      resident += j->mem_bytes;
      j->res.t_end = now();
      j->state = 2;
    }
    cv_done.notify_all();
    cv_q.notify_all();
  }
}

extern "C" {

int pk_server_create(pk_server** out, int device, int n_slots, uint64_t hbm_budget_bytes) {
  if (!out || n_slots < 1 || n_slots > 64) return PK_EINVAL;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return PK_EHIP;
  pk_server* s = new pk_server();
  s->device = device;
  s->n_slots = n_slots;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    delete s;
    return PK_EHIP;
  }
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(fr * 0.9);
  int pools = 0;
  if (hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device) == hipSuccess)
    s->pools = pools != 0;
  if (s->pools) {
    // results freed back to the device's default pool stay there for the
    // next job's results instead of being returned to the system at the
    // next synchronisation (which would make every later result
    // allocation map fresh memory)
    hipMemPool_t pool = nullptr;
    uint64_t keep = UINT64_MAX;
    if (hipDeviceGetDefaultMemPool(&pool, device) != hipSuccess ||
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
      (void)hipGetLastError();
      s->pools = false;
    }
  }
  if (hipStreamCreateWithFlags(&s->free_stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return PK_EHIP;
  }
  for (int i = 0; i < n_slots; i++) {
    pk_ctx* c = nullptr;
    int rc = pk_ctx_create(&c, device, 0);
    // slots grow their buffers on their own streams: a slot's growth must
    // not wait for the other slots' work (under sustained load a
    // device-wide wait stalled every slot for seconds)
    if (rc == PK_OK && s->pools) rc = pk_ctx_set_alloc_mode(c, PK_ALLOC_STREAM_ORDERED);
    if (rc) {
      if (c) pk_ctx_destroy(c);
      for (auto* x : s->ctx) pk_ctx_destroy(x);
      (void)hipStreamDestroy(s->free_stream);
      delete s;
      return rc;
    }
    s->ctx.push_back(c);
  }
  for (int i = 0; i < n_slots; i++) s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return PK_OK;
}

void pk_server_destroy(pk_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    for (Job* j : s->queue) {
      j->state = 2;
      j->res.status = PK_ESTATE;
      std::snprintf(j->res.message, sizeof(j->res.message), "server destroyed before the job ran");
    }
    s->queue.clear();
    s->cv_done.wait(lk, [&] { return s->running == 0; });
    s->stop = true;
  }
  s->cv_q.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  (void)hipSetDevice(s->device);
  for (auto& kv : s->jobs) {
    if (kv.second->mem) {
      if (s->pools) (void)hipFreeAsync(kv.second->mem, s->free_stream);
      else (void)hipFree(kv.second->mem);
    }
    delete kv.second;
  }
  (void)hipStreamSynchronize(s->free_stream);
  (void)hipStreamDestroy(s->free_stream);
  for (auto* c : s->ctx) pk_ctx_destroy(c);
  delete s;
}

double pk_server_now(const pk_server* s) { return s ? s->now() : 0.0; }

const char* pk_server_last_error(const pk_server* s) { return s ? s->err.c_str() : "null server"; }

// Largest cutoff bucket count (pk_cutoff's limit).
constexpr int SERVER_T_MAX = 4096;

static int submit_locked(pk_server* s, const pk_job& spec, uint64_t* ticket) {
  if (spec.k < 8 || spec.k > 64) return s->fail(PK_EINVAL, "k must be in [8, 64]");
  if (spec.t > SERVER_T_MAX) return s->fail(PK_EINVAL, "t must be <= 4096");
  if (!spec.fasta && spec.nbytes) return s->fail(PK_EINVAL, "null fasta");
  if (spec.want_nodes && spec.t < 0) return s->fail(PK_EINVAL, "macro nodes need a cutoff (t >= 0)");
  Job* j = new Job();
  j->spec = spec;
  j->ticket = s->next_ticket++;
  j->res.est_bytes = estimate_bytes(spec);
  j->res.t_submit = s->now();
  s->jobs[j->ticket] = j;
  s->queue.push_back(j);
  *ticket = j->ticket;
  return PK_OK;
}

int pk_server_submit(pk_server* s, const pk_job* job, uint64_t* ticket) {
  if (!s || !job || !ticket) return PK_EINVAL;
  int rc;
  try {
    std::lock_guard<std::mutex> lk(s->mu);
    rc = submit_locked(s, *job, ticket);
  } catch (const std::bad_alloc&) {
    return s->fail_unlocked(PK_ENOMEM, "host allocation failed");
  }
  s->cv_q.notify_all();
  return rc;
}

int pk_server_submit_batch(pk_server* s, const pk_job* jobs, int n, uint64_t* tickets) {
  if (!s || n < 0 || (n && (!jobs || !tickets))) return PK_EINVAL;
  int rc = PK_OK;
  try {
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return jobs[a].nbytes > jobs[b].nbytes; });
    std::lock_guard<std::mutex> lk(s->mu);
    for (int i : order) {
      rc = submit_locked(s, jobs[i], &tickets[i]);
      if (rc) break;
    }
  } catch (const std::bad_alloc&) {
    rc = s->fail_unlocked(PK_ENOMEM, "host allocation failed");
  }
  s->cv_q.notify_all();
  return rc;
}

static int finished(pk_server* s, uint64_t ticket, pk_job_result* out, bool block) {
  if (!s || !out) return PK_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) return s->fail(PK_EINVAL, "unknown ticket");
  Job* j = it->second;
  if (j->state != 2) {
    if (!block) return PK_EAGAIN;
    s->cv_done.wait(lk, [&] { return j->state == 2; });
  }
  *out = j->res;
  return PK_OK;
}

int pk_server_wait(pk_server* s, uint64_t ticket, pk_job_result* out) {
  return finished(s, ticket, out, true);
}

int pk_server_poll(pk_server* s, uint64_t ticket, pk_job_result* out) {
  return finished(s, ticket, out, false);
}

int pk_server_download(pk_server* s, uint64_t ticket, uint64_t* keys, int32_t* counts,
                       uint64_t* kept_keys, int32_t* kept_counts, uint64_t* hist_t,
                       pk_mnode* nodes) {
  if (!s) return PK_EINVAL;
  Job* j;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return s->fail(PK_EINVAL, "unknown ticket");
    j = it->second;
    if (j->state != 2) return s->fail(PK_ESTATE, "job has not finished");
    // pk_server_release of this ticket waits until the copy below is done
    j->readers++;
  }
  const pk_job_result& r = j->res;
  const uint64_t w = (uint64_t)r.table.words * 8;
  auto get = [&](void* dst, const void* src, uint64_t bytes) -> hipError_t {
    if (!dst || !bytes || !src) return hipSuccess;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) e = get(keys, r.table.keys, r.table.n * w);
  if (e == hipSuccess) e = get(counts, r.table.counts, r.table.n * 4);
  if (e == hipSuccess) e = get(kept_keys, r.kept.keys, r.kept.n * w);
  if (e == hipSuccess) e = get(kept_counts, r.kept.counts, r.kept.n * 4);
  if (e == hipSuccess) e = get(nodes, r.nodes.nodes, r.nodes.n * sizeof(pk_mnode));
  if (e == hipSuccess && hist_t && !j->hist.empty()) std::memcpy(hist_t, j->hist.data(), j->hist.size() * 8);
  int rc = PK_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j->readers--;
    if (e != hipSuccess) rc = s->fail(PK_EHIP, std::string("download failed: ") + hipGetErrorString(e));
  }
  s->cv_done.notify_all();
  return rc;
}

int pk_server_set_tuning(pk_server* s, const pk_tune_entry* entries, int n) {
  if (!s || n < 0 || (n && !entries)) return PK_EINVAL;
  for (int i = 0; i < n; i++) {
    const pk_tune_entry& e = entries[i];
    const pk_tuning& t = e.tuning;
    if (e.key_words < 0 || e.key_words > 2 || t.l1_chunks < 0 || t.l1_chunks > 64 || t.l1_bits < 0 ||
        t.l1_bits > PK_TUNING_L1_BITS_MAX || !(t.bucket_factor >= 0.f && t.bucket_factor <= 64.f) ||
        t.ms_scatter < 0 || t.ms_scatter > 2 || t.ms_fold < 0 || t.ms_fold > 2)
      return s->fail(PK_EINVAL, "tuning entry " + std::to_string(i) + " out of range");
  }
  try {
    std::vector<pk_tune_entry> t(entries, entries + n);
    std::sort(t.begin(), t.end(),
              [](const pk_tune_entry& a, const pk_tune_entry& b) { return a.max_bytes < b.max_bytes; });
    std::lock_guard<std::mutex> lk(s->mu);
    s->tune = std::move(t);
  } catch (const std::bad_alloc&) {
    return s->fail_unlocked(PK_ENOMEM, "host allocation failed");
  }
  return PK_OK;
}

int pk_server_release(pk_server* s, uint64_t ticket) {
  if (!s) return PK_EINVAL;
  Job* j;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return s->fail(PK_EINVAL, "unknown ticket");
    if (it->second->state != 2) return s->fail(PK_ESTATE, "job has not finished");
    // a download of the same ticket in progress finishes first (the job is
    // looked up again: a concurrent release may have taken it meanwhile)
    s->cv_done.wait(lk, [&] {
      auto jt = s->jobs.find(ticket);
      return jt == s->jobs.end() || jt->second->readers == 0;
    });
    it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return s->fail(PK_EINVAL, "unknown ticket");
    j = it->second;
    s->jobs.erase(it);
    s->resident -= j->mem_bytes;
  }
  if (j->mem) {
    (void)hipSetDevice(s->device);
    if (s->pools) (void)hipFreeAsync(j->mem, s->free_stream);
    else (void)hipFree(j->mem);
  }
  delete j;
  s->cv_q.notify_all();
  return PK_OK;
}

}  // extern "C"
