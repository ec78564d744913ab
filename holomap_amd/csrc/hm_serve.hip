// Fit-request serving loop over device-resident jobs (host code).
//
// Requests (one dataset each: n x d rows, k, metric) wait in a FIFO queue.
// hm_server_progress, called by the client's loop, retires finished jobs and
// admits queued requests in arrival order into new jobs while job slots and
// the context's HBM budget allow: a request that arrives while others run
// joins the next job as soon as a slot and memory free up (continuous
// batching), and the jobs run concurrently, each on its own stream. Results
// stay in HBM until the client fetches them (hm_server_fetch); a job's
// memory returns to the pool when its last request is fetched.
//
// The graph of a request does not depend on the job it lands in (every
// output is per-row reference-order arithmetic), so a request's result is
// the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/holomap_gpu.h"

namespace {

// A request's rows are copied at submission into pinned host memory, so
// that their upload is an asynchronous DMA (a copy from pageable memory
// blocks the client's thread for its whole length). Blocks come in
// power-of-two size classes (>= 1 MiB) and are reused across requests, so
// pinning is paid once per block; past kPinnedCap bytes pinned in all, or if
// pinning fails, requests fall back to pageable memory.
constexpr uint64_t kPinnedCap = 8ull << 30;
struct PinnedPool {
  std::unordered_map<int, std::vector<void*>> free;  // size class -> blocks
  uint64_t pinned = 0;
  static int size_class(uint64_t bytes) {
    int c = 20;
    while ((1ull << c) < bytes) ++c;
    return c;
  }
  // A block of >= bytes; *cls = its class, or -1 for pageable memory.
  void* get(uint64_t bytes, int* cls) {
    const int c = size_class(bytes);
    auto& v = free[c];
    if (!v.empty()) {
      void* p = v.back();
      v.pop_back();
      *cls = c;
      return p;
    }
    void* p = nullptr;
    if (pinned + (1ull << c) <= kPinnedCap &&
        hipHostMalloc(&p, (size_t)1 << c, hipHostMallocDefault) == hipSuccess) {
      pinned += 1ull << c;
      *cls = c;
      return p;
    }
    (void)hipGetLastError();  // a failed pinning is not an error: pageable memory instead
    *cls = -1;
    return std::malloc(bytes ? bytes : 1);
  }
  void put(void* p, int cls) {
    if (!p) return;
    if (cls >= 0) free[cls].push_back(p);
    else std::free(p);
  }
  ~PinnedPool() {
    for (auto& kv : free)
      for (void* p : kv.second) (void)hipHostFree(p);
  }
};

struct Request {
  uint64_t id = 0;
  uint64_t n = 0;
  uint32_t d = 0, k = 0;
  int32_t metric = 0;
  float* data = nullptr;    // host copy (pinned when possible), kept until the job has run
  int data_cls = -1;        // its pool size class (-1: pageable)
  int state = 0;            // 0 queued, 1 running, 2 done, < 0 error
  int job = -1;             // index into Server::jobs
  uint32_t ds = 0;          // dataset slot within the job
};

struct Job {
  hm_job* job = nullptr;
  std::vector<uint64_t> reqs;
  uint32_t unfetched = 0;
  bool done = false;
};

}  // namespace

struct hm_server {
  PinnedPool pool;  // destroyed last: the requests' blocks are returned before
  hm_ctx* ctx = nullptr;
  hm_server_config cfg;
  uint64_t next_id = 1;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  std::vector<Job> jobs;  // slots; a slot with job == nullptr is free
};

extern "C" {

void hm_server_default_config(hm_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->max_jobs = 4;
  c->max_batch = 32;
  c->max_batch_rows = 1u << 20;
  hm_default_options(&c->opt);
}

hm_status hm_server_create(hm_ctx* ctx, const hm_server_config* cfg, hm_server** out) {
  if (!ctx || !out) return HM_EINVAL;
  hm_server* s = new hm_server();
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg; else hm_server_default_config(&s->cfg);
  if (s->cfg.max_jobs == 0 || s->cfg.max_batch == 0 || s->cfg.max_batch_rows == 0) {
    delete s;
    return HM_EINVAL;
  }
  s->cfg.opt.use_graph = 0;  // every job runs once
  s->cfg.opt.timing = 0;
  *out = s;
  return HM_OK;
}

static void release_data(hm_server* s, Request& r) {
  s->pool.put(r.data, r.data_cls);
  r.data = nullptr;
}

void hm_server_destroy(hm_server* s) {
  if (!s) return;
  for (auto& j : s->jobs)
    if (j.job) hm_job_destroy(j.job);  // synchronises with its stream: uploads are over
  for (auto& kv : s->reqs) release_data(s, kv.second);
  delete s;
}

// dst = src (len floats), and the first index of a non-finite value (len if
// none), in one pass; large requests are split over host threads (the
// client's thread would otherwise spend tens of milliseconds here on a
// 50,000 x 768 request, holding up every other request).
static uint64_t copy_checked(float* dst, const float* src, uint64_t len) {
  auto part = [&](uint64_t b, uint64_t e) -> uint64_t {
    for (uint64_t c = b; c < e; c += 4096) {
      const uint64_t ce = std::min<uint64_t>(e, c + 4096);
      uint32_t any = 0;
      for (uint64_t i = c; i < ce; ++i) {
        uint32_t u;
        std::memcpy(&u, src + i, 4);
        std::memcpy(dst + i, &u, 4);
        any |= (u & 0x7f800000u) == 0x7f800000u ? 1u : 0u;
      }
      if (any)
        for (uint64_t i = c; i < ce; ++i)
          if (!std::isfinite(src[i])) return i;
    }
    return e;
  };
  constexpr uint64_t kPerThread = 4ull << 20;  // floats
  const unsigned nt = (unsigned)std::min<uint64_t>(8, (len + kPerThread - 1) / kPerThread);
  if (nt <= 1) {
    const uint64_t f = part(0, len);
    return f;
  }
  std::vector<uint64_t> first(nt, len);
  std::vector<std::thread> th;
  const uint64_t chunk = (len + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const uint64_t b = std::min<uint64_t>(len, t * chunk), e = std::min<uint64_t>(len, b + chunk);
    th.emplace_back([&, t, b, e] {
      const uint64_t f = part(b, e);
      first[t] = f < e ? f : len;
    });
  }
  for (auto& x : th) x.join();
  return *std::min_element(first.begin(), first.end());
}

hm_status hm_server_submit(hm_server* s, const float* data, uint64_t n, uint32_t d, uint32_t k,
                           int32_t metric, uint64_t* id) {
  if (!s || !data || !id) return HM_EINVAL;
  const hm_status v = hm_validate(nullptr, n * d, d, k, nullptr);  // shape and k
  if (v != HM_OK) return v;
  if (metric != HM_EUCLIDEAN && metric != HM_COSINE) return HM_EINVAL;
  // the limits hm_job_create applies: a request it would refuse never
  // enters the queue (where it would fail the requests batched with it)
  if (k > HM_MAX_NEIGHBORS || n >= (1ull << 31)) return HM_EUNSUPPORTED;
  Request r;
  r.n = n;
  r.d = d;
  r.k = k;
  r.metric = metric;
  r.data = (float*)s->pool.get(n * d * sizeof(float), &r.data_cls);
  if (copy_checked(r.data, data, n * d) < n * d) {  // src/api.rs:142-144
    s->pool.put(r.data, r.data_cls);
    return HM_ENONFINITE;
  }
  r.id = s->next_id++;
  *id = r.id;
  s->queue.push_back(r.id);
  s->reqs.emplace(r.id, std::move(r));
  return HM_OK;
}

// A slot holds a job until its last request is fetched (running jobs are
// limited separately, by cfg.max_jobs).
static int free_slot(hm_server* s) {
  for (size_t i = 0; i < s->jobs.size(); ++i)
    if (!s->jobs[i].job) return (int)i;
  s->jobs.emplace_back();
  return (int)s->jobs.size() - 1;
}

static uint32_t running_jobs(const hm_server* s) {
  uint32_t r = 0;
  for (auto& j : s->jobs)
    if (j.job && !j.done) ++r;
  return r;
}

hm_status hm_server_progress(hm_server* s, uint32_t* completed) {
  if (!s) return HM_EINVAL;
  uint32_t fin = 0;
  // retire finished jobs
  for (auto& j : s->jobs) {
    if (!j.job || j.done) continue;
    const int q = hm_job_done(j.job);
    if (q == 0) continue;
    j.done = true;
    for (uint64_t id : j.reqs) {
      Request& r = s->reqs[id];
      r.state = q == 1 ? 2 : HM_EDEVICE;
      release_data(s, r);
      ++fin;
    }
  }
  // admit queued requests, oldest first
  const uint64_t budget = hm_ctx_budget(s->ctx);
  auto fail_head = [&](hm_status st) {
    Request& r = s->reqs[s->queue.front()];
    r.state = st;
    release_data(s, r);
    s->queue.pop_front();
    ++fin;
  };
  while (!s->queue.empty() && running_jobs(s) < s->cfg.max_jobs) {
    {  // a request larger than the whole budget can never run
      const Request& h = s->reqs[s->queue.front()];
      const uint64_t need1 = hm_job_estimate_bytes(1, &h.n, &h.d, &h.k, &h.metric, &s->cfg.opt);
      if (budget && need1 > budget) {
        fail_head(HM_EBUDGET);
        continue;
      }
    }
    // the longest queue prefix that fits the slot limits and the free budget
    std::vector<uint64_t> ids, n;
    std::vector<uint32_t> d, k;
    std::vector<int32_t> m;
    uint64_t rows = 0;
    for (size_t qi = 0; qi < s->queue.size() && ids.size() < s->cfg.max_batch; ++qi) {
      const Request& r = s->reqs[s->queue[qi]];
      if (!ids.empty() && rows + r.n > s->cfg.max_batch_rows) break;
      n.push_back(r.n);
      d.push_back(r.d);
      k.push_back(r.k);
      m.push_back(r.metric);
      const uint64_t need = hm_job_estimate_bytes((uint32_t)n.size(), n.data(), d.data(),
                                                  k.data(), m.data(), &s->cfg.opt);
      if (budget && hm_ctx_in_use(s->ctx) + need > budget) {
        n.pop_back();
        d.pop_back();
        k.pop_back();
        m.pop_back();
        break;
      }
      ids.push_back(r.id);
      rows += r.n;
    }
    if (ids.empty()) break;  // the head waits for memory to free up
    bool held = false;  // memory that will come back: live jobs
    for (auto& j : s->jobs) held = held || j.job != nullptr;
    hm_job* J = nullptr;
    hm_status st = HM_OK;
    size_t take = ids.size();
    for (;;) {
      st = hm_job_create(s->ctx, (uint32_t)take, n.data(), d.data(), k.data(), m.data(),
                         &s->cfg.opt, HM_STAGE_ALL, &J);
      if (st == HM_OK) break;
      const bool mem = st == HM_EBUDGET || st == HM_ENOMEM;
      if (mem && held) break;  // retried when memory frees up
      if (take == 1) break;
      // a shorter prefix (memory), or the head alone (any other refusal, so
      // that one request's failure does not fail the others)
      take = mem ? take / 2 : 1;
    }
    if (st != HM_OK) {
      if ((st == HM_EBUDGET || st == HM_ENOMEM) && held) break;
      fail_head(st);  // the head alone cannot run
      continue;
    }
    ids.resize(take);
    for (size_t i = 0; i < ids.size(); ++i) s->queue.pop_front();
    if (st == HM_OK) {
      for (size_t i = 0; i < ids.size() && st == HM_OK; ++i)
        st = hm_job_upload(J, (uint32_t)i, s->reqs[ids[i]].data);
      if (st == HM_OK) st = hm_job_run(J);
      if (st != HM_OK) hm_job_destroy(J);
    }
    if (st != HM_OK) {  // (the job, destroyed, has finished its uploads)
      for (uint64_t id : ids) {
        s->reqs[id].state = st;
        release_data(s, s->reqs[id]);
        ++fin;
      }
      continue;
    }
    const int slot = free_slot(s);
    Job& j = s->jobs[slot];
    j.job = J;
    j.reqs = ids;
    j.unfetched = (uint32_t)ids.size();
    j.done = false;
    for (size_t i = 0; i < ids.size(); ++i) {
      Request& r = s->reqs[ids[i]];
      r.state = 1;
      r.job = slot;
      r.ds = (uint32_t)i;
    }
  }
  if (completed) *completed = fin;
  return HM_OK;
}

int hm_server_state(hm_server* s, uint64_t id) {
  if (!s) return HM_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return HM_EINVAL;
  return it->second.state;
}

uint32_t hm_server_pending(const hm_server* s) {
  if (!s) return 0;
  uint32_t p = (uint32_t)s->queue.size();
  for (auto& j : s->jobs)
    if (j.job && !j.done) p += (uint32_t)j.reqs.size();
  return p;
}

hm_status hm_server_result_info(hm_server* s, uint64_t id, uint64_t* n, uint32_t* k,
                                uint64_t* nnz) {
  if (!s) return HM_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second.state != 2) return HM_EINVAL;
  const Request& r = it->second;
  if (n) *n = r.n;
  if (k) *k = r.k;
  if (nnz) {
    hm_graph_out o;
    std::memset(&o, 0, sizeof(o));
    const hm_status st = hm_job_download(s->jobs[r.job].job, r.ds, &o);  // sizes only
    if (st != HM_OK) return st;
    *nnz = o.nnz;
  }
  return HM_OK;
}

hm_status hm_server_fetch(hm_server* s, uint64_t id, hm_graph_out* out) {
  if (!s) return HM_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return HM_EINVAL;
  Request& r = it->second;
  if (r.state < 0) {
    const hm_status st = r.state;
    release_data(s, r);
    s->reqs.erase(it);
    return st;
  }
  if (r.state != 2) return HM_EINVAL;
  hm_status st = HM_OK;
  if (out) st = hm_job_download(s->jobs[r.job].job, r.ds, out);
  if (st != HM_OK) return st;
  Job& j = s->jobs[r.job];
  s->reqs.erase(it);
  if (--j.unfetched == 0) {
    hm_job_destroy(j.job);
    j = Job();
  }
  return HM_OK;
}

}  // extern "C"
