// Request loop for SCF solves on one device (continuous batching).
//
// A server owns n_slots worker threads, each with its own HIP stream and a
// cache of released device blocks, so back-to-back solves on a worker reuse
// memory instead of returning it (hipFree waits for the whole device).
// Requests queue in arrival order; a free slot admits the oldest request
// whose estimated device footprint fits in the free HBM left after the
// solves already running, so solves of different cell sizes run side by side
// and overlap on the device while later arrivals wait for memory or a slot.
// Each solve keeps ψ, ρ and the mixer history resident in its own context
// for its whole lifetime. A request larger than the whole budget fails with
// PWK_E_OOM instead of waiting forever.
//
// Solves are independent and each is deterministic (fixed-order reductions),
// so a result does not depend on what else is running.
//
// Concurrent solves of the same shape (grid, basis size, projector channels, band counts, filter
// degree and precision) share their filter passes: a solve that reaches a pass joins an open
// group of that shape, or opens one and gathers for a short window; the group's passes then run
// as one launch per kernel on the opening solve's stream, ordered after every member's earlier
// work by events, and every member continues once the batch has completed. Each band and k-point
// block of a batched launch runs the code of the one-solve launch, so the batch changes no result.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "common.hpp"

namespace pwk {

thread_local BlockCache* tl_dev_cache = nullptr;
thread_local BlockCache* tl_host_cache = nullptr;

void* BlockCache::take(size_t& bytes) {
  size_t b = (bytes + 255) & ~(size_t)255;
  if (b == 0) b = 256;
  {
    std::lock_guard<std::mutex> lk(mu);
    // smallest cached block that fits, if it wastes at most half of the request
    auto it = free_.lower_bound(b);
    if (it != free_.end() && it->first <= b + b / 2) {
      void* p = it->second;
      bytes = it->first;
      cached_ -= it->first;
      free_.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = pinned ? hipHostMalloc(&p, b, hipHostMallocDefault) : hipMalloc(&p, b);
  if (e == hipErrorOutOfMemory) {  // give the cached blocks back and try once more
    (void)hipGetLastError();
    trim(0);
    e = pinned ? hipHostMalloc(&p, b, hipHostMallocDefault) : hipMalloc(&p, b);
  }
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? PWK_E_OOM : PWK_E_DEVICE,
                std::string(pinned ? "hipHostMalloc" : "hipMalloc") + ": " + hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(mu);
  held_ += b;
  bytes = b;
  return p;
}

void BlockCache::give(void* p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(mu);
    free_.emplace(bytes, p);
    cached_ += bytes;
    if (cached_ <= limit_) return;
  }
  trim(limit_);
}

void BlockCache::trim(size_t limit) {
  std::lock_guard<std::mutex> lk(mu);
  while (cached_ > limit && !free_.empty()) {
    auto it = std::prev(free_.end());
    if (pinned)
      (void)hipHostFree(it->second);
    else
      (void)hipFree(it->second);
    cached_ -= it->first;
    held_ -= it->first;
    free_.erase(it);
  }
}

struct PassBatcher {
  struct Member {
    const PassReq* r;
    hipEvent_t ready, done;
  };
  struct Group {
    PassShape sh;
    std::vector<Member> m;
    bool launched = false;
    int status = PWK_OK;
    std::string error;
  };
  std::mutex mu;
  std::condition_variable cv;
  int window_us = 100;  // how long a group's first solve gathers others; 0: every pass alone
  int running = 0;      // solves in progress (a group of that many launches at once)
  int64_t launches = 0, passes = 0;  // batched launches and the solve passes they carried
  std::vector<std::shared_ptr<Group>> open;
};

// Keeps the batcher's count of solves in progress.
struct PassBatchScope {
  PassBatcher* b;
  explicit PassBatchScope(PassBatcher* pb) : b(pb) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(b->mu);
    ++b->running;
  }
  ~PassBatchScope() {
    if (!b) return;
    {
      std::lock_guard<std::mutex> lk(b->mu);
      --b->running;
    }
    b->cv.notify_all();  // a gathering group may now be complete
  }
};

void pass_batch_join(Context& c, const PassShape& sh, const PassReq& r) {
  PassBatcher& b = *c.batcher;
  PWK_HIP(hipEventRecord(c.batch_ready, c.stream));  // where this solve's pass may start
  std::unique_lock<std::mutex> lk(b.mu);
  std::shared_ptr<PassBatcher::Group> g;
  const bool gather = b.window_us > 0;
  if (gather)
    for (auto& o : b.open)
      if (o->sh == sh && (int)o->m.size() < MAX_PASS_BATCH) {
        g = o;
        break;
      }
  const bool leader = !g;
  if (leader) {
    g = std::make_shared<PassBatcher::Group>();
    g->sh = sh;
    if (gather) b.open.push_back(g);
  }
  g->m.push_back({&r, c.batch_ready, c.batch_done});
  if (!leader) {
    // the group's first solve launches the batch; this stream continues after it
    b.cv.notify_all();
    b.cv.wait(lk, [&] { return g->launched; });
    if (g->status != PWK_OK) throw Error(g->status, g->error);
    lk.unlock();
    // wait on the host: a stream-side wait would hold the hardware queue this stream shares with
    // other solves' streams until the whole batch has run
    PWK_HIP(hipEventSynchronize(c.batch_done));
    return;
  }
  if (gather) {
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(b.window_us);
    b.cv.wait_until(lk, until, [&] { return (int)g->m.size() >= std::min(MAX_PASS_BATCH, b.running); });
    b.open.erase(std::find(b.open.begin(), b.open.end(), g));
  }
  lk.unlock();  // closed: no solve joins after this, so g->m is fixed
  const std::vector<PassBatcher::Member>& mem = g->m;
  int st = PWK_OK;
  std::string err;
  try {
    std::vector<const PassReq*> rq;
    for (size_t i = 0; i < mem.size(); ++i) {
      rq.push_back(mem[i].r);
      if (i > 0) PWK_HIP(hipStreamWaitEvent(c.stream, mem[i].ready, 0));
    }
    launch_pass_batch(c.stream, sh, rq.data(), (int)rq.size());
    for (size_t i = 1; i < mem.size(); ++i) PWK_HIP(hipEventRecord(mem[i].done, c.stream));
  } catch (const Error& e) {
    st = e.code;
    err = e.what();
  } catch (const std::exception& e) {
    st = PWK_E_DEVICE;
    err = e.what();
  }
  lk.lock();
  g->status = st;
  g->error = err;
  g->launched = true;
  b.launches += 1;
  b.passes += (int64_t)mem.size();
  lk.unlock();
  b.cv.notify_all();
  if (st != PWK_OK) throw Error(st, err);
}

struct ServeJob {
  int64_t id = 0;
  pwk_crystal cr{};
  pwk_scf_params p{};
  size_t bytes = 0;
  int status = PWK_OK;
  std::string error;
  pwk_scf_result res{};
  std::vector<double> eig, rho;
  bool done = false;
  std::chrono::steady_clock::time_point t_submit, t_start, t_end;
};

struct Server {
  int device = 0;
  size_t budget = 0;     // bytes the server may hold for running solves
  size_t cache_cap = 0;  // bytes each worker may keep cached between solves
  static constexpr size_t HOST_CACHE_CAP = (size_t)64 << 20;  // pinned host bytes per worker
  size_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<ServeJob>> queue;
  std::map<int64_t, std::shared_ptr<ServeJob>> jobs;
  int64_t next_id = 1;
  bool stopping = false;
  std::vector<std::thread> workers;
  PassBatcher batcher;

  bool admissible(const ServeJob& j) const { return j.bytes > budget || in_use + j.bytes <= budget; }

  // Runs one solve in a fresh context on the worker's stream. The context's buffers come
  // from the worker's block cache and go back to it when the solve ends.
  void solve(ServeJob& job, hipStream_t stream, hipEvent_t sleep_event, hipEvent_t ready, hipEvent_t done,
             int cu_count) {
    if (job.bytes > budget) {
      job.status = PWK_E_OOM;
      job.error = "solve does not fit in the server's device-memory budget";
      return;
    }
    try {
      Context c;
      c.device = device;
      c.stream = stream;
      c.cu_count = cu_count;
      c.sleep_event = sleep_event;
      c.batcher = &batcher;
      c.batch_ready = ready;
      c.batch_done = done;
      struct StreamGuard {
        Context& c;
        ~StreamGuard() {
          (void)hipStreamSynchronize(c.stream);
          c.kd.clear();
          c.stream = nullptr;  // owned by the worker
        }
      } guard{c};
      // stream capture in one worker would be disturbed by allocation and release in another,
      // so served solves launch the filter passes directly
      c.graphs = false;
      // served solves run the host-synchronised iteration: the hardware queues the workers'
      // streams share are the limit here, not the host, and that loop launches fewer kernels
      // per solve than the device-resident one; its filter passes go through the batcher
      c.resident = false;
      int npw, dims[3], nk, nops;
      scf_tables(job.cr, job.p, &npw, dims, &nk, &nops);
      const size_t ns = job.p.nspin == 2 ? 2 : 1;  // eigenvalues of both spin channels
      job.eig.assign(ns * (size_t)nk * (size_t)std::min(job.p.n_bands, npw), 0.0);
      job.rho.assign((size_t)dims[0] * dims[1] * dims[2], 0.0);
      PassBatchScope running(&batcher);
      scf_run(c, job.cr, job.p, job.res, job.eig.data(), job.rho.data());
    } catch (const Error& e) {
      job.status = e.code;
      job.error = e.what();
    } catch (const std::exception& e) {
      job.status = PWK_E_DEVICE;
      job.error = e.what();
    }
  }

  void worker_main() {
    // per-worker stream and block caches, kept across solves
    hipStream_t stream = nullptr;
    hipEvent_t sleep_event = nullptr, ready = nullptr, done = nullptr;
    int cu_count = 0;
    try {
      PWK_HIP(hipSetDevice(device));
      PWK_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      PWK_HIP(hipEventCreateWithFlags(&sleep_event, hipEventBlockingSync | hipEventDisableTiming));
      PWK_HIP(hipEventCreateWithFlags(&ready, hipEventBlockingSync | hipEventDisableTiming));
      PWK_HIP(hipEventCreateWithFlags(&done, hipEventBlockingSync | hipEventDisableTiming));
      hipDeviceProp_t prop;
      PWK_HIP(hipGetDeviceProperties(&prop, device));
      cu_count = prop.multiProcessorCount;
    } catch (const Error&) {
      stream = nullptr;
    }
    BlockCache dev_cache(false, cache_cap), host_cache(true, HOST_CACHE_CAP);
    tl_dev_cache = &dev_cache;
    tl_host_cache = &host_cache;
    for (;;) {
      std::shared_ptr<ServeJob> job;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          if (stopping && queue.empty()) return true;
          for (auto& j : queue)
            if (admissible(*j)) return true;
          return false;
        });
        if (queue.empty()) break;  // only reached when stopping
        for (auto it = queue.begin(); it != queue.end(); ++it)
          if (admissible(**it)) {
            job = *it;
            queue.erase(it);
            break;
          }
        if (job->bytes <= budget) in_use += job->bytes;
      }
      job->t_start = std::chrono::steady_clock::now();
      if (stream && ready && done) {
        solve(*job, stream, sleep_event, ready, done, cu_count);
        dev_cache.trim(cache_cap);
        host_cache.trim(HOST_CACHE_CAP);
      } else {
        job->status = PWK_E_DEVICE;
        job->error = "worker stream could not be created";
      }
      job->t_end = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lk(mu);
        if (job->bytes <= budget) in_use -= job->bytes;
        job->done = true;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    if (stream) (void)hipStreamSynchronize(stream);
    tl_dev_cache = nullptr;
    tl_host_cache = nullptr;
    dev_cache.trim(0);
    host_cache.trim(0);
    if (sleep_event) (void)hipEventDestroy(sleep_event);
    if (ready) (void)hipEventDestroy(ready);
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Device bytes one solve holds at its peak: the buffers Solve::setup and the iteration allocate
// (scf.hip), for the eigensolver setup() will choose (dense H batch and its scratch, or the
// filter blocks and subspace matrices of the Chebyshev path), plus the context's grid and
// stick workspaces.
size_t scf_footprint(const pwk_crystal& cr, const pwk_scf_params& p) {
  int npw, dims[3], nk, nops;
  scf_tables(cr, p, &npw, dims, &nk, &nops);  // validates the request
  const size_t n = (size_t)npw, nb = (size_t)std::min(p.n_bands, npw), N = (size_t)dims[0] * dims[1] * dims[2];
  size_t nch = 0;
  for (int a = 0; a < cr.natoms; ++a) {
    const pwk_species& s = cr.species_data[cr.species[a]];
    for (int i = 0; i < s.nproj; ++i) nch += (size_t)(2 * s.proj_l[i] + 1);
  }
  const size_t ns = p.nspin == 2 ? 2 : 1;
  // the same rule as Solve::setup
  const bool cheb = p.eigensolver == 1 || (p.eigensolver < 0 && !(npw <= 1500 && heev_supported(npw, (int)nb)));
  const size_t nbs = cheb ? std::min(n, nb + std::max<size_t>(8, nb / 4)) : nb;
  // both spin channels go through one eigensolver launch on small cubic grids (at most 2·nk blocks)
  const size_t nke = ns * (size_t)nk;
  size_t b = 0;
  b += N * (12 + 3 * 16 + 3 * 8) + ns * N * (4 * 8 + 16) + 4 * N * 16;  // |G|², radial tables, grids
  if (ns == 2) b += 5 * N * 8;
  b += 2 * N * 24;                                                        // V(r) slots, V_eff(G)
  b += ns * 2 * (size_t)std::max(1, p.mixing_ndim) * N * 8;               // mixer histories
  b += (size_t)nk * n * (12 + nch * 48);                                  // kinetic, B, D·B^H (FP64 + FP32)
  b += ns * (size_t)nk * n * nbs * (p.precision == PWK_FP64 ? 16 : 24);   // eigenvectors (+ FP32 copy)
  if (cheb) {
    b += 8 * nke * n * nbs * 16;                                          // filter and Rayleigh-Ritz blocks
    b += 8 * nke * nbs * nbs * 16 + heev_work_doubles((int)nbs, (int)nbs, (int)nke) * 8;
    b += (size_t)nk * std::max<size_t>(1, nch) * nbs * 16 + 3 * nke * n * 16;  // projections, Lanczos
  } else {
    b += (size_t)nk * n * n * 16 + heev_work_doubles(npw, (int)nb, nk) * 8;    // dense H batch, scratch
  }
  b += (size_t)nk * n * nb * 16 + (size_t)nk * std::max<size_t>(1, nch) * nb * 16;  // final expectations
  // density and H·psi workspaces: stick buffers of the band batch or per-group density partials
  b += std::max(nke * nbs * (size_t)dims[2] * n * 16, (size_t)128 * N * 8) + ((size_t)64 << 20);
  return b;
}

}  // namespace pwk

struct pwk_server {
  pwk::Server s;
};

namespace {
template <class F>
int serve_guarded(F&& f) {
  try {
    f();
    return PWK_OK;
  } catch (const pwk::Error& e) {
    pwk::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    pwk::set_last_error(e.what());
    return PWK_E_DEVICE;
  }
}
}  // namespace

extern "C" {

int pwk_server_create(int device, int n_slots, double hbm_fraction, pwk_server** out) {
  return serve_guarded([&] {
    PWK_REQUIRE(out && n_slots >= 1 && n_slots <= 64 && hbm_fraction > 0.0 && hbm_fraction <= 1.0, PWK_E_INVALID,
                "bad server arguments (1..64 slots, 0 < hbm_fraction <= 1)");
    *out = nullptr;
    int nd = 0;
    PWK_REQUIRE(hipGetDeviceCount(&nd) == hipSuccess && device >= 0 && device < nd, PWK_E_NO_DEVICE,
                "no such HIP device");
    PWK_HIP(hipSetDevice(device));
    size_t free_b = 0, total_b = 0;
    PWK_HIP(hipMemGetInfo(&free_b, &total_b));
    auto srv = std::make_unique<pwk_server>();
    srv->s.device = device;
    // one eighth of the budget is left to the workers' caches of released blocks
    const size_t total = (size_t)((double)free_b * hbm_fraction);
    srv->s.cache_cap = total / 8 / (size_t)n_slots;
    srv->s.budget = total - srv->s.cache_cap * (size_t)n_slots;
    for (int i = 0; i < n_slots; ++i) srv->s.workers.emplace_back([p = &srv->s] { p->worker_main(); });
    *out = srv.release();
  });
}

int pwk_server_submit(pwk_server* srv, const pwk_crystal* crystal, const pwk_scf_params* params, int64_t* ticket) {
  return serve_guarded([&] {
    PWK_REQUIRE(srv && crystal && params && ticket, PWK_E_INVALID, "null argument");
    auto job = std::make_shared<pwk::ServeJob>();
    job->cr = *crystal;
    job->p = *params;
    job->bytes = pwk::scf_footprint(*crystal, *params);
    job->t_submit = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(srv->s.mu);
      PWK_REQUIRE(!srv->s.stopping, PWK_E_STATE, "server is shutting down");
      job->id = srv->s.next_id++;
      srv->s.jobs[job->id] = job;
      srv->s.queue.push_back(job);
    }
    srv->s.cv_work.notify_all();
    *ticket = job->id;
  });
}

int pwk_server_wait(pwk_server* srv, int64_t ticket, pwk_scf_result* result, double* eigenvalues, double* rho,
                    double* latency_s) {
  std::shared_ptr<pwk::ServeJob> job;
  int st = serve_guarded([&] {
    PWK_REQUIRE(srv, PWK_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(srv->s.mu);
    auto it = srv->s.jobs.find(ticket);
    PWK_REQUIRE(it != srv->s.jobs.end(), PWK_E_INVALID, "unknown ticket");
    job = it->second;
    srv->s.cv_done.wait(lk, [&] { return job->done; });
    srv->s.jobs.erase(it);
  });
  if (st != PWK_OK) return st;
  if (result) *result = job->res;
  if (eigenvalues) std::copy(job->eig.begin(), job->eig.end(), eigenvalues);
  if (rho) std::copy(job->rho.begin(), job->rho.end(), rho);
  if (latency_s) *latency_s = std::chrono::duration<double>(job->t_end - job->t_submit).count();
  if (job->status != PWK_OK) pwk::set_last_error(job->error);
  return job->status;
}

int pwk_server_set_pass_batching(pwk_server* srv, int window_us) {
  return serve_guarded([&] {
    PWK_REQUIRE(srv && window_us >= 0 && window_us <= 100000, PWK_E_INVALID,
                "bad batching window (0..100000 microseconds)");
    std::lock_guard<std::mutex> lk(srv->s.batcher.mu);
    srv->s.batcher.window_us = window_us;
  });
}

int pwk_server_pass_stats(pwk_server* srv, int64_t* launches, int64_t* passes) {
  return serve_guarded([&] {
    PWK_REQUIRE(srv && launches && passes, PWK_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->s.batcher.mu);
    *launches = srv->s.batcher.launches;
    *passes = srv->s.batcher.passes;
  });
}

int pwk_server_destroy(pwk_server* srv) {
  return serve_guarded([&] {
    if (!srv) return;
    {
      std::lock_guard<std::mutex> lk(srv->s.mu);
      srv->s.stopping = true;
    }
    srv->s.cv_work.notify_all();
    for (auto& t : srv->s.workers) t.join();
    delete srv;
  });
}

}  // extern "C"
