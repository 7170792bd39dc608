// Render server: the request loop above Integrator::render (no counterpart
// in the reference, whose callers render one image per process).
//
// Continuous batching: jobs wait in a FIFO and are admitted into job slots
// as soon as a slot is free and the HBM budget covers the job's footprint.
// Every slot owns a grow-only workspace and a pinned output buffer.
//
// Path-tracer and gradient-domain jobs of at least pool_min_paths paths
// render in the pool (pool.hip): one set of wavefront queues holding the
// paths of every such job, refilled every bounce from the jobs in order of
// remaining work.  Smaller ones, BDPT jobs and megakernel jobs render in
// chunks on up to `max_concurrent_jobs` lanes (one HIP stream each); a BDPT
// chunk is issued in pieces (camera subpaths, light subpaths, connection
// passes of a quarter chunk), each about 20 ms of device time at the
// default chunk, and a job of at most two chunks' work runs as one chunk.
//
// Scheduling is shortest remaining processing time first over all jobs:
// whenever the pool holds the job with the least remaining work it issues
// iterations (two in flight), otherwise that job issues its next piece and
// the pool issues nothing new; the others wait, so the preemption points are
// pool iterations and pieces.  A job that has issued nothing for kStarveMax
// issues anyway.  Inside the pool a draining job (every path generated) is
// the top one, and the other jobs keep at most pool_fill_paths paths alive
// meanwhile.  Stream priorities are not used.  A job's film depends only on
// its own description (per-job seed, deterministic accumulation in sample
// order, independent of chunking, lanes and what shared the pool), never on
// what it shared the device with.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <thread>

#include "ar_job.h"
#include "ar_pool.h"

// Built with -DAR_SERVER_TRACE, the scheduler prints its events (server
// clock, job, what happened) to stderr.
#ifdef AR_SERVER_TRACE
#include <cstdio>
#define AR_STRACE(...) std::fprintf(stderr, __VA_ARGS__)
#else
#define AR_STRACE(...)                                 \
  do {                                                 \
    if (false) std::fprintf(stderr, __VA_ARGS__);      \
  } while (0)
#endif

namespace ar {

constexpr uint32_t kMaxLanes = 4;            // hardware queues a process gets by default
constexpr uint64_t kServerChunkPaths = 1ull << 23;   // PT / GDPT path-equivalents per chunk (default)
constexpr uint64_t kServerChunkBdpt = 1ull << 22;    // BDPT samples per chunk (default)
constexpr double kStarveMax = 1.0;                   // s a job waits for higher-priority work at most

// Relative device cost of one pixel-sample of each integrator on the room
// (PT path 1 at 223 Mpaths/s; GDPT sample = base + two shifted paths at 215
// Mpaths/s; BDPT sample = both subpaths and all connections at 37 Msamples/s),
// used to estimate device time for scheduling.
static double path_weight(uint32_t integrator) {
  return integrator == AR_INTEGRATOR_BDPT ? 6.0 : (integrator == AR_INTEGRATOR_GDPT ? 3.1 : 1.0);
}

struct Slot {
  hipStream_t stream = nullptr;
  Workspace ws;
  float* pinned = nullptr;
  size_t pinned_bytes = 0;
  bool busy = false;
  size_t footprint = 0;  // of the job holding the slot
};

struct Job {
  uint64_t id = 0;
  JobSpec spec;
  int32_t state = AR_JOB_QUEUED;
  ar_status error = AR_OK;
  std::string error_msg;
  double t_submit = 0, t_start = 0, t_finish = 0;
  size_t footprint = 0;
  int slot = -1;
  bool cancel = false;
  bool issued_all = false;   // every chunk + resolve + copy-out enqueued
  double work_left = 0;      // weighted paths not yet issued
  double work_chunk = 0;     // weighted paths per chunk
  std::unique_ptr<JobRun> run;
  hipEvent_t begun = nullptr;  // after the set-up work begin() enqueued on the slot's stream
  int lane = -1;               // lane holding the job's chunk on the device, -1: none
  bool short_job = false;      // runs as one chunk
  double last_issue = 0;       // server clock of its last chunk
  bool pooled = false;         // PT / GDPT job rendered in the shared pool
  int pool_h = -1;             // its pool handle once started
  std::vector<float> film;
  ar_render_stats stats{};
};

}  // namespace ar

using namespace ar;

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;  // recorded after the chunk in flight
  std::shared_ptr<Job> job;   // the job whose chunk is in flight, null when free
};

struct ar_server {
  const ar_scene* scene = nullptr;
  ar_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;
  std::vector<std::shared_ptr<Job>> active;
  std::vector<std::unique_ptr<Slot>> slots;
  std::vector<Lane> lanes;
  std::vector<hipEvent_t> event_pool;
  int waiters = 0;       // threads inside ar_job_wait (ar_server_destroy waits for them)
  bool broken = false;  // the worker thread ended on an exception
  bool stop = false;
  uint64_t next_id = 1;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;
  int device = 0;
  std::unique_ptr<PathPool> pool;  // shared path-tracer queues (null: every job runs in chunks)
  uint32_t pool_min_paths = 0;
  double pool_last_issue = 0;  // server clock of the pool's last iteration (or last idle moment)

  // PT / GDPT wavefront jobs of at least pool_min_paths paths share the pool.
  bool poolable(const JobSpec& spec) const {
    const ar_render_config& c = spec.cfg;
    if (!pool || c.schedule != AR_SCHED_WAVEFRONT) return false;
    if (c.integrator != AR_INTEGRATOR_PT && c.integrator != AR_INTEGRATOR_GDPT) return false;
    const uint64_t paths = (uint64_t)spec.width * spec.height * c.spp * (c.integrator == AR_INTEGRATOR_GDPT ? 3 : 1);
    return paths >= pool_min_paths;
  }

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    return e;
  }

  void release(Job& j) {  // caller holds mu
    if (j.begun) {
      event_pool.push_back(j.begun);
      j.begun = nullptr;
    }
    if (j.pool_h >= 0) {
      pool->release(j.pool_h);
      j.pool_h = -1;
    }
    if (j.slot >= 0) {
      slots[j.slot]->busy = false;
      slots[j.slot]->footprint = 0;
      j.slot = -1;
    }
    j.run.reset();
  }

  void fail_job(Job& j, ar_status code, const std::string& msg) {  // caller holds mu
    j.state = AR_JOB_FAILED;
    j.error = code;
    j.error_msg = msg;
    j.t_finish = now();
  }

  // Last resort of the worker thread: an exception escaping the scheduling
  // loop (host memory) fails every unfinished job, so that no waiter blocks
  // forever, and the server takes no new jobs.  Device work already issued
  // keeps its buffers until ar_server_destroy synchronises the streams.
  void fail_all(const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    broken = true;
    for (auto& kv : jobs) {
      Job& j = *kv.second;
      if (j.state == AR_JOB_QUEUED || j.state == AR_JOB_RUNNING) fail_job(j, AR_ERR_OUT_OF_BUDGET, msg);
    }
    pending.clear();
    cv_done.notify_all();
  }

  // Device bytes the slots hold: a busy slot at least its job's footprint,
  // an idle one what its workspace retains (workspaces grow and are kept for
  // the next job).  Only the worker thread changes workspaces.
  size_t committed_bytes() const {
    size_t total = 0;
    for (auto& sp : slots) total += sp->busy ? std::max(sp->ws.bytes(), sp->footprint) : sp->ws.bytes();
    return total;
  }

  // Picks queued jobs for free slots within the budget (caller holds mu) and
  // reserves the slot; the device work of starting them (JobRun::begin, which
  // may allocate and synchronise the slot's stream) runs in start() without
  // the lock.  An idle slot's retained workspace is freed when the budget
  // needs its bytes.
  void admit(std::vector<std::shared_ptr<Job>>& starting) {
    while (!pending.empty()) {
      auto it = jobs.find(pending.front());
      if (it == jobs.end() || it->second->state != AR_JOB_QUEUED) {
        pending.pop_front();
        continue;
      }
      Job& j = *it->second;
      if (j.footprint > cfg.hbm_budget_bytes) {
        pending.pop_front();
        fail_job(j, AR_ERR_OUT_OF_BUDGET, "job footprint exceeds the HBM budget");
        cv_done.notify_all();
        continue;
      }
      int free_slot = -1;
      for (int pass = 0; pass < 2 && free_slot < 0; ++pass) {
        const size_t committed = committed_bytes();
        for (size_t s = 0; s < slots.size(); ++s) {
          const Slot& sl = *slots[s];
          if (sl.busy) continue;
          const size_t held = sl.ws.bytes();
          if (committed - held + std::max(held, j.footprint) <= cfg.hbm_budget_bytes) {
            free_slot = (int)s;
            break;
          }
        }
        if (free_slot >= 0 || pass == 1) break;
        // no free slot fits as things stand: free the idle workspaces (their
        // streams have drained: a slot is idle once its job completed)
        bool trimmed = false;
        for (auto& sp : slots)
          if (!sp->busy && sp->ws.bytes() > 0) {
            (void)hipStreamSynchronize(sp->stream);
            sp->ws.trim();
            trimmed = true;
          }
        if (!trimmed) break;
      }
      if (free_slot < 0) break;
      pending.pop_front();
      Slot& sl = *slots[free_slot];
      sl.busy = true;
      sl.footprint = j.footprint;
      j.slot = free_slot;
      j.state = AR_JOB_RUNNING;
      j.t_start = now();
      j.last_issue = j.t_start;
      starting.push_back(it->second);
    }
  }

  // Device-side start of an admitted job (worker thread, mu not held).
  void start(const std::shared_ptr<Job>& jp) {
    Job& j = *jp;
    Slot& sl = *slots[j.slot];
    const size_t npix = (size_t)j.spec.width * j.spec.height;
    const size_t out_bytes = npix * 3 * sizeof(float);
    ar_status e = AR_OK;
    std::string msg;
    if (sl.ws.out.bytes < out_bytes) {
      (void)hipStreamSynchronize(sl.stream);
      if (sl.ws.out.alloc(out_bytes) != hipSuccess) {
        e = AR_ERR_DEVICE;
        msg = "output buffer allocation failed";
      }
    }
    if (e == AR_OK && sl.pinned_bytes < out_bytes) {
      if (sl.pinned) (void)hipHostFree(sl.pinned);
      sl.pinned = nullptr;
      sl.pinned_bytes = 0;
      if (hipHostMalloc((void**)&sl.pinned, out_bytes, hipHostMallocDefault) != hipSuccess) {
        e = AR_ERR_DEVICE;
        msg = "pinned buffer allocation failed";
      } else {
        sl.pinned_bytes = out_bytes;
      }
    }
    if (e == AR_OK && j.pooled) {
      e = pool->add(j.spec, &sl.ws, sl.ws.out.as<float>(), sl.pinned, sl.stream, j.id, &j.pool_h);
      if (e != AR_OK) msg = std::string("pool: ") + ar_last_error();
    } else if (e == AR_OK) {
      if (j.spec.cfg.sample_chunk == 0) {
        // a job of up to two chunks' work runs as one chunk; larger jobs are
        // split so that the device never holds much of them at a time
        uint64_t per_sample = npix * (j.spec.cfg.integrator == AR_INTEGRATOR_GDPT ? 3 : 1);
        uint64_t target = j.spec.cfg.integrator == AR_INTEGRATOR_BDPT ? cfg.bdpt_chunk_samples : cfg.chunk_paths;
        if (per_sample * j.spec.cfg.spp <= 2 * target) {
          j.spec.cfg.sample_chunk = j.spec.cfg.spp;
          j.short_job = true;
        }
        else
          j.spec.cfg.sample_chunk = (uint32_t)std::max<uint64_t>(
              1, std::min<uint64_t>(j.spec.cfg.spp, target / std::max<uint64_t>(per_sample, 1)));
      }
      const double w = path_weight(j.spec.cfg.integrator) * (double)npix;
      j.work_left = w * j.spec.cfg.spp;
      j.work_chunk = w * chunk_samples(j.spec);
      j.run = make_job(scene, j.spec, &sl.ws, sl.stream, sl.ws.out.as<float>(), nullptr);
      e = j.run->begin();
      if (e != AR_OK) {
        msg = std::string("job begin: ") + ar_last_error();
      } else {
        j.begun = get_event();
        if (!j.begun || hipEventRecord(j.begun, sl.stream) != hipSuccess) {
          e = AR_ERR_DEVICE;
          msg = "event record failed";
        }
      }
    }
    std::lock_guard<std::mutex> lk(mu);
    if (e != AR_OK) {
      fail_job(j, e, msg);
      release(j);
      cv_done.notify_all();
      return;
    }
    active.push_back(jp);
  }

  // Remaining work of an active job in path-equivalents, the order of
  // shortest remaining processing time first: a pooled job counts the paths
  // it has still to generate (0 while its last paths drain), a chunked job
  // the work of its chunks not yet issued.
  double remaining(const Job& j) const {
    if (j.pooled) {
      if (j.pool_h < 0) return 0.0;
      const double per = j.spec.cfg.integrator == AR_INTEGRATOR_GDPT ? path_weight(AR_INTEGRATOR_GDPT) / 3.0 : 1.0;
      return per * (double)pool->remaining_paths(j.pool_h);
    }
    return j.work_left;
  }

  // Completes a job whose work has drained (caller holds mu).
  void complete(const std::shared_ptr<Job>& jp) {
    Job& j = *jp;
    release(j);
    active.erase(std::find(active.begin(), active.end(), jp));
    cv_done.notify_all();
    cv_work.notify_all();
  }

  // One pass: retire finished chunks and pool iterations, complete drained
  // jobs, then let the job with the least remaining work issue work: the
  // pool when one of its jobs is that job, otherwise that job's next chunk
  // on a free lane (shortest remaining processing time first, preempting at
  // pool iterations and chunk boundaries).  The pool fills what its top job
  // leaves of each iteration with its other jobs' paths.  A job that has
  // issued nothing for kStarveMax issues anyway.  Returns true if any
  // progress was made.  Called without mu held; takes it only for
  // bookkeeping.
  bool pump() {
    bool progress = false;
    for (auto& ln : lanes) {
      if (!ln.job || hipEventQuery(ln.done) != hipSuccess) continue;
      Job& j = *ln.job;
      j.lane = -1;
      ln.job.reset();
      progress = true;
    }
    std::vector<std::shared_ptr<Job>> act;
    {
      std::lock_guard<std::mutex> lk(mu);
      act = active;
    }
    auto pool_failed = [&](ar_status e) {
      // a device error in the shared queues ends every pooled job
      (void)hipStreamSynchronize(pool->stream());
      std::lock_guard<std::mutex> lk(mu);
      for (auto& jp : act)
        if (jp->pooled) {
          fail_job(*jp, e, std::string("pool: ") + ar_last_error());
          complete(jp);
        }
      return true;
    };
    if (pool) {
      // cancellations reach the pool before it plans another iteration
      for (auto& jp : act)
        if (jp->pooled && jp->pool_h >= 0 && jp->cancel) pool->cancel(jp->pool_h);
      const ar_status e = pool->pump(false, &progress);
      if (e != AR_OK) return pool_failed(e);
    }
    for (auto& jp : act) {
      Job& j = *jp;
      if (j.lane >= 0) continue;
      std::lock_guard<std::mutex> lk(mu);
      if (j.pooled) {
        // film resolved and copied (state 1, done event complete) or
        // cancelled and drained (state 2)
        if (j.pool_h < 0) continue;
        const int ps = pool->state(j.pool_h);
        if (ps == 1 && hipEventQuery(pool->done_event(j.pool_h)) == hipSuccess) {
          Slot& sl = *slots[j.slot];
          size_t n = (size_t)j.spec.width * j.spec.height * 3;
          j.film.assign(sl.pinned, sl.pinned + n);
          pool->stats(j.pool_h, &j.stats);
          j.state = j.cancel ? AR_JOB_CANCELLED : AR_JOB_DONE;
        } else if (ps == 2) {
          j.state = AR_JOB_CANCELLED;
        } else {
          continue;
        }
        j.t_finish = now();
        AR_STRACE("T %.4f done pooled job %llu\n", j.t_finish, (unsigned long long)j.id);
        complete(jp);
        progress = true;
        continue;
      }
      if (!j.issued_all && !j.cancel) continue;
      // drained: complete the job
      if (j.cancel && !j.issued_all) {
        j.state = AR_JOB_CANCELLED;
        j.t_finish = now();
      } else {
        Slot& sl = *slots[j.slot];
        size_t n = (size_t)j.spec.width * j.spec.height * 3;
        j.film.assign(sl.pinned, sl.pinned + n);
        (void)j.run->read_stats(&j.stats);
        j.state = j.cancel ? AR_JOB_CANCELLED : AR_JOB_DONE;
        j.t_finish = now();
      }
      AR_STRACE("T %.4f done job %llu\n", j.t_finish, (unsigned long long)j.id);
      complete(jp);
      progress = true;
    }
    // priorities: the least remaining work of the pool's jobs and of the
    // chunked jobs (those with a chunk in flight included)
    double pool_min = std::numeric_limits<double>::infinity();
    double chunk_min = std::numeric_limits<double>::infinity();
    std::shared_ptr<Job> chunk_top;
    const double tnow = now();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& jp : active) {
        Job& j = *jp;
        if (j.cancel) continue;
        if (j.pooled) {
          if (j.pool_h >= 0 && pool->state(j.pool_h) == 0) pool_min = std::min(pool_min, remaining(j));
          continue;
        }
        if (!j.run || (j.issued_all && j.lane < 0)) continue;
        const double r = remaining(j);
        if (!chunk_top || r < chunk_min || (r == chunk_min && j.t_submit < chunk_top->t_submit)) {
          chunk_min = r;
          chunk_top = jp;
        }
      }
    }
    if (pool && !pool->idle()) {
      const bool turn = pool_min <= chunk_min || tnow - pool_last_issue >= kStarveMax;
      const uint32_t inflight0 = pool->in_flight();
      const ar_status e = pool->pump(turn, &progress);
      if (e != AR_OK) return pool_failed(e);
      if (turn) pool_last_issue = tnow;
      if (pool->in_flight() > inflight0)
        AR_STRACE("T %.4f pool iterations %u in flight, pool_min %.0f chunk_min %.0f\n", tnow, pool->in_flight(),
                  pool_min, chunk_min);
    } else {
      pool_last_issue = tnow;
    }
    for (size_t li = 0; li < lanes.size(); ++li) {
      Lane& ln = lanes[li];
      if (ln.job) continue;
      // the chunked job with the least remaining work, if no pooled job has
      // less; any job that has waited kStarveMax since its last chunk
      std::shared_ptr<Job> pick;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& jp : active) {
          Job& j = *jp;
          if (j.pooled || j.lane >= 0 || j.issued_all || j.cancel || !j.run) continue;
          const bool top = jp == chunk_top && chunk_min < pool_min;
          if (!top && tnow - j.last_issue < kStarveMax) continue;
          if (!pick || j.work_left < pick->work_left ||
              (j.work_left == pick->work_left && j.t_submit < pick->t_submit))
            pick = jp;
        }
      }
      if (!pick) break;
      Job& j = *pick;
      ar_status e = AR_OK;
      if (j.begun && hipStreamWaitEvent(ln.stream, j.begun, 0) != hipSuccess)
        e = fail(AR_ERR_DEVICE, "hipStreamWaitEvent failed");
      j.run->stream = ln.stream;
      bool done = false;
      if (e == AR_OK) e = j.run->step(&done);
      j.work_left = std::max(0.0, j.work_left - j.work_chunk * j.run->last_step_share());
      if (e == AR_OK && done) {
        e = j.run->finish();
        if (e == AR_OK) {
          Slot& sl = *slots[j.slot];
          size_t bytes = (size_t)j.spec.width * j.spec.height * 3 * sizeof(float);
          if (hipMemcpyAsync(sl.pinned, sl.ws.out.p, bytes, hipMemcpyDeviceToHost, ln.stream) != hipSuccess)
            e = fail(AR_ERR_DEVICE, "copy-out failed");
        }
        j.issued_all = true;
      }
      if (e == AR_OK && hipEventRecord(ln.done, ln.stream) != hipSuccess) e = fail(AR_ERR_DEVICE, "event record failed");
      if (e != AR_OK) {
        (void)hipStreamSynchronize(ln.stream);
        std::lock_guard<std::mutex> lk(mu);
        fail_job(j, e, ar_last_error());
        complete(pick);
        progress = true;
        continue;
      }
      j.lane = (int)li;
      j.last_issue = now();
      AR_STRACE("T %.4f issue job %llu lane %zu share %.3f left %.0f pool_min %.0f\n", j.last_issue,
                (unsigned long long)j.id, li, j.run->last_step_share(), j.work_left, pool_min);
      ln.job = pick;
      progress = true;
    }
    return progress;
  }

  void loop() {
    (void)hipSetDevice(device);
    std::vector<std::shared_ptr<Job>> starting;
    for (;;) {
      starting.clear();
      {
        std::unique_lock<std::mutex> lk(mu);
        if (stop && active.empty() && pending.empty()) break;
        admit(starting);
        if (active.empty() && starting.empty()) {
          if (stop) continue;
          cv_work.wait_for(lk, std::chrono::milliseconds(5));
          continue;
        }
      }
      for (auto& jp : starting) {
        const double t0s = now();
        start(jp);
        AR_STRACE("T %.4f start job %llu integ %u %ux%u spp %u pooled %d took %.2f ms\n", t0s,
                  (unsigned long long)jp->id, jp->spec.cfg.integrator, jp->spec.width, jp->spec.height, jp->spec.cfg.spp,
                  (int)jp->pooled, (now() - t0s) * 1e3);
      }
      if (!pump()) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
  }
};

extern "C" {

ar_status ar_server_create(const ar_scene* s, const ar_server_config* cfg, ar_server** out) try {
  if (!s || !out) return fail(AR_ERR_INVALID_ARG, "ar_server_create: NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_server_create: scene not committed");
  std::unique_ptr<ar_server> srv(new ar_server());
  srv->scene = s;
  if (cfg) srv->cfg = *cfg;
  if (srv->cfg.hbm_budget_bytes == 0) srv->cfg.hbm_budget_bytes = 200ull << 30;
  if (srv->cfg.chunk_paths == 0) srv->cfg.chunk_paths = (uint32_t)kServerChunkPaths;
  if (srv->cfg.bdpt_chunk_samples == 0) srv->cfg.bdpt_chunk_samples = (uint32_t)kServerChunkBdpt;
  if (srv->cfg.max_active_jobs == 0) srv->cfg.max_active_jobs = 8;
  if (srv->cfg.max_active_jobs > 64) return fail(AR_ERR_INVALID_ARG, "max_active_jobs above 64");
  if (srv->cfg.max_concurrent_jobs == 0) srv->cfg.max_concurrent_jobs = std::min<uint32_t>(3, srv->cfg.max_active_jobs);
  if (srv->cfg.max_concurrent_jobs > kMaxLanes)
    return fail(AR_ERR_INVALID_ARG, "max_concurrent_jobs above 4 (the hardware queues of a process)");
  (void)hipGetDevice(&srv->device);
  // The shared path-tracer queues take their device memory out of the
  // budget; a budget too small for them leaves every job in chunks.
  if (srv->cfg.pool_paths != AR_POOL_OFF) {
    const uint32_t cap = srv->cfg.pool_paths ? srv->cfg.pool_paths : kPoolDefaultPaths;
    if (cap > (1u << 26)) return fail(AR_ERR_INVALID_ARG, "pool_paths above 2^26");
    const size_t fp = PathPool::footprint(cap);
    if (fp < srv->cfg.hbm_budget_bytes / 2) {
      const uint32_t fill = srv->cfg.pool_fill_paths ? srv->cfg.pool_fill_paths : cap / 8;
      const ar_status e = PathPool::create(s, cap, fill, &srv->pool);
      if (e != AR_OK) return e;
      srv->cfg.hbm_budget_bytes -= fp;
      srv->pool_min_paths = srv->cfg.pool_min_paths ? srv->cfg.pool_min_paths : (1u << 22);
    }
  }
  auto drop_streams = [&srv]() {
    for (auto& sp : srv->slots)
      if (sp->stream) (void)hipStreamDestroy(sp->stream);
    for (auto& ln : srv->lanes) {
      if (ln.stream) (void)hipStreamDestroy(ln.stream);
      if (ln.done) (void)hipEventDestroy(ln.done);
    }
  };
  for (uint32_t i = 0; i < srv->cfg.max_active_jobs; ++i) {
    srv->slots.emplace_back(new Slot());
    const hipError_t e = hipStreamCreateWithFlags(&srv->slots.back()->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      drop_streams();
      return fail(AR_ERR_DEVICE, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
  }
  for (uint32_t i = 0; i < srv->cfg.max_concurrent_jobs; ++i) {
    srv->lanes.emplace_back();
    Lane& ln = srv->lanes.back();
    hipError_t e = hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ln.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      drop_streams();
      return fail(AR_ERR_DEVICE, std::string("lane stream: ") + hipGetErrorString(e));
    }
  }
  srv->t0 = std::chrono::steady_clock::now();
  ar_server* p = srv.get();
  try {
    p->worker = std::thread([p]() {
      try {
        p->loop();
      } catch (const std::exception& e) {
        p->fail_all(std::string("server worker: ") + e.what());
      }
    });
  } catch (...) {
    drop_streams();
    throw;
  }
  srv.release();
  *out = p;
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_server_create");
}

ar_status ar_server_destroy(ar_server* srv) {
  if (!srv) return AR_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint64_t id : srv->pending) {
      auto it = srv->jobs.find(id);
      if (it != srv->jobs.end() && it->second->state == AR_JOB_QUEUED) it->second->state = AR_JOB_CANCELLED;
    }
    srv->pending.clear();
    for (auto& j : srv->active) j->cancel = true;
    srv->stop = true;
    srv->cv_done.notify_all();
  }
  srv->cv_work.notify_all();
  srv->worker.join();
  {
    // every job is terminal now; let threads still in ar_job_wait return
    // before the condition variable goes away
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.notify_all();
    srv->cv_done.wait(lk, [srv]() { return srv->waiters == 0; });
  }
  srv->pool.reset();  // drains its stream
  for (auto& ln : srv->lanes) {
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
    if (ln.stream) (void)hipStreamDestroy(ln.stream);
    if (ln.done) (void)hipEventDestroy(ln.done);
  }
  for (auto& sp : srv->slots) {
    Slot& sl = *sp;
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto e : srv->event_pool) (void)hipEventDestroy(e);
  delete srv;
  return AR_OK;
}

ar_status ar_job_submit(ar_server* srv, const ar_job_desc* desc, uint64_t* out_id) try {
  if (!srv || !desc) return fail(AR_ERR_INVALID_ARG, "ar_job_submit: NULL argument");
  auto j = std::make_shared<Job>();
  j->spec.width = desc->width;
  j->spec.height = desc->height;
  j->spec.camera = desc->camera;
  j->spec.cfg = desc->config;
  ar_status e = validate_job(srv->scene, j->spec);
  if (e != AR_OK) return e;
  // BDPT connection passes of at most a quarter chunk: the pieces the
  // server issues between its scheduling decisions stay short
  if (j->spec.cfg.integrator == AR_INTEGRATOR_BDPT)
    j->spec.conn_pass = std::max<uint64_t>(64, srv->cfg.bdpt_chunk_samples / 4);
  j->pooled = srv->poolable(j->spec);
  j->footprint = j->pooled ? PathPool::job_bytes(j->spec) : job_footprint(srv->scene, j->spec);
  if (j->footprint > srv->cfg.hbm_budget_bytes) return fail(AR_ERR_OUT_OF_BUDGET, "job exceeds the HBM budget");
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(AR_ERR_STATE, "server is shutting down");
    if (srv->broken) return fail(AR_ERR_STATE, "the server's worker thread failed; destroy the server");
    j->id = srv->next_id++;
    j->t_submit = srv->now();
    srv->jobs[j->id] = j;
    srv->pending.push_back(j->id);
    if (out_id) *out_id = j->id;
  }
  srv->cv_work.notify_all();
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_job_submit");
}

ar_status ar_job_poll(ar_server* srv, uint64_t id, int32_t* state) {
  if (!srv || !state) return fail(AR_ERR_INVALID_ARG, "ar_job_poll: NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(id);
  if (it == srv->jobs.end()) return fail(AR_ERR_NOT_FOUND, "unknown job id");
  *state = it->second->state;
  return AR_OK;
}

ar_status ar_job_wait(ar_server* srv, uint64_t id, float* out_rgb, ar_job_stats* stats) {
  if (!srv) return fail(AR_ERR_INVALID_ARG, "ar_job_wait: NULL server");
  std::shared_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(id);
    if (it == srv->jobs.end()) return fail(AR_ERR_NOT_FOUND, "unknown job id");
    j = it->second;
    srv->waiters++;
    srv->cv_done.wait(lk, [&]() {
      return j->state == AR_JOB_DONE || j->state == AR_JOB_CANCELLED || j->state == AR_JOB_FAILED;
    });
    srv->jobs.erase(id);
    if (--srv->waiters == 0) srv->cv_done.notify_all();
  }
  if (stats) {
    stats->submit_time = j->t_submit;
    stats->start_time = j->t_start;
    stats->finish_time = j->t_finish;
    stats->paths = j->stats.paths;
    stats->rays = j->stats.rays;
    stats->state = j->state;
  }
  if (j->state == AR_JOB_CANCELLED) return fail(AR_ERR_CANCELLED, "job was cancelled");
  if (j->state == AR_JOB_FAILED) return fail(j->error, j->error_msg);
  if (out_rgb) memcpy(out_rgb, j->film.data(), j->film.size() * sizeof(float));
  return AR_OK;
}

ar_status ar_job_cancel(ar_server* srv, uint64_t id) {
  if (!srv) return fail(AR_ERR_INVALID_ARG, "ar_job_cancel: NULL server");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(id);
  if (it == srv->jobs.end()) return fail(AR_ERR_NOT_FOUND, "unknown job id");
  Job& j = *it->second;
  if (j.state == AR_JOB_QUEUED) {
    j.state = AR_JOB_CANCELLED;
    j.t_finish = srv->now();
    srv->cv_done.notify_all();
  } else if (j.state == AR_JOB_RUNNING) {
    j.cancel = true;
  }
  return AR_OK;
}

double ar_server_clock(const ar_server* srv) { return srv ? srv->now() : 0.0; }

}  // extern "C"
