// Serving pool: continuous batching of path-tracer and gradient-domain jobs
// in one set of wavefront queues (no counterpart in the reference, whose
// PathTracer::render, path_tracer.rs:159-191, renders one image per call).
//
// The pool keeps up to C paths in flight, taken from any of the jobs it
// holds.  Their state sits in two compacted structure-of-arrays buffers:
// the paths alive at the start of an iteration occupy places [0, n) of one
// buffer, and shading writes the survivors densely into the other.  One
// iteration:
//   begin     the host's generation plan (jobs, highest priority first) is
//             written to device memory; the other buffer and the shadow
//             queue are emptied;
//   generate  places [n, C) are refilled with new paths from the plan's jobs
//             in order, each job handing out the paths of its current
//             window in index order;
//   advance   each job's cursor moves past the paths generated;
//   extend    closest hit for every path (the path tracer's stage);
//   shade     one bounce of path_tracer.rs:59-156 (pt_shade); a surviving
//             path moves to the other buffer, a finished one writes its
//             radiance into its job's window buffer;
//   shadow    any hit; an unoccluded contribution is added where the path's
//             radiance now is (the other buffer, or the window buffer);
// then the jobs' progress counters are copied to pinned host memory.
//
// A path carries its job and its index within the job's window; the index
// fixes pixel, sample and variant exactly as a chunk of ar_render does, and
// every path runs the same functions in the same order, so its radiance is
// ar_render's bit for bit.  Once every path of a window has finished, the
// host adds the window to the job's film in sample order: the film equals
// ar_render's bit for bit whatever the job shared the pool with.  Jobs of
// the plan are ordered by paths still to generate, so the job closest to
// completion takes the freed places first and the others fill the rest
// (its last bounces, which alone would leave most of the device idle, run
// beside their paths).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ar_pool.h"
#include "ar_wf.h"

namespace ar {

namespace {

#define AR_TRY(call)               \
  do {                             \
    const ar_status s_ = (call);   \
    if (s_ != AR_OK) return s_;    \
  } while (0)

constexpr int kPoolBlock = 512;        // generate / shade workgroup
constexpr uint32_t kPoolRing = 4;      // progress snapshots in flight
constexpr uint32_t kPoolInFlight = 2;  // iterations issued ahead of the host

// One job as the iterations read it (written by the host between iterations).
struct PoolJob {
  PtParams P;       // s0, k: the window being generated
  float4* Lout[2];  // window buffers of ring slots 0 and 1: [variant][sample][pixel], as a chunk's
  uint32_t ring;    // ring slot of the window being generated
  uint32_t end;     // its paths
};

// A job's progress, updated by the iterations and copied back after each.
struct PoolProg {
  uint32_t epoch;       // occupant of the table entry (snapshots of an earlier one are ignored)
  uint32_t gen_win;     // window being generated
  uint32_t cursor;      // its next path
  uint32_t win[2];      // window held by each ring slot
  uint32_t retired[2];  // paths of that window whose radiance is final
  uint32_t pad;
  unsigned long long rays;  // extension + shadow rays
};

struct PoolPlan {
  uint32_t n;
  uint32_t order[kPoolMaxJobs];  // table entries, highest priority first
  uint32_t limit[kPoolMaxJobs];  // paths each may hand out this iteration at most
};

struct PoolState {  // one compacted state buffer
  float4 *ray_o, *ray_d, *beta, *L, *prev_p;
  uint4* meta;        // (pixel, sample, dim, depth)
  uint32_t* job;      // job | ring << 16
  uint32_t* pidx;     // index within the job's window
};

struct PoolDev {
  PoolState S[2];
  uint32_t* iota;  // 0, 1, ..., C-1: the extension queue of a compacted buffer
  float4* hit;
  uint32_t* hit_inst;
  float4 *sh_o, *sh_d, *sh_c;
  float4** sh_tgt;                   // where an unoccluded contribution goes
  uint32_t* cnt;                     // [0], [1] paths in S[0], S[1]; [2] shadow rays; [3] generated;
                                     // [4] paths alive when the iteration began
  unsigned long long* stats;         // [0] extension, [1] shadow rays (whole pool)
  uint32_t* work;                    // walker counters of the extension and shadow launches (walk_items)
  PoolJob* jobs;
  PoolProg* prog;
  PoolPlan* plan;
  uint32_t cap;
};

// Pointers loaded from memory (window buffers, shadow targets) are generic;
// casting them to the global address space keeps their accesses global
// loads and stores instead of flat ones, which inside the traversal loop
// would also wait on the LDS counter of the stack.
__device__ __forceinline__ AR_GLOBAL_AS float4* as_global(float4* p) { return (AR_GLOBAL_AS float4*)p; }

__device__ __forceinline__ void pool_load(const PoolState& S, uint32_t i, PathState& s, uint32_t& job, uint32_t& p) {
  float4 o = S.ray_o[i], d = S.ray_d[i], b = S.beta[i], l = S.L[i], pp = S.prev_p[i];
  uint4 m = S.meta[i];
  s.ray.o = mk(o.x, o.y, o.z);
  s.ray.tmin = o.w;
  s.ray.d = mk(d.x, d.y, d.z);
  s.ray.tmax = d.w;
  s.beta = mk(b.x, b.y, b.z);
  s.prev_pdf = b.w;
  s.L = mk(l.x, l.y, l.z);
  s.prev_p = mk(pp.x, pp.y, pp.z);
  s.pixel = m.x;
  s.sample = m.y;
  s.dim = m.z;
  s.depth = m.w;
  job = S.job[i];
  p = S.pidx[i];
}

__device__ __forceinline__ void pool_store(const PoolState& S, uint32_t i, const PathState& s, uint32_t job,
                                           uint32_t p) {
  S.ray_o[i] = make_float4(s.ray.o.x, s.ray.o.y, s.ray.o.z, s.ray.tmin);
  S.ray_d[i] = make_float4(s.ray.d.x, s.ray.d.y, s.ray.d.z, s.ray.tmax);
  S.beta[i] = make_float4(s.beta.x, s.beta.y, s.beta.z, s.prev_pdf);
  S.L[i] = make_float4(s.L.x, s.L.y, s.L.z, 0.0f);
  S.prev_p[i] = make_float4(s.prev_p.x, s.prev_p.y, s.prev_p.z, 0.0f);
  S.meta[i] = make_uint4(s.pixel, s.sample, s.dim, s.depth);
  S.job[i] = job;
  S.pidx[i] = p;
}

// Adds v of every lane with `on` set to lds[key]: one LDS atomic per distinct
// key in the wave (the lanes of a wave mostly belong to one job).  Every lane
// of the wave calls it.
__device__ __forceinline__ void wave_keyed_add(uint32_t* lds, bool on, uint32_t key, uint32_t v) {
  uint64_t pending = __ballot(on);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t k = __shfl(key, leader);
    const bool mine = on && key == k && ((pending >> __lane_id()) & 1ull);
    const uint64_t m = __ballot(mine);
    uint32_t s = mine ? v : 0u;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((int)__lane_id() == leader) atomicAdd(&lds[k], s);
    pending &= ~m;
  }
}

__global__ void k_pool_init(PoolDev D) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 5) D.cnt[t] = 0;
  if (t < 2) D.stats[t] = 0ull;
  if (t < kWalkWorkBytes / sizeof(uint32_t)) D.work[t] = 0u;
  if (t < kPoolMaxJobs) {
    PoolProg p{};
    p.win[0] = p.win[1] = 0xffffffffu;
    D.prog[t] = p;
  }
  if (t == 0) D.plan->n = 0;
  for (uint32_t i = t; i < D.cap; i += gridDim.x * blockDim.x) D.iota[i] = i;
}

// A new job in table entry j; its first window goes to ring slot 0.
__global__ void k_pool_set_job(PoolDev D, uint32_t j, PoolJob J, uint32_t epoch) {
  if (threadIdx.x != 0) return;
  D.jobs[j] = J;
  PoolProg p{};
  p.epoch = epoch;
  p.win[0] = 0;
  p.win[1] = 0xffffffffu;
  D.prog[j] = p;
}

// Window w of job j starts generating in ring slot `ring` (whose previous
// window has been accumulated).
__global__ void k_pool_set_window(PoolDev D, uint32_t j, uint32_t w, uint32_t ring, uint32_t s0, uint32_t k,
                                  uint32_t end) {
  if (threadIdx.x != 0) return;
  PoolJob* J = &D.jobs[j];
  J->P.s0 = s0;
  J->P.k = k;
  J->ring = ring;
  J->end = end;
  PoolProg* p = &D.prog[j];
  p->gen_win = w;
  p->cursor = 0;
  p->win[ring] = w;
  p->retired[ring] = 0;
}

__global__ void k_pool_begin(PoolDev D, PoolPlan plan, int cur) {
  if (threadIdx.x != 0) return;
  *D.plan = plan;
  D.cnt[cur ^ 1] = 0;
  D.cnt[2] = 0;
  D.cnt[3] = 0;
  // generate appends to cnt[cur] while its workgroups run: they size their
  // loops (and so their barriers) from this copy, which nothing changes
  D.cnt[4] = D.cnt[cur];
}

// Paths plan entry e may hand out this iteration.
__device__ __forceinline__ uint32_t pool_entry_rem(const PoolDev& D, uint32_t e) {
  const uint32_t j = D.plan->order[e];
  const uint32_t rem = D.jobs[j].end - D.prog[j].cursor;
  return rem < D.plan->limit[e] ? rem : D.plan->limit[e];
}

// Paths the plan's jobs still have to hand out, and how many of them fit.
__device__ __forceinline__ uint32_t pool_gen_count(const PoolDev& D, uint32_t nplan, uint32_t live0) {
  uint32_t total = 0;
  for (uint32_t e = 0; e < nplan; ++e) total += pool_entry_rem(D, e);
  const uint32_t room = live0 < D.cap ? D.cap - live0 : 0u;
  return total < room ? total : room;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pool_generate(PoolDev D, int cur) {
  __shared__ uint32_t qsh[2][1][BLOCK / 64];
  __shared__ uint32_t dead[kPoolMaxJobs * 2];
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs * 2; i += BLOCK) dead[i] = 0;
  const uint32_t nplan = min(D.plan->n, kPoolMaxJobs);
  const uint32_t live0 = D.cnt[4];
  const uint32_t gen = pool_gen_count(D, nplan, live0);
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t rounds = (gen + nthreads - 1) / nthreads;
  const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
  if (gtid == 0) D.cnt[3] = gen;
  uint32_t* const ctr[1] = {&D.cnt[cur]};
  __syncthreads();
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t t = r * nthreads + gtid;
    bool live = false;
    uint32_t j = 0, ring = 0, p = 0;
    PathState s;
    if (t < gen) {
      // the plan entry that hands out place t
      uint32_t base = 0, e = 0, rem = 0;
      for (; e < nplan; ++e) {
        j = D.plan->order[e];
        rem = pool_entry_rem(D, e);
        if (t - base < rem) break;
        base += rem;
      }
      if (e < nplan) {
        const PoolJob* J = &D.jobs[j];
        p = D.prog[j].cursor + (t - base);
        ring = J->ring;
        const PtParams P = J->P;
        Rng rng;
        live = gen_path(P, p, s, rng) && P.max_depth > 0;
        if (!live) {
          as_global(J->Lout[ring])[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          atomicAdd(&dead[j * 2 + ring], 1u);
        }
      }
    }
    const bool fl[1] = {live};
    uint32_t slot[1];
    block_push<BLOCK, 1>(qsh, r & 1, ctr, fl, slot);
    if (live) pool_store(D.S[cur], slot[0], s, j | (ring << 16), p);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs * 2; i += BLOCK)
    if (dead[i]) atomicAdd(&D.prog[i >> 1].retired[i & 1], dead[i]);
}

// Moves the plan's cursors past what k_pool_generate handed out.
__global__ void k_pool_advance(PoolDev D) {
  if (threadIdx.x != 0) return;
  const uint32_t nplan = min(D.plan->n, kPoolMaxJobs);
  uint32_t left = D.cnt[3];
  for (uint32_t e = 0; e < nplan && left; ++e) {
    const uint32_t j = D.plan->order[e];
    const uint32_t rem = pool_entry_rem(D, e);
    const uint32_t take = rem < left ? rem : left;
    D.prog[j].cursor += take;
    left -= take;
  }
}

// At least 4 waves per SIMD, as k_wf_shade (two 512-thread workgroups per CU).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void k_pool_shade(ShadeView v, PoolDev D, int cur) {
  __shared__ uint32_t qsh[2][2][BLOCK / 64];
  __shared__ uint32_t ret[kPoolMaxJobs * 2];
  __shared__ uint32_t rays[kPoolMaxJobs];
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs * 2; i += BLOCK) ret[i] = 0;
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs; i += BLOCK) rays[i] = 0;
  const uint32_t n = D.cnt[cur];
  const PoolState Sc = D.S[cur], Sn = D.S[cur ^ 1];
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t rounds = (n + nthreads - 1) / nthreads;
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  uint32_t* const ctr[2] = {&D.cnt[2], &D.cnt[cur ^ 1]};
  __syncthreads();
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t i = r * nthreads + gtid;
    const bool active = i < n;
    bool cont = false;
    ShadowReq sh;
    sh.valid = false;
    PathState s;
    uint32_t jr = 0, p = 0;
    if (active) {
      pool_load(Sc, i, s, jr, p);
      const float4 hv = D.hit[i];
      WorldHit h;
      h.t = hv.x;
      h.pos = __float_as_uint(hv.y);
      h.u = hv.z;
      h.v = hv.w;
      h.inst = D.hit_inst[i];
      const PoolJob* J = &D.jobs[jr & 0xffffu];
      PtParams P{};  // pt_shade reads the depth limits
      P.max_depth = J->P.max_depth;
      P.rr_depth = J->P.rr_depth;
      Rng rng;
      rng.init(J->P.seed, s.pixel, s.sample, 0, s.dim);
      cont = pt_shade(v, P, s, h, rng, sh);
      s.dim = rng.dim;
      if (!cont) as_global(J->Lout[jr >> 16])[p] = make_float4(s.L.x, s.L.y, s.L.z, 0.0f);
    }
    const uint32_t j = jr & 0xffffu;
    wave_keyed_add(rays, active, j, sh.valid ? 2u : 1u);
    wave_keyed_add(ret, active && !cont, j * 2 + (jr >> 16), 1u);
    const bool fl[2] = {active && sh.valid, active && cont};
    uint32_t slot[2];
    block_push<BLOCK, 2>(qsh, r & 1, ctr, fl, slot);
    if (active && cont) pool_store(Sn, slot[1], s, jr, p);
    if (active && sh.valid) {
      const uint32_t si = slot[0];
      D.sh_o[si] = make_float4(sh.ray.o.x, sh.ray.o.y, sh.ray.o.z, sh.ray.tmin);
      D.sh_d[si] = make_float4(sh.ray.d.x, sh.ray.d.y, sh.ray.d.z, sh.ray.tmax);
      D.sh_c[si] = make_float4(sh.c.x, sh.c.y, sh.c.z, 0.0f);
      D.sh_tgt[si] = cont ? &Sn.L[slot[1]] : &D.jobs[j].Lout[jr >> 16][p];
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs * 2; i += BLOCK)
    if (ret[i]) atomicAdd(&D.prog[i >> 1].retired[i & 1], ret[i]);
  for (uint32_t i = threadIdx.x; i < kPoolMaxJobs; i += BLOCK)
    if (rays[i]) atomicAdd(&D.prog[i].rays, (unsigned long long)rays[i]);
}

// Shadow rays of the iteration: any hit; an unoccluded ray adds its
// contribution to the radiance its path holds now (one ray per path).
template <int BLOCK, int S, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_pool_shadow(SceneView sv, PoolDev D,
                                                                           int2* __restrict__ spill,
                                                                           uint32_t* __restrict__ work,
                                                                           uint32_t refill_min, uint32_t search_min) {
  __shared__ int2 lds[S * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK, S>(lds, spill, gtid, nthreads);
  const uint32_t n = D.cnt[2];
  walk_items<true, true, SLOTS>(
      sv, stk, n, gtid / 64, nthreads / 64, work, refill_min, search_min,
      [&](uint64_t i) { return make_ray4(D.sh_o[i], D.sh_d[i]); },
      [&](uint64_t i, const Walker<true, true>& w) {
        if (w.bpos == kNoHit) {
          AR_GLOBAL_AS float4* t = as_global(D.sh_tgt[i]);
          const float4 c = D.sh_c[i], l = *t;
          *t = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
        }
      });
  if (gtid == 0) atomicAdd(&D.stats[1], (unsigned long long)n);
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct HJob {
  bool used = false;
  uint32_t epoch = 0;
  uint64_t order = 0;
  JobSpec spec;
  PtParams P{};
  bool gdpt = false;
  uint32_t variants = 1, npix = 0, k = 1, nwin = 0;
  float4* Lout[2] = {nullptr, nullptr};
  float4* acc = nullptr;
  float* films = nullptr;
  float* out = nullptr;
  float* pinned = nullptr;
  Workspace* ws = nullptr;
  hipStream_t slot_stream = nullptr;
  hipEvent_t fin = nullptr, done = nullptr;
  uint32_t set_win = 0;   // windows set on the device
  uint32_t acc_win = 0;   // windows accumulated into the film
  bool ring_busy[2] = {false, false};
  uint32_t gen_win = 0, cursor = 0;  // last snapshot: window being generated and its cursor
  uint32_t retired[2] = {0, 0}, win[2] = {0xffffffffu, 0xffffffffu};
  unsigned long long rays = 0;
  bool cancel = false;
  uint64_t cancel_it = 0;  // first iteration issued without the job in its plan
  int state = 0;  // 0 rendering, 1 resolved, 2 cancelled and drained

  uint32_t win_samples(uint32_t w) const { return std::min<uint32_t>(k, spec.cfg.spp - w * k); }
  uint32_t win_paths(uint32_t w) const { return variants * npix * win_samples(w); }
  uint64_t total_paths() const { return (uint64_t)variants * npix * spec.cfg.spp; }
  // windows before the last are full (k samples each)
  uint64_t full_paths(uint32_t windows) const { return (uint64_t)windows * variants * npix * k; }
  uint64_t generated() const { return full_paths(std::min(gen_win, nwin)) + cursor; }
  bool generating() const { return !cancel && state == 0 && generated() < total_paths(); }
  // every path generated, some still in flight
  bool draining() const { return !cancel && state == 0 && generated() >= total_paths(); }
  uint64_t retired_total() const {
    uint64_t r = full_paths(std::min(acc_win, nwin - 1)) + (acc_win == nwin ? win_paths(nwin - 1) : 0);
    for (uint32_t w = acc_win; w < set_win; ++w)
      if (win[w & 1] == w) r += retired[w & 1];
    return r;
  }
  uint64_t live() const {
    const uint64_t g = generated(), r = retired_total();
    return g > r ? g - r : 0;
  }
};

class PoolImpl : public PathPool {
 public:
  const ar_scene* scene = nullptr;
  hipStream_t st = nullptr;
  uint32_t cap = 0;
  uint32_t fill = 0;  // live paths the other jobs may keep while a job drains
  DevBuf buf, spill_buf;
  size_t spill_words = 0;
  PoolDev D{};
  PoolProg* snaps = nullptr;  // pinned, kPoolRing x kPoolMaxJobs
  hipEvent_t ev[kPoolRing] = {};
  uint64_t issued = 0, retired_it = 0;
  uint32_t epoch_next = 1;
  HJob jobs[kPoolMaxJobs];
  ShadeView view{};

  ~PoolImpl() override {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& j : jobs) {
      if (j.fin) (void)hipEventDestroy(j.fin);
      if (j.done) (void)hipEventDestroy(j.done);
    }
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (snaps) (void)hipHostFree(snaps);
    if (st) (void)hipStreamDestroy(st);
  }

  ar_status init(const ar_scene* s, uint32_t c, uint32_t f) {
    scene = s;
    cap = c;
    fill = f;
    view = shade_view(s);
    AR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : ev) AR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& j : jobs) {
      AR_HIP(hipEventCreateWithFlags(&j.fin, hipEventDisableTiming));
      AR_HIP(hipEventCreateWithFlags(&j.done, hipEventDisableTiming));
    }
    AR_HIP(hipHostMalloc((void**)&snaps, sizeof(PoolProg) * kPoolMaxJobs * kPoolRing, hipHostMallocDefault));
    std::memset(snaps, 0, sizeof(PoolProg) * kPoolMaxJobs * kPoolRing);
    size_t bytes = 0;
    std::vector<std::pair<void**, size_t>> parts;
    auto add = [&](void** p, size_t n) { parts.push_back({p, n}); };
    for (int b = 0; b < 2; ++b) {
      PoolState& S = D.S[b];
      for (float4** f : {&S.ray_o, &S.ray_d, &S.beta, &S.L, &S.prev_p}) add((void**)f, (size_t)c * sizeof(float4));
      add((void**)&S.meta, (size_t)c * sizeof(uint4));
      add((void**)&S.job, (size_t)c * 4);
      add((void**)&S.pidx, (size_t)c * 4);
    }
    add((void**)&D.iota, (size_t)c * 4);
    add((void**)&D.hit, (size_t)c * sizeof(float4));
    add((void**)&D.hit_inst, (size_t)c * 4);
    add((void**)&D.sh_o, (size_t)c * sizeof(float4));
    add((void**)&D.sh_d, (size_t)c * sizeof(float4));
    add((void**)&D.sh_c, (size_t)c * sizeof(float4));
    add((void**)&D.sh_tgt, (size_t)c * sizeof(float4*));
    add((void**)&D.cnt, 64);
    add((void**)&D.stats, 64);
    add((void**)&D.work, kWalkWorkBytes);
    add((void**)&D.jobs, sizeof(PoolJob) * kPoolMaxJobs);
    add((void**)&D.prog, sizeof(PoolProg) * kPoolMaxJobs);
    add((void**)&D.plan, sizeof(PoolPlan));
    for (auto& pt : parts) bytes += align256(pt.second);
    AR_HIP(buf.alloc(bytes));
    size_t off = 0;
    for (auto& pt : parts) {
      *pt.first = (char*)buf.p + off;
      off += align256(pt.second);
    }
    D.cap = c;
    hipLaunchKernelGGL(k_pool_init, dim3(std::min<uint32_t>((c + 255) / 256, 4096)), dim3(256), 0, st, D);
    AR_HIP(hipGetLastError());
    return AR_OK;
  }

  // Traversal stack spill for the scene's current launch configurations.
  ar_status reserve_spill() {
    size_t words = 0;
    for (const ar_trace_config* tc : {&scene->tcfg, &scene->tcfg_any}) {
      const size_t gt = (size_t)(device_cus() * tc->blocks_per_cu / 8 * 8) * tc->workgroup_size;
      const size_t e = scene->stack_depth > tc->lds_stack ? scene->stack_depth - tc->lds_stack : 0;
      words = std::max(words, e * gt);
    }
    if (words > spill_words || !spill_buf.p) {
      AR_HIP(hipStreamSynchronize(st));
      AR_HIP(spill_buf.alloc(std::max<size_t>(words, 1) * sizeof(int2)));
      spill_words = std::max<size_t>(words, 1);
    }
    return AR_OK;
  }

  ar_status add(const JobSpec& spec, Workspace* ws, float* out_dev, float* pinned, hipStream_t slot_stream,
                uint64_t order, int* handle) override {
    int h = -1;
    for (int i = 0; i < (int)kPoolMaxJobs; ++i)
      if (!jobs[i].used) {
        h = i;
        break;
      }
    if (h < 0) return fail(AR_ERR_STATE, "the serving pool holds its maximum number of jobs");
    HJob& J = jobs[h];
    const ar_render_config& c = spec.cfg;
    J.spec = spec;
    J.order = order;
    J.gdpt = c.integrator == AR_INTEGRATOR_GDPT;
    J.variants = J.gdpt ? 3 : 1;
    J.npix = spec.width * spec.height;
    J.k = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(c.spp, kPoolWindowPaths / ((uint64_t)J.variants * J.npix)));
    J.nwin = (c.spp + J.k - 1) / J.k;
    const size_t ring_bytes = align256((size_t)J.variants * J.npix * J.k * sizeof(float4));
    const size_t acc_bytes = align256((size_t)J.npix * J.variants * sizeof(float4));
    const size_t film_bytes = J.gdpt ? align256((size_t)J.npix * 9 * sizeof(float)) : 0;
    const size_t need = 2 * ring_bytes + acc_bytes + film_bytes;
    if (ws->aux.bytes < need || (J.gdpt && ws->pois.bytes < poisson_scratch_bytes(spec.width, spec.height))) {
      AR_HIP(hipStreamSynchronize(slot_stream));
      AR_HIP(ws->aux.alloc(need));
      if (J.gdpt) AR_HIP(ws->pois.alloc(poisson_scratch_bytes(spec.width, spec.height)));
    }
    char* b = (char*)ws->aux.p;
    J.Lout[0] = (float4*)b;
    J.Lout[1] = (float4*)(b + ring_bytes);
    J.acc = (float4*)(b + 2 * ring_bytes);
    J.films = J.gdpt ? (float*)(b + 2 * ring_bytes + acc_bytes) : nullptr;
    J.out = out_dev;
    J.pinned = pinned;
    J.ws = ws;
    J.slot_stream = slot_stream;
    PtParams& P = J.P;
    P = PtParams{};
    P.width = spec.width;
    P.height = spec.height;
    P.spp = c.spp;
    P.max_depth = c.max_depth;
    P.rr_depth = c.rr_depth;
    P.npix = J.npix;
    P.seed = c.seed;
    P.cam = cam_from(spec.camera);
    P.variants = J.variants;
    P.s0 = 0;
    P.k = J.win_samples(0);
    PoolJob pj{};
    pj.P = P;
    pj.Lout[0] = J.Lout[0];
    pj.Lout[1] = J.Lout[1];
    pj.ring = 0;
    pj.end = J.win_paths(0);
    J.epoch = epoch_next++;
    // the slot's buffers were last used on its own stream (a previous job)
    hipEvent_t ready = J.fin;
    AR_HIP(hipEventRecord(ready, slot_stream));
    AR_HIP(hipStreamWaitEvent(st, ready, 0));
    hipLaunchKernelGGL(k_pool_set_job, dim3(1), dim3(64), 0, st, D, (uint32_t)h, pj, J.epoch);
    AR_HIP(hipGetLastError());
    J.used = true;
    J.set_win = 1;
    J.acc_win = 0;
    J.ring_busy[0] = true;
    J.ring_busy[1] = false;
    J.gen_win = 0;
    J.cursor = 0;
    J.retired[0] = J.retired[1] = 0;
    J.win[0] = 0;
    J.win[1] = 0xffffffffu;
    J.rays = 0;
    J.cancel = false;
    J.state = 0;
    *handle = h;
    return AR_OK;
  }

  void cancel(int h) override {
    if (jobs[h].cancel) return;
    jobs[h].cancel = true;
    jobs[h].cancel_it = issued;
  }

  // Generated paths of window w that can still be in flight, per the snapshot.
  uint32_t win_generated(const HJob& J, uint32_t w) const {
    if (w < J.gen_win) return J.win_paths(w);
    if (w == J.gen_win) return J.cursor;
    return 0;
  }

  ar_status on_snapshot(const PoolProg* snap) {
    for (uint32_t h = 0; h < kPoolMaxJobs; ++h) {
      HJob& J = jobs[h];
      if (!J.used || J.state != 0 || snap[h].epoch != J.epoch) continue;
      const PoolProg& p = snap[h];
      J.gen_win = p.gen_win;
      J.cursor = p.cursor;
      J.retired[0] = p.retired[0];
      J.retired[1] = p.retired[1];
      J.win[0] = p.win[0];
      J.win[1] = p.win[1];
      J.rays = p.rays;
    }
    return AR_OK;
  }

  // Accumulates completed windows, sets the next windows, finishes jobs.
  ar_status advance_jobs(bool* progress) {
    for (uint32_t h = 0; h < kPoolMaxJobs; ++h) {
      HJob& J = jobs[h];
      if (!J.used || J.state != 0) continue;
      if (J.cancel) {
        // Drained once every path generated for it has retired.  Iterations
        // issued before the cancel may still generate its paths, so only
        // snapshots taken after all of them count.
        if (retired_it < J.cancel_it) continue;
        bool drained = true;
        for (uint32_t w = J.acc_win; w < J.set_win; ++w) {
          const uint32_t r = w & 1;
          if (J.win[r] != w || J.retired[r] != win_generated(J, w)) drained = false;
        }
        if (drained) {
          J.state = 2;
          *progress = true;
        }
        continue;
      }
      while (J.acc_win < J.set_win) {
        const uint32_t w = J.acc_win, r = w & 1;
        if (J.win[r] != w || J.retired[r] != J.win_paths(w)) break;
        const uint32_t eg = std::min<uint32_t>((J.npix + 255) / 256, 4096);
        if (J.gdpt)
          hipLaunchKernelGGL(k_gdpt_accumulate, dim3(eg), dim3(256), 0, st, J.Lout[r], J.acc, J.P.width, J.P.height,
                             J.win_samples(w), w == 0 ? 1u : 0u);
        else
          hipLaunchKernelGGL(k_film_accumulate, dim3(eg), dim3(256), 0, st, J.Lout[r], J.acc, J.npix,
                             J.win_samples(w), w == 0 ? 1u : 0u);
        AR_HIP(hipGetLastError());
        J.ring_busy[r] = false;
        J.acc_win++;
        *progress = true;
      }
      // the window being generated is complete: start the next one in the other ring slot
      if (J.set_win < J.nwin && J.gen_win == J.set_win - 1 && J.cursor == J.win_paths(J.gen_win)) {
        const uint32_t w = J.set_win, r = w & 1;
        if (!J.ring_busy[r]) {
          hipLaunchKernelGGL(k_pool_set_window, dim3(1), dim3(64), 0, st, D, h, w, r, w * J.k, J.win_samples(w),
                             J.win_paths(w));
          AR_HIP(hipGetLastError());
          J.ring_busy[r] = true;
          J.set_win++;
          *progress = true;
        }
      }
      if (J.acc_win == J.nwin) {
        // every window is in the film: resolve on the job's own stream
        AR_HIP(hipEventRecord(J.fin, st));
        AR_HIP(hipStreamWaitEvent(J.slot_stream, J.fin, 0));
        const uint32_t eg = std::min<uint32_t>((J.npix + 255) / 256, 4096);
        if (!J.gdpt) {
          hipLaunchKernelGGL(k_film_resolve, dim3(eg), dim3(256), 0, J.slot_stream, J.acc, J.out, J.npix,
                             (float)J.spec.cfg.spp);
          AR_HIP(hipGetLastError());
        } else {
          hipLaunchKernelGGL(k_gdpt_resolve, dim3(eg), dim3(256), 0, J.slot_stream, J.acc, J.films, J.npix,
                             1.0f / (float)J.spec.cfg.spp);
          AR_HIP(hipGetLastError());
          const ar_render_config& c = J.spec.cfg;
          ar_status e = poisson_solve_device(J.films, J.films + (size_t)J.npix * 3, J.films + (size_t)J.npix * 6,
                                             J.P.width, J.P.height, c.alpha, c.poisson_iterations, c.poisson_sign,
                                             J.out, nullptr, J.ws->pois, J.slot_stream);
          if (e != AR_OK) return e;
        }
        AR_HIP(hipMemcpyAsync(J.pinned, J.out, (size_t)J.npix * 3 * sizeof(float), hipMemcpyDeviceToHost,
                              J.slot_stream));
        AR_HIP(hipEventRecord(J.done, J.slot_stream));
        J.state = 1;
        *progress = true;
      }
    }
    return AR_OK;
  }

  template <class F>
  static bool pick_variant(const ar_trace_config& c, F&& f) {
    bool ok = false;
#define AR_PICK(BB, SS)                  \
  if (c.workgroup_size == BB && c.lds_stack == SS) { \
    f(std::integral_constant<int, BB>(), std::integral_constant<int, SS>()); \
    ok = true;                           \
  }
    AR_TRACE_VARIANTS(AR_PICK)
#undef AR_PICK
    return ok;
  }

  ar_status issue_iteration() {
    AR_TRY(reserve_spill());
    PoolPlan plan{};
    std::vector<uint32_t> gen;
    for (uint32_t h = 0; h < kPoolMaxJobs; ++h)
      if (jobs[h].used && jobs[h].generating()) gen.push_back(h);
    std::sort(gen.begin(), gen.end(), [&](uint32_t a, uint32_t b) {
      const uint64_t ra = jobs[a].total_paths() - jobs[a].generated();
      const uint64_t rb = jobs[b].total_paths() - jobs[b].generated();
      return ra != rb ? ra < rb : jobs[a].order < jobs[b].order;
    });
    plan.n = (uint32_t)gen.size();
    // While a job drains its last paths (the job with the least remaining
    // work), the others may keep at most `fill` paths alive: its last
    // bounces, which alone leave most of the device idle, run beside a
    // bounded amount of their work instead of behind a full pool.
    bool drain = false;
    uint64_t live_gen = 0;
    for (const HJob& J : jobs)
      if (J.used && J.state == 0 && !J.cancel) {
        if (J.draining())
          drain = true;
        else
          live_gen += J.live();
      }
    uint64_t budget = drain ? (fill > live_gen ? fill - live_gen : 0) : ~0ull;
    for (size_t e = 0; e < gen.size(); ++e) {
      plan.order[e] = gen[e];
      const uint64_t want = jobs[gen[e]].total_paths() - jobs[gen[e]].generated();
      const uint64_t lim = std::min<uint64_t>(std::min(want, budget), 0xffffffffull);
      plan.limit[e] = (uint32_t)lim;
      if (drain) budget -= lim;
    }
    const int cur = (int)(issued & 1);
    const ar_trace_config tc = scene->tcfg, ta = scene->tcfg_any;
    const uint32_t Gp = (uint32_t)(device_cus() * kTraceBlocksPerCU * kTraceBlock / kPoolBlock / 8 * 8);
    WaveBufs B{};
    B.ray_o = D.S[cur].ray_o;
    B.ray_d = D.S[cur].ray_d;
    B.q[cur] = D.iota;
    B.hit = D.hit;
    B.hit_inst = D.hit_inst;
    B.cnt = D.cnt;
    B.stats = D.stats;
    B.work = D.work;
    int2* spill = spill_buf.as<int2>();
    hipLaunchKernelGGL(k_pool_begin, dim3(1), dim3(64), 0, st, D, plan, cur);
    hipLaunchKernelGGL((k_pool_generate<kPoolBlock>), dim3(Gp), dim3(kPoolBlock), 0, st, D, cur);
    hipLaunchKernelGGL(k_pool_advance, dim3(1), dim3(64), 0, st, D);
    const bool ok_e = pick_variant(tc, [&](auto bb, auto ss) {
      constexpr int BB = decltype(bb)::value, SS = decltype(ss)::value;
      const void* kfn = tc.park_slots == 4 ? (const void*)&k_wf_extend<BB, SS, 4> : (const void*)&k_wf_extend<BB, SS, 3>;
      const uint32_t G = clamp_resident((uint32_t)(device_cus() * tc.blocks_per_cu / 8 * 8), kfn, BB);
      if (tc.park_slots == 4)
        hipLaunchKernelGGL((k_wf_extend<BB, SS, 4>), dim3(G), dim3(BB), 0, st, view.sv, B, cur, spill,
                           tc.balance ? D.work : nullptr, tc.refill_min, tc.search_min);
      else
        hipLaunchKernelGGL((k_wf_extend<BB, SS, 3>), dim3(G), dim3(BB), 0, st, view.sv, B, cur, spill,
                           tc.balance ? D.work : nullptr, tc.refill_min, tc.search_min);
    });
    hipLaunchKernelGGL((k_pool_shade<kPoolBlock>), dim3(Gp), dim3(kPoolBlock), 0, st, view, D, cur);
    const bool ok_s = pick_variant(ta, [&](auto bb, auto ss) {
      constexpr int BB = decltype(bb)::value, SS = decltype(ss)::value;
      const void* kfn =
          ta.park_slots == 4 ? (const void*)&k_pool_shadow<BB, SS, 4> : (const void*)&k_pool_shadow<BB, SS, 3>;
      const uint32_t G = clamp_resident((uint32_t)(device_cus() * ta.blocks_per_cu / 8 * 8), kfn, BB);
      if (ta.park_slots == 4)
        hipLaunchKernelGGL((k_pool_shadow<BB, SS, 4>), dim3(G), dim3(BB), 0, st, view.sv, D, spill,
                           ta.balance ? D.work : nullptr, ta.refill_min, ta.search_min);
      else
        hipLaunchKernelGGL((k_pool_shadow<BB, SS, 3>), dim3(G), dim3(BB), 0, st, view.sv, D, spill,
                           ta.balance ? D.work : nullptr, ta.refill_min, ta.search_min);
    });
    if (!ok_e || !ok_s) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
    AR_HIP(hipGetLastError());
    const uint32_t slot = (uint32_t)(issued % kPoolRing);
    AR_HIP(hipMemcpyAsync(snaps + (size_t)slot * kPoolMaxJobs, D.prog, sizeof(PoolProg) * kPoolMaxJobs,
                          hipMemcpyDeviceToHost, st));
    AR_HIP(hipEventRecord(ev[slot], st));
    issued++;
    return AR_OK;
  }

  bool has_work() const {
    for (const HJob& J : jobs)
      if (J.used && J.state == 0) return true;
    return false;
  }

  ar_status pump(bool issue, bool* progress) override {
    while (retired_it < issued) {
      const uint32_t slot = (uint32_t)(retired_it % kPoolRing);
      const hipError_t q = hipEventQuery(ev[slot]);
      if (q == hipErrorNotReady) break;
      AR_HIP(q);
      on_snapshot(snaps + (size_t)slot * kPoolMaxJobs);
      retired_it++;
      *progress = true;
    }
    AR_TRY(advance_jobs(progress));
    while (issue && has_work() && issued - retired_it < kPoolInFlight) {
      AR_TRY(issue_iteration());
      *progress = true;
    }
    return AR_OK;
  }

  int state(int h) override { return jobs[h].state; }
  hipEvent_t done_event(int h) override { return jobs[h].done; }
  void stats(int h, ar_render_stats* s) override {
    const HJob& J = jobs[h];
    s->paths = J.total_paths();
    s->rays = J.rays;
    s->shadow_rays = 0;
  }
  void release(int h) override { jobs[h].used = false; }
  uint64_t pending_paths() override {
    uint64_t n = 0;
    for (const HJob& J : jobs)
      if (J.used && J.generating()) n += J.total_paths() - J.generated();
    return n;
  }
  uint64_t remaining_paths(int h) override {
    const HJob& J = jobs[h];
    return J.total_paths() - std::min<uint64_t>(J.total_paths(), J.generated());
  }
  bool idle() override { return !has_work(); }
  bool full() override {
    for (const HJob& J : jobs)
      if (!J.used) return false;
    return true;
  }
  uint32_t in_flight() override { return (uint32_t)(issued - retired_it); }
  hipStream_t stream() override { return st; }
};

}  // namespace

size_t PathPool::footprint(uint32_t capacity) {
  const size_t per = 2 * (5 * 16 + 16 + 4 + 4) + 4 + 16 + 4 + 3 * 16 + 8;
  return (size_t)capacity * per + (64ull << 20);  // + stack spill and tables
}

size_t PathPool::job_bytes(const JobSpec& spec) {
  const bool gdpt = spec.cfg.integrator == AR_INTEGRATOR_GDPT;
  const uint64_t variants = gdpt ? 3 : 1, npix = (uint64_t)spec.width * spec.height;
  const uint64_t k = std::max<uint64_t>(1, std::min<uint64_t>(spec.cfg.spp, kPoolWindowPaths / (variants * npix)));
  size_t b = 2 * align256(variants * npix * k * 16) + align256(npix * variants * 16) + npix * 12;
  if (gdpt) b += align256(npix * 36) + poisson_scratch_bytes(spec.width, spec.height);
  return b;
}

ar_status PathPool::create(const ar_scene* s, uint32_t capacity, uint32_t fill, std::unique_ptr<PathPool>* out) {
  std::unique_ptr<PoolImpl> p(new PoolImpl());
  ar_status e = p->init(s, capacity, fill);
  if (e != AR_OK) return e;
  *out = std::move(p);
  return AR_OK;
}

}  // namespace ar
