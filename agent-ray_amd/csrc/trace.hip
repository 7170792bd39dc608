// Batch traversal kernels behind ar_trace / ar_occluded
// (TLAS::intersect, agent-ray src/accel/bvh.rs:413-447; Scene::occluded,
// src/scene.rs:166-168).
//
// One ray per lane, persistent grid: each wave walks chunks of 64 rays, a
// static share (every W-th chunk, W = waves in the grid) and then chunks
// taken from a counter shared by the launch, and its lanes take the wave's
// rays in order as they finish (walk_items), so a wave never waits for its
// slowest ray and the waves of a launch end together.  The
// grid is a multiple of 8 and block ids are remapped so that each XCD
// (blocks are dispatched round-robin over the 8 XCDs) sweeps a contiguous
// range of rays, keeping neighbouring (coherent) rays in the same L2.
#include "ar_internal.h"
#include "ar_traverse.h"
#include "ar_kernels.h"

#include <map>

namespace ar {

__device__ __forceinline__ Ray load_ray(const ar_ray* rays, uint64_t i) {
  const float4* p = reinterpret_cast<const float4*>(rays + i);
  float4 a = p[0], b = p[1];
  Ray r;
  r.o = mk(a.x, a.y, a.z);
  r.tmin = a.w;
  r.d = mk(b.x, b.y, b.z);
  r.tmax = b.w;
  return r;
}

template <bool ANY, bool NEAR, int S, int BLOCK, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_trace(SceneView sv, const ar_ray* __restrict__ rays, uint64_t n,
                                                 ar_hit* __restrict__ hits, uint32_t* __restrict__ inst_ids,
                                                 uint32_t* __restrict__ occ, int2* __restrict__ spill,
                                                 uint32_t* __restrict__ work, uint32_t refill_min,
                                                 uint32_t search_min) {
  __shared__ int2 lds[S * BLOCK];
  const uint32_t G = gridDim.x;
  const uint32_t b = blockIdx.x;
  const uint32_t lb = (G % 8 == 0) ? (b % 8) * (G / 8) + (b / 8) : b;
  const uint32_t gtid = lb * BLOCK + threadIdx.x;
  Stack stk;
  stk.lds = (LdsEntryPtr)(lds + threadIdx.x);
  stk.stride = BLOCK;
  stk.spill = (SpillEntryPtr)(spill + gtid);
  stk.spill_stride = G * BLOCK;
  stk.S = S;
  stk.sp = 0;
  const uint64_t waves = (uint64_t)G * (BLOCK / 64);
  const uint64_t wave = (uint64_t)lb * (BLOCK / 64) + threadIdx.x / 64;
  walk_items<ANY, NEAR, SLOTS>(
      sv, stk, n, wave, waves, work, refill_min, search_min, [&](uint64_t i) { return load_ray(rays, i); },
      [&](uint64_t i, const Walker<ANY, NEAR>& w) {
        if (ANY) {
          occ[i] = w.bpos != kNoHit ? 1u : 0u;
        } else {
          float t = INFINITY;
          uint32_t prim = kNoHit;
          if (w.bpos != kNoHit) {
            t = w.bt;
            prim = __float_as_uint(sv.tris[w.bpos].v0.w) & ~kLastBit;
          }
          *reinterpret_cast<float2*>(hits + i) = make_float2(t, __uint_as_float(prim));
          if (inst_ids) inst_ids[i] = w.bpos != kNoHit ? (sv.insts[w.binst].id & ~kLastBit) : kNoHit;
        }
      });
}

// Stack entries one query can need (bounded by the tree depths at commit).
static uint32_t stack_need(const ar_scene* s) { return s->stack_depth; }

template <bool ANY, bool NEAR, int BLOCK, int S, uint32_t SLOTS>
static void launch_slots(const ar_scene* s, uint32_t G, const ar_ray* rays, uint64_t n, ar_hit* hits,
                         uint32_t* inst, uint32_t* occ, int2* spill, uint32_t* work, hipStream_t st,
                         const ar_trace_config& c) {
  G = clamp_resident(G, reinterpret_cast<const void*>(&k_trace<ANY, NEAR, S, BLOCK, SLOTS>), BLOCK);
  hipLaunchKernelGGL((k_trace<ANY, NEAR, S, BLOCK, SLOTS>), dim3(G), dim3(BLOCK), 0, st, s->view, rays, n, hits,
                     inst, occ, spill, work, c.refill_min, c.search_min);
}

template <bool ANY, bool NEAR, int BLOCK, int S>
static void launch_variant(const ar_scene* s, uint32_t G, const ar_ray* rays, uint64_t n, ar_hit* hits,
                           uint32_t* inst, uint32_t* occ, int2* spill, uint32_t* work, hipStream_t st,
                           const ar_trace_config& c) {
  if (c.park_slots == 4)
    launch_slots<ANY, NEAR, BLOCK, S, 4>(s, G, rays, n, hits, inst, occ, spill, work, st, c);
  else
    launch_slots<ANY, NEAR, BLOCK, S, 3>(s, G, rays, n, hits, inst, occ, spill, work, st, c);
}

// Launches over n rays in pieces of < 2^31 (32-bit ray positions inside a
// wave) with grid G, the given spill area and walker counters (the pieces
// run one after another on st and share them); false for an unsupported
// variant.
template <bool ANY, bool NEAR>
static bool launch_pieces(const ar_scene* s, uint32_t G, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst,
                          uint32_t* occ, int2* spill, uint32_t* work, hipStream_t st, const ar_trace_config& c) {
  bool ok = false;
  const uint64_t kMaxLaunch = 1ull << 31;
  for (uint64_t off = 0; off < n; off += kMaxLaunch) {
    const uint64_t m = n - off < kMaxLaunch ? n - off : kMaxLaunch;
    const ar_ray* r = rays + off;
    ar_hit* h = hits ? hits + off : nullptr;
    uint32_t* in = inst ? inst + off : nullptr;
    uint32_t* oc = occ ? occ + off : nullptr;
    ok = false;
#define AR_LAUNCH(B, SS)                                                           \
  if (c.workgroup_size == B && c.lds_stack == SS) {                                \
    launch_variant<ANY, NEAR, B, SS>(s, G, r, m, h, in, oc, spill, work, st, c);      \
    ok = true;                                                                     \
  }
    AR_TRACE_VARIANTS(AR_LAUNCH)
#undef AR_LAUNCH
    if (!ok) break;
  }
  return ok;
}

template <bool ANY, bool NEAR>
static ar_status launch_trace(const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst,
                              uint32_t* occ, hipStream_t st, const ar_trace_config& c) {
  uint32_t G = trace_grid_cfg(n, c);
  uint32_t need = stack_need(s);
  size_t bytes = need > c.lds_stack ? (size_t)(need - c.lds_stack) * G * c.workgroup_size * sizeof(int2) : 0;
  ar_scene* ms = const_cast<ar_scene*>(s);
  char* base = nullptr;
  ar_status e = scratch_acquire(ms, kWalkWorkBytes + bytes, st, (void**)&base);
  if (e != AR_OK) return e;
  uint32_t* work = (uint32_t*)base;
  int2* spill = (int2*)(base + kWalkWorkBytes);
  const bool ok = launch_pieces<ANY, NEAR>(s, G, rays, n, hits, inst, occ, spill, c.balance ? work : nullptr, st, c);
  hipError_t le = hipGetLastError();
  scratch_release(ms, st);
  if (!ok) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
  if (le != hipSuccess) return fail(AR_ERR_DEVICE, std::string("k_trace launch: ") + hipGetErrorString(le));
  return AR_OK;
}

// The grid a launch of ar_trace / ar_occluded uses for n rays under c (its
// kernel variant's resident-block clamp), 0 for an unsupported variant.
uint32_t trace_launch_grid(uint64_t n, const ar_trace_config& c, bool any, bool near) {
  const uint32_t G = trace_grid_cfg(n, c);
  uint32_t out = 0;
#define AR_GRID_OF(B, SS)                                                                                     \
  if (c.workgroup_size == B && c.lds_stack == SS) {                                                          \
    const void* k3 = any ? (near ? (const void*)&k_trace<true, true, SS, B, 3> : (const void*)&k_trace<true, false, SS, B, 3>) \
                         : (near ? (const void*)&k_trace<false, true, SS, B, 3> : (const void*)&k_trace<false, false, SS, B, 3>); \
    const void* k4 = any ? (near ? (const void*)&k_trace<true, true, SS, B, 4> : (const void*)&k_trace<true, false, SS, B, 4>) \
                         : (near ? (const void*)&k_trace<false, true, SS, B, 4> : (const void*)&k_trace<false, false, SS, B, 4>); \
    out = clamp_resident(G, c.park_slots == 4 ? k4 : k3, B);                                                 \
  }
  AR_TRACE_VARIANTS(AR_GRID_OF)
#undef AR_GRID_OF
  return out;
}

bool trace_variant_supported(uint32_t block, uint32_t stack) {
  bool ok = false;
#define AR_CHECK(B, SS) ok |= (block == B && stack == SS);
  AR_TRACE_VARIANTS(AR_CHECK)
#undef AR_CHECK
  return ok;
}

uint32_t clamp_resident(uint32_t G, const void* kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, uint32_t> cache;
  uint32_t nb;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({kernel, block});
    if (it == cache.end()) {
      int v = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, block, 0) != hipSuccess || v <= 0) v = 1;
      it = cache.emplace(std::make_pair(kernel, block), (uint32_t)v).first;
    }
    nb = it->second;
  }
  uint32_t cap = (uint32_t)device_cus() * nb / 8 * 8;
  if (cap == 0) cap = 8;
  return G < cap ? G : cap;
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

uint32_t trace_grid_cfg(uint64_t n, const ar_trace_config& c) {
  uint64_t want = (n + c.workgroup_size - 1) / c.workgroup_size;
  uint64_t cap = (uint64_t)device_cus() * c.blocks_per_cu;
  uint64_t g = want < cap ? want : cap;
  g = (g + 7) / 8 * 8;
  if (g == 0) g = 8;
  return (uint32_t)g;
}

__global__ void k_clear_walk_work(uint32_t* p) { p[threadIdx.x] = 0u; }

hipError_t clear_walk_work(void* p, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_walk_work, dim3(1), dim3(kWalkWorkBytes / sizeof(uint32_t)), 0, st, (uint32_t*)p);
  return hipGetLastError();
}

uint32_t trace_grid(uint64_t n, int block) {
  ar_trace_config c{(uint32_t)block, (uint32_t)kTraceLdsStack, (uint32_t)kTraceBlocksPerCU, kTraceRefillMin, 8, 3, 0, 0.0f};
  return trace_grid_cfg(n, c);
}

ar_status scratch_acquire(ar_scene* s, size_t bytes, hipStream_t st, void** out) {
  s->mu.lock();
  if (!s->scratch_event) {
    hipError_t e = hipEventCreateWithFlags(&s->scratch_event, hipEventDisableTiming);
    if (e != hipSuccess) {
      s->mu.unlock();
      return fail(AR_ERR_DEVICE, "hipEventCreate failed");
    }
  } else {
    hipError_t e = hipStreamWaitEvent(st, s->scratch_event, 0);
    if (e != hipSuccess) {
      s->mu.unlock();
      return fail(AR_ERR_DEVICE, "hipStreamWaitEvent failed");
    }
  }
  if (bytes > 0 && s->d_scratch.bytes < bytes) {
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = s->d_scratch.alloc(bytes);
    if (e == hipSuccess) e = clear_walk_work(s->d_scratch.p, st);
    if (e != hipSuccess) {
      s->mu.unlock();
      return fail(AR_ERR_DEVICE, std::string("scratch alloc: ") + hipGetErrorString(e));
    }
  }
  *out = s->d_scratch.p;
  return AR_OK;
}

void scratch_release(ar_scene* s, hipStream_t st) {
  (void)hipEventRecord(s->scratch_event, st);
  s->mu.unlock();
}

// ---------------------------------------------------------------------------
// Allocation-free tracing (ar_trace_noalloc).  The walker counters and the
// traversal stack spill of such calls live in `scratch`, a fixed array of
// kNoallocSpillBytes in every GPU's memory for the life of the process: the
// counters in its first kWalkWorkBytes (zero when the code object is loaded,
// and set back to zero by every launch's last wave), the spill after them.
// Its device address is looked up once per device (hipGetSymbolAddress, at
// ar_init or at the first call on that device) and only that address is
// handed to kernels: the array's name in host code is a host-side stand-in,
// not a device address.  The spill needs no clearing (every entry is
// written before it is read).  Calls that overlap on one GPU share the
// array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocSpillBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocSpillBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

ar_status noalloc_setup_current_device() {
  int dev = 0;
  AR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(AR_ERR_DEVICE, "device index beyond the allocation-free table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return AR_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(AR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_noalloc_ptr[dev] = p;
  return AR_OK;
}

static void* noalloc_scratch_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc_ptr[dev];
}

}  // namespace ar

using namespace ar;

namespace {

// Stages host buffers when AR_HOST_PTRS is set.
struct Staged {
  DevBuf rays, out, out2;
};

}  // namespace

extern "C" {

ar_status ar_trace(const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst_ids,
                   uint32_t flags, void* stream) {
  if (!s || (!rays && n) || (!hits && n)) return fail(AR_ERR_INVALID_ARG, "ar_trace: NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_trace: scene not committed");
  if (n == 0) return AR_OK;
  hipStream_t st = pick_stream(s, stream);
  const ar_ray* d_rays = rays;
  ar_hit* d_hits = hits;
  uint32_t* d_inst = inst_ids;
  Staged sg;
  bool host = flags & AR_HOST_PTRS;
  if (host) {
    AR_HIP(sg.rays.alloc(n * sizeof(ar_ray)));
    AR_HIP(sg.out.alloc(n * sizeof(ar_hit)));
    if (inst_ids) AR_HIP(sg.out2.alloc(n * sizeof(uint32_t)));
    AR_HIP(hipMemcpyAsync(sg.rays.p, rays, n * sizeof(ar_ray), hipMemcpyHostToDevice, st));
    d_rays = sg.rays.as<ar_ray>();
    d_hits = sg.out.as<ar_hit>();
    d_inst = inst_ids ? sg.out2.as<uint32_t>() : nullptr;
  }
  ar_status e = (flags & AR_REF_ORDER) ? launch_trace<false, false>(s, d_rays, n, d_hits, d_inst, nullptr, st, s->tcfg)
                                       : launch_trace<false, true>(s, d_rays, n, d_hits, d_inst, nullptr, st, s->tcfg);
  if (e != AR_OK) return e;
  if (host) {
    AR_HIP(hipMemcpyAsync(hits, d_hits, n * sizeof(ar_hit), hipMemcpyDeviceToHost, st));
    if (inst_ids) AR_HIP(hipMemcpyAsync(inst_ids, d_inst, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  }
  if (host || !(flags & AR_ASYNC)) AR_HIP(hipStreamSynchronize(st));
  return AR_OK;
}

ar_status ar_trace_noalloc(const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst_ids,
                           uint32_t flags, void* stream) {
  if (!s || (!rays && n) || (!hits && n)) return fail(AR_ERR_INVALID_ARG, "ar_trace_noalloc: NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_trace_noalloc: scene not committed");
  if (flags & AR_HOST_PTRS)
    return fail(AR_ERR_INVALID_ARG, "ar_trace_noalloc: device pointers only (staging host buffers would allocate)");
  if (n == 0) return AR_OK;
  ar_status e = noalloc_setup_current_device();
  if (e != AR_OK) return e;
  char* base = static_cast<char*>(noalloc_scratch_current_device());
  if (!base) return fail(AR_ERR_DEVICE, "ar_trace_noalloc: no scratch address for the current device");
  uint32_t* work = (uint32_t*)base;
  int2* spill = (int2*)(base + kWalkWorkBytes);
  constexpr size_t kSpillBytes = kNoallocSpillBytes - kWalkWorkBytes;
  const ar_trace_config c = s->tcfg;
  const uint32_t need = stack_need(s);
  const uint64_t per_thread = need > c.lds_stack ? (uint64_t)(need - c.lds_stack) * sizeof(int2) : 0;
  uint32_t G = trace_grid_cfg(n, c);
  if (per_thread) {
    // the spill is [entry][thread]: fit the grid to the fixed array (the
    // answer does not depend on the grid)
    const uint64_t gmax = kSpillBytes / (per_thread * c.workgroup_size) / 8 * 8;
    if (gmax < 8)
      return fail(AR_ERR_OUT_OF_BUDGET,
                  "ar_trace_noalloc: this scene's stack spill needs " + std::to_string(per_thread * c.workgroup_size * 8) +
                      " bytes for the smallest grid; the allocation-free array holds " +
                      std::to_string(kSpillBytes));
    G = (uint32_t)std::min<uint64_t>(G, gmax);
  }
  hipStream_t st = pick_stream(s, stream);
  const bool ok = (flags & AR_REF_ORDER)
                      ? launch_pieces<false, false>(s, G, rays, n, hits, inst_ids, nullptr, spill,
                                                    c.balance ? work : nullptr, st, c)
                      : launch_pieces<false, true>(s, G, rays, n, hits, inst_ids, nullptr, spill,
                                                   c.balance ? work : nullptr, st, c);
  if (!ok) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) return fail(AR_ERR_DEVICE, std::string("k_trace launch: ") + hipGetErrorString(le));
  if (!(flags & AR_ASYNC)) AR_HIP(hipStreamSynchronize(st));
  return AR_OK;
}

ar_status ar_occluded(const ar_scene* s, const ar_ray* rays, uint64_t n, uint32_t* occluded, uint32_t flags,
                      void* stream) {
  if (!s || (!rays && n) || (!occluded && n)) return fail(AR_ERR_INVALID_ARG, "ar_occluded: NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_occluded: scene not committed");
  if (n == 0) return AR_OK;
  hipStream_t st = pick_stream(s, stream);
  const ar_ray* d_rays = rays;
  uint32_t* d_occ = occluded;
  Staged sg;
  bool host = flags & AR_HOST_PTRS;
  if (host) {
    AR_HIP(sg.rays.alloc(n * sizeof(ar_ray)));
    AR_HIP(sg.out.alloc(n * sizeof(uint32_t)));
    AR_HIP(hipMemcpyAsync(sg.rays.p, rays, n * sizeof(ar_ray), hipMemcpyHostToDevice, st));
    d_rays = sg.rays.as<ar_ray>();
    d_occ = sg.out.as<uint32_t>();
  }
  ar_status e = (flags & AR_REF_ORDER) ? launch_trace<true, false>(s, d_rays, n, nullptr, nullptr, d_occ, st, s->tcfg_any)
                                       : launch_trace<true, true>(s, d_rays, n, nullptr, nullptr, d_occ, st, s->tcfg_any);
  if (e != AR_OK) return e;
  if (host) AR_HIP(hipMemcpyAsync(occluded, d_occ, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  if (host || !(flags & AR_ASYNC)) AR_HIP(hipStreamSynchronize(st));
  return AR_OK;
}

}  // extern "C"

extern "C" {

static ar_status check_trace_config(const ar_scene* s, const ar_trace_config* c, const char* who) {
  if (!s || !c) return fail(AR_ERR_INVALID_ARG, std::string(who) + ": NULL argument");
  if (!trace_variant_supported(c->workgroup_size, c->lds_stack))
    return fail(AR_ERR_UNSUPPORTED_TYPE, "no traversal variant for this workgroup size / stack depth");
  if (c->blocks_per_cu == 0 || c->blocks_per_cu > 64) return fail(AR_ERR_INVALID_ARG, "blocks_per_cu out of range");
  if (c->refill_min == 0 || c->refill_min > 64) return fail(AR_ERR_INVALID_ARG, "refill_min out of range (1..64)");
  if (c->search_min > 63) return fail(AR_ERR_INVALID_ARG, "search_min out of range (0..63)");
  if (c->park_slots != 3 && c->park_slots != 4) return fail(AR_ERR_INVALID_ARG, "park_slots must be 3 or 4");
  if (c->balance > 1) return fail(AR_ERR_INVALID_ARG, "balance must be 0 or 1");
  return AR_OK;
}

ar_status ar_scene_set_trace_config(ar_scene* s, const ar_trace_config* c) {
  ar_status e = check_trace_config(s, c, "ar_scene_set_trace_config");
  if (e != AR_OK) return e;
  std::lock_guard<std::mutex> lk(s->mu);
  s->tcfg = *c;
  s->tcfg.ms = 0.0f;
  s->tcfg_any = s->tcfg;
  s->tcfg_set = true;
  return AR_OK;
}

ar_status ar_scene_set_anyhit_config(ar_scene* s, const ar_trace_config* c) {
  ar_status e = check_trace_config(s, c, "ar_scene_set_anyhit_config");
  if (e != AR_OK) return e;
  std::lock_guard<std::mutex> lk(s->mu);
  s->tcfg_any = *c;
  s->tcfg_any.ms = 0.0f;
  s->tcfg_set = true;
  return AR_OK;
}

ar_status ar_scene_get_trace_config(const ar_scene* s, ar_trace_config* c) {
  if (!s || !c) return fail(AR_ERR_INVALID_ARG, "ar_scene_get_trace_config: NULL argument");
  *c = s->tcfg;
  return AR_OK;
}

ar_status ar_scene_get_anyhit_config(const ar_scene* s, ar_trace_config* c) {
  if (!s || !c) return fail(AR_ERR_INVALID_ARG, "ar_scene_get_anyhit_config: NULL argument");
  *c = s->tcfg_any;
  return AR_OK;
}

// Times every (variant, blocks per CU, refill threshold, search quorum, park
// slots) candidate on the sample, near-first order, one warm-up launch each,
// once for closest hit and once for any hit (their best launches differ), and
// keeps the fastest of each.  `balance` keeps its current value: its effect
// shows on launches of at least 4 chunks of 64 rays per wave, which a ray
// sample is usually too small for (the job-level tuner times it).
ar_status ar_scene_autotune(ar_scene* s, const ar_ray* rays, uint64_t n, uint32_t flags, void* stream,
                            ar_trace_config* best) {
  if (!s || !rays || n == 0) return fail(AR_ERR_INVALID_ARG, "ar_scene_autotune: NULL or empty sample");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_scene_autotune: scene not committed");
  if (flags & AR_HOST_PTRS) return fail(AR_ERR_INVALID_ARG, "ar_scene_autotune: sample must be on the device");
  hipStream_t st = pick_stream(s, stream);
  DevBuf hits, occ;
  AR_HIP(hits.alloc(n * sizeof(ar_hit)));
  AR_HIP(occ.alloc(n * sizeof(uint32_t)));
  hipEvent_t e0, e1;
  AR_HIP(hipEventCreate(&e0));
  AR_HIP(hipEventCreate(&e1));
  ar_trace_config winners[2] = {s->tcfg, s->tcfg_any};
  const uint32_t balance[2] = {s->tcfg.balance, s->tcfg_any.balance};
  float best_ms[2] = {INFINITY, INFINITY};
  const uint32_t per_cu[3] = {4, 8, 16};
  const uint32_t refills[3] = {8, 24, 48};
  const uint32_t searches[3] = {8, 16, 32};
  const uint32_t slot_counts[2] = {3, 4};
  const uint32_t variants[][2] = {
#define AR_LIST(B, SS) {B, SS},
      AR_TRACE_VARIANTS(AR_LIST)
#undef AR_LIST
  };
  ar_status err = AR_OK;
  // with neither flag both are tuned on the sample; with one, only that one
  const bool only_c = (flags & AR_TUNE_CLOSEST) && !(flags & AR_TUNE_ANYHIT);
  const bool only_a = (flags & AR_TUNE_ANYHIT) && !(flags & AR_TUNE_CLOSEST);
  for (int q = only_a ? 1 : 0; q < (only_c ? 1 : 2) && err == AR_OK; ++q) {
    auto launch = [&](const ar_trace_config& c) {
      return q == 0 ? launch_trace<false, true>(s, rays, n, hits.as<ar_hit>(), nullptr, nullptr, st, c)
                    : launch_trace<true, true>(s, rays, n, nullptr, nullptr, occ.as<uint32_t>(), st, c);
    };
    auto consider = [&](const ar_trace_config& c) {
      err = launch(c);
      float ms = INFINITY;  // best of two timed launches after a warm-up
      for (int rep = 0; rep < 2 && err == AR_OK; ++rep) {
        (void)hipEventRecord(e0, st);
        err = launch(c);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float m1 = 0.0f;
        (void)hipEventElapsedTime(&m1, e0, e1);
        ms = std::min(ms, m1);
      }
      if (err == AR_OK && ms < best_ms[q]) {
        best_ms[q] = ms;
        winners[q] = c;
      }
    };
    for (auto& vv : variants)
      for (uint32_t pc : per_cu)
        for (uint32_t rf : refills)
          for (uint32_t sm : searches)
            for (uint32_t ps : slot_counts)
              if (err == AR_OK) consider(ar_trace_config{vv[0], vv[1], pc, rf, sm, ps, balance[q], 0.0f});
  }
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != AR_OK) return err;
  winners[0].ms = best_ms[0];
  winners[1].ms = best_ms[1];
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (!only_a) s->tcfg = winners[0];
    if (!only_c) s->tcfg_any = winners[1];
    s->tcfg_set = true;
  }
  if (best) *best = only_a ? winners[1] : winners[0];
  return AR_OK;
}

}  // extern "C"
