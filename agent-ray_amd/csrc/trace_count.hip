// Counting build of the batch traversal kernels behind ar_trace_count: the
// walker of ar_traverse.h compiled with its instrumentation hooks defined,
// in a namespace of its own, launched with the production launch's grid and
// configuration, so it makes the same visiting decisions as the kernel that
// ar_trace / ar_occluded time.  Each lane counts into its own slot of a
// [kind][thread] array (plain increments, no cross-lane operations), summed
// on the host afterwards.  The counts are the traversal's algorithmic work:
// node and triangle records read, instance records read, stack entries
// spilled to and read back from global memory.
#include "ar_kernels.h"

namespace ar_count {

constexpr int kKinds = 5;  // node, triangle, instance, spill write, spill read
__device__ uint32_t* g_counts;  // [kKinds][threads of the launch]

__host__ __device__ __forceinline__ void bump(int kind) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t n = gridDim.x * blockDim.x;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  g_counts[(size_t)kind * n + t] += 1u;
#else
  (void)kind;
#endif
}

}  // namespace ar_count

#define AR_TRAVERSE_NS ar_count
#define AR_COUNT_NODE() ::ar_count::bump(0)
#define AR_COUNT_TRI() ::ar_count::bump(1)
#define AR_COUNT_INST() ::ar_count::bump(2)
#define AR_COUNT_SPILL_W() ::ar_count::bump(3)
#define AR_COUNT_SPILL_R() ::ar_count::bump(4)
#include "ar_traverse.h"

namespace ar_count {

using namespace ::ar;

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

// Same structure as k_trace (block remap, stack layout, walk_items, result
// writes), so the per-lane work equals the production kernel's.
template <bool ANY, bool NEAR, int S, int BLOCK, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK) void k_trace_count(SceneView sv, const ar_ray* __restrict__ rays, uint64_t n,
                                                       ar_hit* __restrict__ hits, uint32_t* __restrict__ occ,
                                                       int2* __restrict__ spill, uint32_t* __restrict__ work,
                                                       uint32_t refill_min, uint32_t search_min) {
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
        }
      });
}

template <bool ANY, bool NEAR, int S, int BLOCK>
static void launch(uint32_t G, const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* occ,
                   int2* spill, uint32_t* work, hipStream_t st, const ar_trace_config& c) {
  if (c.park_slots == 4)
    hipLaunchKernelGGL((k_trace_count<ANY, NEAR, S, BLOCK, 4>), dim3(G), dim3(BLOCK), 0, st, s->view, rays, n, hits,
                       occ, spill, work, c.refill_min, c.search_min);
  else
    hipLaunchKernelGGL((k_trace_count<ANY, NEAR, S, BLOCK, 3>), dim3(G), dim3(BLOCK), 0, st, s->view, rays, n, hits,
                       occ, spill, work, c.refill_min, c.search_min);
}

}  // namespace ar_count

using namespace ar;

extern "C" ar_status ar_trace_count(const ar_scene* s, const ar_ray* rays, uint64_t n, uint32_t flags, void* stream,
                                    ar_trace_counts* out) {
  if (!s || !rays || !out || n == 0) return fail(AR_ERR_INVALID_ARG, "ar_trace_count: NULL argument or no rays");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_trace_count: scene not committed");
  if (flags & AR_HOST_PTRS) return fail(AR_ERR_INVALID_ARG, "ar_trace_count: rays must be on the device");
  if (n >= (1ull << 31)) return fail(AR_ERR_INVALID_ARG, "ar_trace_count: at most 2^31 - 1 rays");
  const bool any = (flags & AR_COUNT_ANYHIT) != 0, near = !(flags & AR_REF_ORDER);
  const ar_trace_config c = any ? s->tcfg_any : s->tcfg;
  const uint32_t G = trace_launch_grid(n, c, any, near);
  if (G == 0) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
  const uint64_t threads = (uint64_t)G * c.workgroup_size;
  hipStream_t st = pick_stream(s, stream);
  DevBuf counts, res, spill, work;
  AR_HIP(counts.alloc(threads * ar_count::kKinds * sizeof(uint32_t)));
  AR_HIP(res.alloc(n * (any ? sizeof(uint32_t) : sizeof(ar_hit))));
  const uint32_t need = s->stack_depth;
  const size_t spill_bytes = need > c.lds_stack ? (size_t)(need - c.lds_stack) * threads * sizeof(int2) : 0;
  AR_HIP(spill.alloc(spill_bytes));
  AR_HIP(work.alloc(kWalkWorkBytes));
  AR_HIP(hipMemsetAsync(counts.p, 0, counts.bytes, st));
  AR_HIP(hipMemsetAsync(work.p, 0, work.bytes, st));
  uint32_t* const wk = c.balance ? work.as<uint32_t>() : nullptr;
  uint32_t* cp = counts.as<uint32_t>();
  AR_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(ar_count::g_counts), &cp, sizeof(cp), 0, hipMemcpyHostToDevice, st));
  bool ok = false;
#define AR_COUNT_LAUNCH(B, SS)                                                                                \
  if (c.workgroup_size == B && c.lds_stack == SS) {                                                          \
    if (any && near)                                                                                          \
      ar_count::launch<true, true, SS, B>(G, s, rays, n, nullptr, res.as<uint32_t>(), spill.as<int2>(), wk, st, c); \
    else if (any)                                                                                             \
      ar_count::launch<true, false, SS, B>(G, s, rays, n, nullptr, res.as<uint32_t>(), spill.as<int2>(), wk, st, c); \
    else if (near)                                                                                            \
      ar_count::launch<false, true, SS, B>(G, s, rays, n, res.as<ar_hit>(), nullptr, spill.as<int2>(), wk, st, c); \
    else                                                                                                      \
      ar_count::launch<false, false, SS, B>(G, s, rays, n, res.as<ar_hit>(), nullptr, spill.as<int2>(), wk, st, c); \
    ok = true;                                                                                                \
  }
  AR_TRACE_VARIANTS(AR_COUNT_LAUNCH)
#undef AR_COUNT_LAUNCH
  if (!ok) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
  AR_HIP(hipGetLastError());
  std::vector<uint32_t> h(threads * ar_count::kKinds);
  AR_HIP(hipMemcpyAsync(h.data(), counts.p, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  AR_HIP(hipStreamSynchronize(st));
  uint64_t tot[ar_count::kKinds] = {};
  for (int k = 0; k < ar_count::kKinds; ++k)
    for (uint64_t t = 0; t < threads; ++t) tot[k] += h[(size_t)k * threads + t];
  out->rays = n;
  out->node_visits = tot[0];
  out->tri_tests = tot[1];
  out->inst_tests = tot[2];
  out->spill_writes = tot[3];
  out->spill_reads = tot[4];
  out->grid = G;
  out->workgroup_size = c.workgroup_size;
  return AR_OK;
}
