// Wavefront path state and the traversal stages shared by the path
// tracer's chunked bounce loop (pt.hip) and the serving pool (pool.hip):
// per-path state in structure-of-arrays form, extension rays (closest hit)
// and shadow rays (any hit) over device queues.
#pragma once
#include "ar_pt.h"
#include "ar_render.h"

namespace ar {

__device__ __forceinline__ void load_state(const WaveBufs& B, uint32_t i, PathState& s) {
  float4 o = B.ray_o[i], d = B.ray_d[i], b = B.beta[i], l = B.L[i], pp = B.prev_p[i];
  uint4 m = B.meta[i];
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
}

__device__ __forceinline__ void store_state(const WaveBufs& B, uint32_t i, const PathState& s) {
  B.ray_o[i] = make_float4(s.ray.o.x, s.ray.o.y, s.ray.o.z, s.ray.tmin);
  B.ray_d[i] = make_float4(s.ray.d.x, s.ray.d.y, s.ray.d.z, s.ray.tmax);
  B.beta[i] = make_float4(s.beta.x, s.beta.y, s.beta.z, s.prev_pdf);
  B.L[i] = make_float4(s.L.x, s.L.y, s.L.z, 0.0f);
  B.prev_p[i] = make_float4(s.prev_p.x, s.prev_p.y, s.prev_p.z, 0.0f);
  B.meta[i] = make_uint4(s.pixel, s.sample, s.dim, s.depth);
}

// Extension rays of the live paths (queue qi): closest hit, near-first order.
// Lanes take the queue's rays through walk_items, so a wave keeps all its
// lanes busy until its share of the queue is done.
template <int BLOCK, int S, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_wf_extend(SceneView sv, WaveBufs B, int qi, int2* __restrict__ spill,
                                                     uint32_t* __restrict__ work, uint32_t refill_min,
                                                     uint32_t search_min) {
  __shared__ int2 lds[S * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK, S>(lds, spill, gtid, nthreads);
  const uint32_t n = B.cnt[qi];
  const uint32_t* q = B.q[qi];
  walk_items<false, true, SLOTS>(
      sv, stk, n, gtid / 64, nthreads / 64, work, refill_min, search_min,
      [&](uint64_t i) {
        uint32_t p = q[i];
        return make_ray4(B.ray_o[p], B.ray_d[p]);
      },
      [&](uint64_t i, const Walker<false, true>& w) {
        B.hit[i] = make_float4(w.bt, __uint_as_float(w.bpos), w.bu, w.bv);
        B.hit_inst[i] = w.binst;
      });
  if (gtid == 0) atomicAdd(&B.stats[0], (unsigned long long)n);
}

// Shadow rays of this bounce: any hit; an unoccluded ray adds its
// contribution to its path's radiance (one shadow ray per path per bounce).
template <int BLOCK, int S, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_wf_shadow(SceneView sv, WaveBufs B, int2* __restrict__ spill,
                                                     uint32_t* __restrict__ work, uint32_t refill_min,
                                                     uint32_t search_min) {
  __shared__ int2 lds[S * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK, S>(lds, spill, gtid, nthreads);
  const uint32_t n = B.cnt[2];
  walk_items<true, true, SLOTS>(
      sv, stk, n, gtid / 64, nthreads / 64, work, refill_min, search_min,
      [&](uint64_t i) { return make_ray4(B.sh_o[i], B.sh_d[i]); },
      [&](uint64_t i, const Walker<true, true>& w) {
        if (w.bpos == kNoHit) {
          uint32_t p = B.sh_path[i];
          float4 c = B.sh_c[i], l = B.L[p];
          B.L[p] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
        }
      });
  if (gtid == 0) atomicAdd(&B.stats[1], (unsigned long long)n);
}

// Film kernels of the path tracer (pt.hip).
__global__ void k_film_accumulate(const float4* __restrict__ L, float4* __restrict__ acc, uint32_t npix, uint32_t k,
                                  uint32_t first);
__global__ void k_film_resolve(const float4* __restrict__ acc, float* __restrict__ out, uint32_t npix, float spp);
__global__ void k_gdpt_accumulate(const float4* __restrict__ L, float4* __restrict__ acc, uint32_t w, uint32_t h,
                                  uint32_t k, uint32_t first);
__global__ void k_gdpt_resolve(const float4* __restrict__ acc, float* __restrict__ films, uint32_t npix, float inv_spp);

}  // namespace ar
