// Path-tracing bounce logic shared by the megakernel, the wavefront stages
// and the gradient-domain tracer: agent-ray src/integrators/path_tracer.rs
// :59-156 (PathTracer::li), split at the two traversal queries so that a
// wavefront schedule can run each query as its own stage.  Both schedules
// call these same functions in the same order, so they produce identical
// per-path radiance.
//
// Extension over the reference: emitters (quad lights) add Le on camera hits
// and Le * w_bsdf (power heuristic against light sampling) on BSDF-sampled
// hits; the reference's non-delta MIS branch is otherwise followed as is.
#pragma once
#include "ar_kernels.h"
#include "ar_shade.h"
#include "ar_traverse.h"

namespace ar {

ShadeView shade_view(const ar_scene* s);  // device views of a committed scene

struct PtParams {
  uint32_t width, height, spp, max_depth, rr_depth;
  uint32_t npix;       // width * height
  uint32_t s0;         // first sample index of this chunk
  uint32_t k;          // samples per pixel in this chunk
  uint64_t seed;
  Cam cam;
  uint32_t variants;  // 1 for PT; 3 for GDPT (base, x+1 offset, y+1 offset)
};

struct PathState {
  Ray ray;
  f3 beta, L, prev_p;
  float prev_pdf;
  uint32_t pixel, sample, dim, depth;
};

struct ShadowReq {
  Ray ray;
  f3 c;
  bool valid;
};

__device__ __forceinline__ void wave_add_u64(unsigned long long* counter, uint32_t v) {
  uint32_t s = v;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (__lane_id() == 0 && s) atomicAdd(counter, (unsigned long long)s);
}

// A lane's traversal stack: S entries in LDS (lane-strided), the rest
// spilled to spill[entry * nthreads + gtid].
template <int BLOCK, int S = kTraceLdsStack>
__device__ __forceinline__ Stack make_stack(int2* lds, int2* spill, uint32_t gtid, uint32_t nthreads) {
  Stack stk;
  stk.lds = (LdsEntryPtr)(lds + threadIdx.x);
  stk.stride = BLOCK;
  stk.spill = (SpillEntryPtr)(spill + gtid);
  stk.spill_stride = nthreads;
  stk.S = S;
  stk.sp = 0;
  return stk;
}

// Block index remapped so that each of the 8 XCDs (which take workgroups
// round-robin) sweeps a contiguous range of blocks.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t G = gridDim.x, b = blockIdx.x;
  return (G % 8 == 0) ? (b % 8) * (G / 8) + (b / 8) : b;
}

// Workgroup-aggregated queue appends for NQ queues at once: every wave
// ballots its flags, the waves' counts meet in LDS, and one lane per queue
// takes the workgroup's range with a single atomic.  Appends from every wave
// of the device to one counter serialise in L2 (about 11 ns each here), so
// one atomic per workgroup instead of per wave is what keeps the generate
// and shade stages from being bound by them.  Lanes get consecutive slots in
// wave order.  Every thread of the workgroup must call it the same number of
// times (it contains two barriers); `par` alternates between calls so one
// call's LDS row is not overwritten while another wave still reads it.
template <int BLOCK, int NQ>
__device__ __forceinline__ void block_push(uint32_t (&sh)[2][NQ][BLOCK / 64], uint32_t par,
                                           uint32_t* const (&counter)[NQ], const bool (&flag)[NQ],
                                           uint32_t (&slot)[NQ]) {
  const uint32_t wave = threadIdx.x / 64, lane = __lane_id();
  uint64_t m[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    m[k] = __ballot(flag[k]);
    if (lane == 0) sh[par][k][wave] = (uint32_t)__popcll(m[k]);
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int k = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      const uint32_t c = sh[par][k][w];
      sh[par][k][w] = tot;
      tot += c;
    }
    const uint32_t base = tot ? atomicAdd(counter[k], tot) : 0u;
    for (int w = 0; w < BLOCK / 64; ++w) sh[par][k][w] += base;
  }
  __syncthreads();
  const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int k = 0; k < NQ; ++k) slot[k] = sh[par][k][wave] + (uint32_t)__popcll(m[k] & below);
}

__device__ __forceinline__ Ray make_ray4(float4 o, float4 d) {
  Ray r;
  r.o = mk(o.x, o.y, o.z);
  r.tmin = o.w;
  r.d = mk(d.x, d.y, d.z);
  r.tmax = d.w;
  return r;
}

// Camera ray for (x, y) with the path's first two dimensions as jitter
// (path_tracer.rs:176-182).
__device__ __forceinline__ void pt_begin(const PtParams& P, uint32_t x, uint32_t y, Rng& rng, PathState& s) {
  float jx = rng.next(), jy = rng.next();
  float u = ((float)x + jx) / (float)P.width;
  float v = 1.0f - ((float)y + jy) / (float)P.height;
  s.ray = camera_ray(P.cam, u, v);
  s.beta = mk(1.0f, 1.0f, 1.0f);
  s.L = splat(0.0f);
  s.prev_p = splat(0.0f);
  s.prev_pdf = 0.0f;
  s.depth = 0;
}

// One bounce after the closest-hit query.  Returns true when the path
// continues (s.ray is the next extension ray); `sh` receives the next-event
// shadow request whose contribution is added only if unoccluded.
__device__ __forceinline__ bool pt_shade(const ShadeView& v, const PtParams& P, PathState& s, const WorldHit& h,
                                         Rng& rng, ShadowReq& sh) {
  sh.valid = false;
  if (h.pos == kNoHit) return false;  // miss: black environment
  ShadingPoint sp;
  make_shading(v, s.ray, h.t, h.pos, h.inst, h.u, h.v, sp);
  const DMaterial m = material_at(v, sp.mat, sp.tu, sp.tv);
  f3 wo = -s.ray.d;
  if (m.light_id >= 0) {  // emitter (black surface): add Le and stop
    if (dot(sp.ng, wo) > 0.0f) {
      f3 le = mk(m.emission[0], m.emission[1], m.emission[2]);
      if (s.depth == 0 || v.dist_mode == AR_LIGHTDIST_NONE || v.n_lights == 0) {
        s.L = s.L + s.beta * le;
      } else {
        const DLight& Lt = v.lights[m.light_id];
        float pdf_l = Lt.pmf * light_pdf_li(Lt, s.prev_p, sp.p, length(sp.p - s.prev_p));
        float w = power_heuristic(s.prev_pdf, pdf_l);
        s.L = s.L + s.beta * le * w;
      }
    }
    return false;
  }
  f3 wo_l = to_local(sp.onb, wo);
  if (v.dist_mode != AR_LIGHTDIST_NONE && v.n_lights > 0) {  // path_tracer.rs:79-117
    float u_sel = rng.next();
    uint32_t li;
    float sel_pmf;
    light_select(v, u_sel, &li, &sel_pmf);
    const DLight& Lt = v.lights[li];
    float ux = rng.next(), uy = rng.next();
    LightSample ls;
    if (light_sample(Lt, sp.p, ux, uy, ls)) {
      f3 wi_l = to_local(sp.onb, ls.wi);
      float cos_wi = fabsf(wi_l.z);
      if (cos_wi > 0.0f) {
        f3 f = bsdf_eval(m, wi_l, wo_l);
        f3 c;
        if (Lt.type == AR_LIGHT_POINT) {
          c = s.beta * f * cos_wi * ls.li / sel_pmf;
        } else {
          float pdf_light = ls.pdf * sel_pmf;
          float pdf_b = bsdf_pdf(m, wi_l, wo_l);
          float w = power_heuristic(pdf_light, pdf_b);
          c = s.beta * f * cos_wi * ls.li * w / pdf_light;
        }
        if (!is_zero(c)) {
          f3 off = sp.n * 1e-4f * signum(dot(ls.wi, sp.n));
          // path_tracer.rs:88-92 for a point light.  A quad stops a relative
          // 1e-4 short of its emitter plane measured from the offset origin,
          // so the emitter's own triangles never decide occlusion by rounding.
          float t_end = ls.dist;
          if (Lt.type != AR_LIGHT_POINT) t_end = ls.dist + dot(off, mk(Lt.n[0], Lt.n[1], Lt.n[2])) / ls.cos_l;
          sh.ray = make_ray(sp.p + off, ls.wi, 0.0f, t_end * (1.0f - 1e-4f));
          sh.c = c;
          sh.valid = true;
        }
      }
    }
  }
  float u_sel = rng.next();  // path_tracer.rs:122-152
  float ux = rng.next(), uy = rng.next();
  BsdfSample bs;
  if (!bsdf_sample(m, wo_l, u_sel, ux, uy, bs)) return false;
  if (bs.pdf == 0.0f || is_zero(bs.f)) return false;
  float cos_wi = fabsf(bs.wi.z);
  s.beta = s.beta * (bs.f * cos_wi / bs.pdf);
  s.prev_pdf = bs.pdf;
  s.prev_p = sp.p;
  if (s.depth + 1 >= P.rr_depth) {
    float surv = fminf(maxc(s.beta), 0.95f);
    if (rng.next() > surv) return false;
    s.beta = s.beta / surv;
  }
  f3 wi_w = to_world(sp.onb, bs.wi);
  f3 off = sp.n * 1e-4f * signum(dot(wi_w, sp.n));
  s.ray = make_ray(sp.p + off, wi_w, 1e-4f, kF32Max);
  s.depth += 1;
  return s.depth < P.max_depth;
}

// Path p of a chunk: variant = p / (k * npix), then sample-major, pixel-minor.
// The RNG is keyed by the base pixel; GDPT's offset variants shift the camera
// sample by one pixel and replay the base path's random numbers
// (gradient_domain_pt.rs:292-327).  Returns false for an offset that falls
// outside the image (the reference then uses the base radiance).
__device__ __forceinline__ bool gen_path(const PtParams& P, uint32_t p, PathState& s, Rng& rng) {
  uint32_t per = P.npix * P.k;
  uint32_t variant = p / per;
  uint32_t rem = p - variant * per;
  uint32_t pixel = rem % P.npix;
  uint32_t x = pixel % P.width, y = pixel / P.width;
  uint32_t xs = x + (variant == 1 ? 1u : 0u), ys = y + (variant == 2 ? 1u : 0u);
  s.pixel = pixel;
  s.sample = P.s0 + rem / P.npix;
  rng.init(P.seed, pixel, s.sample, 0);
  if (xs >= P.width || ys >= P.height) return false;
  pt_begin(P, xs, ys, rng, s);
  s.dim = rng.dim;
  return true;
}

// Whole path in one thread (megakernel and GDPT).  Returns the radiance;
// `nrays` counts traversal queries.
template <bool NEAR>
__device__ __forceinline__ f3 pt_li(const ShadeView& v, const PtParams& P, PathState& s, Rng& rng, Stack& stk,
                                    uint32_t& nrays) {
  if (P.max_depth == 0) return s.L;
  for (;;) {
    WorldHit h;
    trace_scene<false, NEAR>(v.sv, s.ray, h, stk);
    nrays++;
    ShadowReq sh;
    bool cont = pt_shade(v, P, s, h, rng, sh);
    if (sh.valid) {
      WorldHit hh;
      nrays++;
      if (!trace_scene<true, NEAR>(v.sv, sh.ray, hh, stk)) s.L = s.L + sh.c;
    }
    if (!cont) break;
  }
  return s.L;
}

}  // namespace ar
