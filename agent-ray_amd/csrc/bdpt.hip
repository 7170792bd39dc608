// Bidirectional path tracer with MIS on gfx950 behind ar_render
// (BidirectionalPathTracer::render, agent-ray
// src/integrators/bidirectional_path_tracer.rs:1221-1357).
//
// Phase 1, subpaths: the camera subpath (:338-437) and the light subpath
// (:443-559) of every (pixel, sample) go into a structure-of-arrays vertex
// buffer in HBM, [side][vertex slot][path], so lanes of a wave touch
// consecutive addresses for the same slot.  Wavefront schedule (default):
// all camera subpaths, then all light subpaths, as bounce loops of an extend
// stage (closest hit over the live queue, walk_items, the scene's
// closest-hit launch) and a shade stage.  Megakernel schedule
// (k_bdpt_subpaths): one lane builds both subpaths of its path.
// Phase 2, connection: one lane per path enumerates the strategies in the
// reference order (t outer, s inner; :1281-1334) and evaluates each
// contribution with its MIS weight (:594-770, fp64 sums); a contribution
// counts if its any-hit shadow ray is unoccluded.  Wavefront schedule
// (default): k_bdpt_conn_prep stores every positive contribution and its
// shadow ray in the path's slots and queues the slot, k_bdpt_conn_vis traces
// the whole queue in one any-hit stage (walk_items, the scene's any-hit
// launch), k_bdpt_conn_gather adds the unoccluded ones in enumeration order.
// Megakernel schedule: k_bdpt_connect traces each shadow ray where it is
// enumerated.  Both give the same film bit for bit.  Non-splat contributions
// go to the path's radiance slot (reduced per pixel in sample order like the
// path tracer); light-tracing (s = 1) splats are added as 64-bit fixed-point
// integers, so the splat film is the same whatever order the atomics land in.
//
// Reference behaviour kept (bug-compatible): s = 0 not implemented;
// point-light t = 1 omits 1/sel_pmf; the camera vertex normal is the first
// ray direction; surface vertices are never delta; the first camera-side
// MIS step reuses vertex s-1 in the loop.  Quad lights (extension) connect
// at t = 1 to the sampled emission point with the general formula, and the
// t = 0 strategy the reference leaves as a TODO for area lights (:810-818)
// is built: a camera subpath vertex on an emitter's front face contributes
// beta * Le (no shadow ray), weighted by the reference's walk with the
// light side's pdfs of generating it (the emitter point: sel_pmf / area;
// the vertex before it: the emission pdf in area measure); with an area
// light the light-side walk also counts the t' = 0 alternative.
#include "ar_kernels.h"
#include "ar_pt.h"
#include "ar_job.h"

namespace ar {

enum : uint32_t { VT_CAMERA = 0, VT_LIGHT = 1, VT_SURFACE = 2 };
constexpr float kSplatScale = 16777216.0f;  // 2^24 fixed point

struct BdptParams {
  PtParams pt;
  uint32_t V;  // vertex slots per subpath = max_depth + 1
  uint32_t mis_mode;
  float mis_beta;
};

struct Verts {
  float4* pf;    // (p.xyz, pdf_fwd)          [2][V][stride]
  float4* nr;    // (n.xyz, pdf_rev)
  float4* bm;    // (beta.xyz, meta bits: type | delta << 2 | emitter << 3 | id << 4)
  float2* uv;    // tex_uv of surface vertices (textured albedo)
  uint2* count;  // (camera vertices, light vertices) per path
  uint32_t stride;  // path slots per vertex plane: the workspace's capacity (a chunk uses npix * k of them)
  uint32_t V;
  __device__ __forceinline__ size_t at(uint32_t side, uint32_t k, uint32_t p) const {
    return ((size_t)side * V + k) * stride + p;
  }
};

struct PV {
  f3 p, n, beta;
  float pdf_fwd, pdf_rev;
  uint32_t type, delta, id;  // id: material (surface) or light index
  uint32_t emit = 0;         // camera-side surface vertex on the front face of an emitter
  float tu = 0.0f, tv = 0.0f;  // tex_uv (surface vertices)
};

__device__ __forceinline__ PV load_v(const Verts& Vt, uint32_t side, uint32_t k, uint32_t p) {
  size_t i = Vt.at(side, k, p);
  float4 a = Vt.pf[i], b = Vt.nr[i], c = Vt.bm[i];
  PV v;
  v.p = mk(a.x, a.y, a.z);
  v.pdf_fwd = a.w;
  v.n = mk(b.x, b.y, b.z);
  v.pdf_rev = b.w;
  v.beta = mk(c.x, c.y, c.z);
  uint32_t m = __float_as_uint(c.w);
  v.type = m & 3u;
  v.delta = (m >> 2) & 1u;
  v.emit = (m >> 3) & 1u;
  v.id = m >> 4;
  const float2 t = Vt.uv[i];
  v.tu = t.x;
  v.tv = t.y;
  return v;
}
__device__ __forceinline__ void store_v(const Verts& Vt, uint32_t side, uint32_t k, uint32_t p, const PV& v) {
  size_t i = Vt.at(side, k, p);
  uint32_t m = v.type | (v.delta << 2) | (v.emit << 3) | (v.id << 4);
  Vt.pf[i] = make_float4(v.p.x, v.p.y, v.p.z, v.pdf_fwd);
  Vt.nr[i] = make_float4(v.n.x, v.n.y, v.n.z, v.pdf_rev);
  Vt.bm[i] = make_float4(v.beta.x, v.beta.y, v.beta.z, __uint_as_float(m));
  Vt.uv[i] = make_float2(v.tu, v.tv);
}

// :305-314
__device__ __forceinline__ float pdf_sa_to_area(float pdf, f3 from, f3 to, f3 n_to) {
  f3 d = to - from;
  float d2 = dot(d, d);
  if (d2 < 1e-12f) return 0.0f;
  f3 dir = d / sqrtf(d2);
  float c = fabsf(dot(n_to, dir));
  return pdf * c / d2;
}
// :318-328
__device__ __forceinline__ float geometry_term(f3 pa, f3 na, f3 pb, f3 nb) {
  f3 d = pb - pa;
  float d2 = dot(d, d);
  if (d2 < 1e-12f) return 0.0f;
  f3 dir = d / sqrtf(d2);
  float ca = fabsf(dot(na, dir));
  float cb = fabsf(dot(nb, -dir));
  return ca * cb / d2;
}

// PinholeCamera::pdf_we / sample_we (cameras.rs:175-248)
__device__ __forceinline__ float cam_pdf_we(const Cam& c, f3 d) {
  float ct = dot(d, c.fwd);
  if (ct <= 0.0f) return 0.0f;
  return 1.0f / (c.area * ct * ct * ct);
}
struct WeSample {
  float px, py;
  f3 wi;
  float dist, we;
};
__device__ __forceinline__ bool cam_sample_we(const Cam& c, f3 p, uint32_t w, uint32_t h, WeSample& o) {
  f3 to_cam = c.origin - p;
  float d2 = dot(to_cam, to_cam);
  float dist = sqrtf(d2);
  if (dist < 1e-6f) return false;
  f3 wi = to_cam / dist;
  f3 dir = -wi;
  float ct = dot(dir, c.fwd);
  if (ct <= 1e-6f) return false;
  f3 right = normalize(c.horiz), up = normalize(c.vert);
  float vw = length(c.horiz), vh = length(c.vert);
  f3 ds = dir / ct;
  float u = dot(ds, right) / vw + 0.5f;
  float v = dot(ds, up) / vh + 0.5f;
  if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return false;
  float px = u * (float)w;
  float py = (1.0f - v) * (float)h;
  if (px < 0.0f || px >= (float)w || py < 0.0f || py >= (float)h) return false;
  float c2 = ct * ct;
  float c4 = c2 * c2;
  o.px = px;
  o.py = py;
  o.wi = wi;
  o.dist = dist;
  o.we = 1.0f / (c.area * c4);
  return true;
}

struct Emission {
  Ray ray;
  f3 le, n_light;
  float pdf_pos, pdf_dir;
};
// PointLight::sample_emission (point.rs:75-90); quad: uniform position,
// cosine-weighted direction about the normal (extension).
__device__ __forceinline__ bool light_emission(const DLight& L, float upx, float upy, float udx, float udy,
                                               Emission& e) {
  if (L.type == AR_LIGHT_POINT) {
    float ct = 1.0f - 2.0f * udx;
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    float phi = (2.0f * kPi) * udy;
    float s, c;
    sincos_f32(phi, &s, &c);
    f3 dir = mk(st * c, st * s, ct);
    e.ray = make_ray(mk(L.pos[0], L.pos[1], L.pos[2]), dir, 1e-4f, kF32Max);
    e.le = mk(L.intensity[0], L.intensity[1], L.intensity[2]);
    e.pdf_pos = 1.0f;
    e.pdf_dir = 1.0f / (4.0f * kPi);
    e.n_light = dir;
    return true;
  }
  f3 n = mk(L.n[0], L.n[1], L.n[2]);
  f3 x = (mk(L.pos[0], L.pos[1], L.pos[2]) + mk(L.eu[0], L.eu[1], L.eu[2]) * upx) + mk(L.ev[0], L.ev[1], L.ev[2]) * upy;
  f3 l = cosine_hemisphere(udx, udy);
  Onb o = onb_from_normal(n);
  f3 dir = to_world(o, l);
  if (l.z <= 0.0f) return false;
  e.ray = make_ray(x, dir, 1e-4f, kF32Max);
  e.le = mk(L.intensity[0], L.intensity[1], L.intensity[2]);
  e.pdf_pos = 1.0f / L.area;
  e.pdf_dir = l.z * kInvPi;
  e.n_light = n;
  return true;
}
__device__ __forceinline__ float light_pdf_emission_dir(const DLight& L, f3 dir) {
  if (L.type == AR_LIGHT_POINT) return 1.0f / (4.0f * kPi);
  float c = dot(mk(L.n[0], L.n[1], L.n[2]), dir);
  return c > 0.0f ? c * kInvPi : 0.0f;
}

__device__ __forceinline__ bool has_light_dist(const ShadeView& v) {
  return v.dist_mode != AR_LIGHTDIST_NONE && v.n_lights > 0;
}

// t = 0 (extension): a camera subpath vertex sv reaching the front face of
// emitter material m records it, with the reverse pdfs of the light side
// generating it -- the emitter point (light-origin pdf, area measure) in sv,
// and, from there, the previous vertex (emission pdf toward it, converted to
// area), whose stored pdf_rev is returned through prev_pdf_rev.  Without a
// light distribution no light subpath exists and both stay 0.
__device__ __forceinline__ void mark_emitter(const ShadeView& v, const DMaterial& m, const ShadingPoint& sp,
                                             const Ray& ray, const PV& prev, PV& sv, float& prev_pdf_rev) {
  prev_pdf_rev = -1.0f;  // unchanged
  if (m.light_id < 0 || !(dot(sp.ng, -ray.d) > 0.0f)) return;
  sv.emit = 1u;
  if (!has_light_dist(v)) return;
  const DLight& L = v.lights[m.light_id];
  sv.pdf_rev = L.pmf / L.area;
  prev_pdf_rev = pdf_sa_to_area(light_pdf_emission_dir(L, normalize(prev.p - sp.p)), sp.p, prev.p, prev.n);
}

// Shared subpath bounce loop (:368-434, :498-556).  Vertex 0 is already
// stored; returns the vertex count.
template <bool NEAR>
__device__ __forceinline__ uint32_t extend_subpath(const ShadeView& v, const BdptParams& P, const Verts& Vt,
                                                   uint32_t side, uint32_t p, Ray ray, f3 beta, float prev_pdf_dir,
                                                   PV prev, Rng& rng, Stack& stk, uint32_t& nrays) {
  uint32_t nv = 1;
  PV before = prev;  // vertex nv-2 (for the retroactive pdf_rev write)
  for (uint32_t depth = 0; depth < P.pt.max_depth; ++depth) {
    WorldHit h;
    trace_scene_wave<false, NEAR>(v.sv, ray, h, stk);
    nrays++;
    if (h.pos == kNoHit) break;
    ShadingPoint sp;
    make_shading(v, ray, h.t, h.pos, h.inst, h.u, h.v, sp);
    PV sv;
    sv.type = VT_SURFACE;
    sv.delta = 0;
    sv.id = sp.mat;
    sv.tu = sp.tu;
    sv.tv = sp.tv;
    sv.p = sp.p;
    sv.n = sp.n;
    sv.beta = beta;
    sv.pdf_fwd = pdf_sa_to_area(prev_pdf_dir, prev.p, sp.p, sp.n);
    sv.pdf_rev = 0.0f;
    const DMaterial m = material_at(v, sp.mat, sp.tu, sp.tv);
    if (side == 0) {
      float pr;
      mark_emitter(v, m, sp, ray, prev, sv, pr);
      if (pr >= 0.0f) {
        const size_t pi = Vt.at(side, nv - 1, p);
        const float4 nr = Vt.nr[pi];
        Vt.nr[pi] = make_float4(nr.x, nr.y, nr.z, pr);
        prev.pdf_rev = pr;
      }
    }
    store_v(Vt, side, nv, p, sv);
    before = prev;
    prev = sv;
    nv++;
    f3 wo_l = to_local(sp.onb, -ray.d);
    float u_sel = rng.next();
    float ux = rng.next(), uy = rng.next();
    BsdfSample bs;
    if (!bsdf_sample(m, wo_l, u_sel, ux, uy, bs)) break;
    if (bs.pdf == 0.0f || is_zero(bs.f)) break;
    float cos_wi = fabsf(bs.wi.z);
    beta = beta * (bs.f * cos_wi / bs.pdf);
    prev_pdf_dir = bs.pdf;
    float pdf_rev_dir = bsdf_pdf(m, wo_l, bs.wi);
    before.pdf_rev = pdf_sa_to_area(pdf_rev_dir, sp.p, before.p, before.n);
    size_t bi = Vt.at(side, nv - 2, p);
    float4 nr = Vt.nr[bi];
    Vt.nr[bi] = make_float4(nr.x, nr.y, nr.z, before.pdf_rev);
    if (depth + 1 >= P.pt.rr_depth) {
      float surv = fminf(maxc(beta), 0.95f);
      if (rng.next() > surv || surv <= 0.0f) break;
      beta = beta / surv;
    }
    f3 wi_w = to_world(sp.onb, bs.wi);
    f3 off = sp.n * 1e-4f * signum(dot(wi_w, sp.n));
    ray = make_ray(sp.p + off, wi_w, 1e-4f, kF32Max);
  }
  return nv;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bdpt_subpaths(ShadeView v, BdptParams P, Verts Vt, int2* __restrict__ spill,
                                                         unsigned long long* __restrict__ stats) {
  __shared__ int2 lds[kTraceLdsStack * BLOCK];
  const uint32_t G = gridDim.x, b = blockIdx.x;
  const uint32_t lb = (G % 8 == 0) ? (b % 8) * (G / 8) + (b / 8) : b;
  const uint32_t gtid = lb * BLOCK + threadIdx.x;
  const uint32_t nthreads = G * BLOCK;
  Stack stk;
  stk.lds = (LdsEntryPtr)(lds + threadIdx.x);
  stk.stride = BLOCK;
  stk.spill = (SpillEntryPtr)(spill + gtid);
  stk.spill_stride = nthreads;
  stk.S = kTraceLdsStack;
  stk.sp = 0;
  uint32_t nrays = 0;
  const uint32_t n = P.pt.npix * P.pt.k;
  for (uint32_t p = gtid; p < n; p += nthreads) {
    uint32_t pixel = p % P.pt.npix;
    uint32_t smp = P.pt.s0 + p / P.pt.npix;
    Rng rng;
    rng.init(P.pt.seed, pixel, smp, 0);
    float jx = rng.next(), jy = rng.next();
    float u = ((float)(pixel % P.pt.width) + jx) / (float)P.pt.width;
    float vv = 1.0f - ((float)(pixel / P.pt.width) + jy) / (float)P.pt.height;
    Ray ray = camera_ray(P.pt.cam, u, vv);
    // camera vertex (:349-361)
    PV cv;
    cv.type = VT_CAMERA;
    cv.delta = 1;
    cv.id = 0;
    cv.p = P.pt.cam.origin;
    cv.n = ray.d;
    cv.beta = mk(1.0f, 1.0f, 1.0f);
    cv.pdf_fwd = 1.0f;
    cv.pdf_rev = 0.0f;
    store_v(Vt, 0, 0, p, cv);
    uint32_t nc = extend_subpath<true>(v, P, Vt, 0, p, ray, mk(1.0f, 1.0f, 1.0f), cam_pdf_we(P.pt.cam, ray.d), cv, rng,
                                       stk, nrays);
    // light subpath (:451-496)
    uint32_t nl = 0;
    if (v.dist_mode != AR_LIGHTDIST_NONE && v.n_lights > 0) {
      float u_sel = rng.next();
      uint32_t li;
      float sel_pmf;
      light_select(v, u_sel, &li, &sel_pmf);
      const DLight& L = v.lights[li];
      float upx = rng.next(), upy = rng.next(), udx = rng.next(), udy = rng.next();
      Emission em;
      if (light_emission(L, upx, upy, udx, udy, em)) {
        PV lv;
        lv.type = VT_LIGHT;
        lv.delta = L.type == AR_LIGHT_POINT ? 1u : 0u;
        lv.id = li;
        lv.p = em.ray.o;
        lv.n = em.n_light;
        lv.pdf_fwd = em.pdf_pos * sel_pmf;
        lv.pdf_rev = 0.0f;
        lv.beta = em.le / (em.pdf_pos * sel_pmf);
        store_v(Vt, 1, 0, p, lv);
        nl = 1;
        if (P.pt.max_depth > 0) {
          f3 thr = em.le / (em.pdf_pos * sel_pmf * em.pdf_dir);
          float cos_emit = fabsf(dot(em.n_light, em.ray.d));
          thr = thr * cos_emit;
          nl = extend_subpath<true>(v, P, Vt, 1, p, em.ray, thr, em.pdf_dir, lv, rng, stk, nrays);
        }
      }
    }
    Vt.count[p] = make_uint2(nc, nl);
  }
  uint32_t s = nrays;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (__lane_id() == 0 && s) atomicAdd(&stats[0], (unsigned long long)s);
}

// ---- wavefront subpaths: the camera subpaths of every path, then the light
// subpaths (a path's light subpath draws its random numbers after its camera
// subpath's, so the two sides run one after the other), each as a bounce
// loop of extend (closest hit over the live queue) and shade (one iteration
// of extend_subpath's loop) stages.  Live subpath state: next ray in
// B.ray_o / B.ray_d, (beta, pdf of the sampled direction) in B.beta,
// (vertices stored, depth, random-number dimension, -) in B.meta.  The
// vertices and counts equal those of k_bdpt_subpaths bit for bit.

__device__ __forceinline__ void sub_state(const WaveBufs& B, uint32_t p, const Ray& r, f3 beta, float pdf_dir,
                                          uint32_t nv, uint32_t depth, uint32_t dim) {
  B.ray_o[p] = make_float4(r.o.x, r.o.y, r.o.z, r.tmin);
  B.ray_d[p] = make_float4(r.d.x, r.d.y, r.d.z, r.tmax);
  B.beta[p] = make_float4(beta.x, beta.y, beta.z, pdf_dir);
  B.meta[p] = make_uint4(nv, depth, dim, 0u);
}

// Camera vertex and first ray of every path (:338-361); queue 0 = all paths.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bdpt_cam_gen(const BdptParams* __restrict__ Pd, Verts Vt, WaveBufs B) {
  const BdptParams P = *Pd;
  const uint32_t n = P.pt.npix * P.pt.k;
  const uint32_t nthreads = gridDim.x * BLOCK;
  for (uint32_t p = xcd_block() * BLOCK + threadIdx.x; p < n; p += nthreads) {
    const uint32_t pixel = p % P.pt.npix, smp = P.pt.s0 + p / P.pt.npix;
    Rng rng;
    rng.init(P.pt.seed, pixel, smp, 0);
    float jx = rng.next(), jy = rng.next();
    float u = ((float)(pixel % P.pt.width) + jx) / (float)P.pt.width;
    float vv = 1.0f - ((float)(pixel / P.pt.width) + jy) / (float)P.pt.height;
    Ray ray = camera_ray(P.pt.cam, u, vv);
    PV cv;
    cv.type = VT_CAMERA;
    cv.delta = 1;
    cv.id = 0;
    cv.p = P.pt.cam.origin;
    cv.n = ray.d;
    cv.beta = mk(1.0f, 1.0f, 1.0f);
    cv.pdf_fwd = 1.0f;
    cv.pdf_rev = 0.0f;
    store_v(Vt, 0, 0, p, cv);
    sub_state(B, p, ray, mk(1.0f, 1.0f, 1.0f), cam_pdf_we(P.pt.cam, ray.d), 1u, 0u, rng.dim);
    Vt.count[p] = make_uint2(1u, 0u);
    B.q[0][p] = p;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    B.cnt[0] = P.pt.max_depth > 0 ? n : 0u;
    B.cnt[1] = 0u;
  }
}

// Light vertex and first ray of every path (:451-496), its random numbers
// continuing at the camera subpath's last dimension; emitting lights go to
// queue 0.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bdpt_light_gen(ShadeView v, const BdptParams* __restrict__ Pd, Verts Vt,
                                                          WaveBufs B) {
  __shared__ uint32_t qsh[2][1][BLOCK / 64];
  const BdptParams P = *Pd;
  const uint32_t n = P.pt.npix * P.pt.k;
  uint32_t* const ctr[1] = {&B.cnt[0]};
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t rounds = (n + nthreads - 1) / nthreads;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t p = r * nthreads + gtid;
    bool push = false;
    if (p < n && v.dist_mode != AR_LIGHTDIST_NONE && v.n_lights > 0) {
      const uint32_t pixel = p % P.pt.npix, smp = P.pt.s0 + p / P.pt.npix;
      Rng rng;
      rng.init(P.pt.seed, pixel, smp, 0, B.meta[p].z);
      float u_sel = rng.next();
      uint32_t li;
      float sel_pmf;
      light_select(v, u_sel, &li, &sel_pmf);
      const DLight& L = v.lights[li];
      float upx = rng.next(), upy = rng.next(), udx = rng.next(), udy = rng.next();
      Emission em;
      if (light_emission(L, upx, upy, udx, udy, em)) {
        PV lv;
        lv.type = VT_LIGHT;
        lv.delta = L.type == AR_LIGHT_POINT ? 1u : 0u;
        lv.id = li;
        lv.p = em.ray.o;
        lv.n = em.n_light;
        lv.pdf_fwd = em.pdf_pos * sel_pmf;
        lv.pdf_rev = 0.0f;
        lv.beta = em.le / (em.pdf_pos * sel_pmf);
        store_v(Vt, 1, 0, p, lv);
        Vt.count[p].y = 1u;
        if (P.pt.max_depth > 0) {
          f3 thr = em.le / (em.pdf_pos * sel_pmf * em.pdf_dir);
          float cos_emit = fabsf(dot(em.n_light, em.ray.d));
          thr = thr * cos_emit;
          sub_state(B, p, em.ray, thr, em.pdf_dir, 1u, 0u, rng.dim);
          push = true;
        }
      }
    }
    const bool fl[1] = {push};
    uint32_t slot[1];
    block_push<BLOCK, 1>(qsh, r & 1, ctr, fl, slot);
    if (push) B.q[0][slot[0]] = p;
  }
}

// Closest hit of the live subpaths of queue qi (near-first order).
template <int BLOCK, int S, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_bdpt_sub_extend(SceneView sv, WaveBufs B, int qi,
                                                                               int2* __restrict__ spill,
                                                                               uint32_t* __restrict__ work,
                                                                               uint32_t refill_min,
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
        const uint32_t p = q[i];
        return make_ray4(B.ray_o[p], B.ray_d[p]);
      },
      [&](uint64_t i, const Walker<false, true>& w) {
        B.hit[i] = make_float4(w.bt, __uint_as_float(w.bpos), w.bu, w.bv);
        B.hit_inst[i] = w.binst;
      });
  if (gtid == 0) atomicAdd(&B.stats[0], (unsigned long long)n);
}

// One iteration of extend_subpath's loop (:368-434, :498-556) for the live
// subpaths of side `side` in queue qi; survivors go to queue qi ^ 1, the
// others record their vertex count (and, on the camera side, the dimension
// where the light subpath's random numbers start).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bdpt_sub_shade(ShadeView v, const BdptParams* __restrict__ Pd, Verts Vt,
                                                          WaveBufs B, int qi, uint32_t side) {
  __shared__ uint32_t qsh[2][1][BLOCK / 64];
  const BdptParams P = *Pd;
  uint32_t* const ctr[1] = {&B.cnt[qi ^ 1]};
  const uint32_t n = B.cnt[qi];
  const uint32_t* q = B.q[qi];
  uint32_t* qn = B.q[qi ^ 1];
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t rounds = (n + nthreads - 1) / nthreads;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t i = r * nthreads + gtid;
    bool cont = false;
    uint32_t p = 0;
    if (i < n) {
      p = q[i];
      const float4 hv = B.hit[i], st = B.beta[p];
      const uint4 mt = B.meta[p];
      uint32_t nv = mt.x, depth = mt.y;
      f3 beta = mk(st.x, st.y, st.z);
      Ray ray = make_ray4(B.ray_o[p], B.ray_d[p]);
      Rng rng;
      rng.init(P.pt.seed, p % P.pt.npix, P.pt.s0 + p / P.pt.npix, 0, mt.z);
      const uint32_t pos = __float_as_uint(hv.y);
      if (pos != kNoHit) {
        ShadingPoint sp;
        make_shading(v, ray, hv.x, pos, B.hit_inst[i], hv.z, hv.w, sp);
        PV prev = load_v(Vt, side, nv - 1, p);
        PV sv;
        sv.type = VT_SURFACE;
        sv.delta = 0;
        sv.id = sp.mat;
        sv.tu = sp.tu;
        sv.tv = sp.tv;
        sv.p = sp.p;
        sv.n = sp.n;
        sv.beta = beta;
        sv.pdf_fwd = pdf_sa_to_area(st.w, prev.p, sp.p, sp.n);
        sv.pdf_rev = 0.0f;
        const DMaterial m = material_at(v, sp.mat, sp.tu, sp.tv);
        if (side == 0) {
          float pr;
          mark_emitter(v, m, sp, ray, prev, sv, pr);
          if (pr >= 0.0f) {
            const size_t pi = Vt.at(side, nv - 1, p);
            const float4 nr = Vt.nr[pi];
            Vt.nr[pi] = make_float4(nr.x, nr.y, nr.z, pr);
          }
        }
        store_v(Vt, side, nv, p, sv);
        nv++;
        f3 wo_l = to_local(sp.onb, -ray.d);
        float u_sel = rng.next();
        float ux = rng.next(), uy = rng.next();
        BsdfSample bs;
        if (bsdf_sample(m, wo_l, u_sel, ux, uy, bs) && !(bs.pdf == 0.0f || is_zero(bs.f))) {
          float cos_wi = fabsf(bs.wi.z);
          beta = beta * (bs.f * cos_wi / bs.pdf);
          float pdf_rev_dir = bsdf_pdf(m, wo_l, bs.wi);
          const float pdf_rev = pdf_sa_to_area(pdf_rev_dir, sp.p, prev.p, prev.n);
          size_t bi = Vt.at(side, nv - 2, p);
          float4 nr = Vt.nr[bi];
          Vt.nr[bi] = make_float4(nr.x, nr.y, nr.z, pdf_rev);
          bool alive = true;
          if (depth + 1 >= P.pt.rr_depth) {
            float surv = fminf(maxc(beta), 0.95f);
            if (rng.next() > surv || surv <= 0.0f)
              alive = false;
            else
              beta = beta / surv;
          }
          if (alive && depth + 1 < P.pt.max_depth) {
            f3 wi_w = to_world(sp.onb, bs.wi);
            f3 off = sp.n * 1e-4f * signum(dot(wi_w, sp.n));
            sub_state(B, p, make_ray(sp.p + off, wi_w, 1e-4f, kF32Max), beta, bs.pdf, nv, depth + 1, rng.dim);
            cont = true;
          }
        }
      }
      if (!cont) {
        if (side == 0) {
          Vt.count[p].x = nv;
          B.meta[p] = make_uint4(nv, depth, rng.dim, 0u);
        } else {
          Vt.count[p].y = nv;
        }
      }
    }
    const bool fl[1] = {cont};
    uint32_t slot[1];
    block_push<BLOCK, 1>(qsh, r & 1, ctr, fl, slot);
    if (cont) qn[slot[0]] = p;
  }
}

// End of a bounce: the consumed queue becomes empty.  With qi < 0 both.
__global__ void k_bdpt_sub_reset(WaveBufs B, int qi) {
  if (threadIdx.x == 0) {
    if (qi < 0) {
      B.cnt[0] = 0u;
      B.cnt[1] = 0u;
    } else {
      B.cnt[qi] = 0u;
    }
  }
}

__device__ __forceinline__ double powb(double x, float beta) {
  return beta == 2.0f ? x * x : pow(x, (double)beta);
}

// (pdf_fwd, pdf_rev) of vertex i of one side of path p, as the MIS walk
// reads them: from the vertex planes, or from the lane's LDS copy.
struct PdfsGlobal {
  const Verts& Vt;
  uint32_t p;
  __device__ __forceinline__ float2 operator()(uint32_t side, uint32_t i) const {
    const size_t a = Vt.at(side, i, p);
    return make_float2(Vt.pf[a].w, Vt.nr[a].w);
  }
};
struct PdfsLds {
  const float2* base;  // the lane's first entry; [side][vertex] at a stride of the workgroup size
  uint32_t V, stride;
  __device__ __forceinline__ float2 operator()(uint32_t side, uint32_t i) const {
    return base[(side * V + i) * stride];
  }
};

// :594-770.  rev_z2 / rev_y2: the pdfs of z[s-2] / y[t-2] generated from
// z[s-1] / y[t-1] arriving from the connection, used by the textbook walk
// only (mis_mode 2).
template <class Pdfs>
__device__ __forceinline__ float mis_weight(const BdptParams& P, const Pdfs& pd, uint32_t nc, uint32_t nl, uint32_t s,
                                            uint32_t t, float pdf_cc, float pdf_cl, bool area_light, float rev_z2,
                                            float rev_y2) {
  if (P.mis_mode == 1) {  // :605-636
    uint32_t len = s + t, count = 0;
    uint32_t tmx = len < nl ? len : nl;
    for (uint32_t t2 = 0; t2 <= tmx; ++t2) {
      uint32_t s2 = len - t2;
      if (s2 > nc || s2 + t2 < 2 || s2 == 0) continue;
      count++;
    }
    return count == 0 ? 1.0f : 1.0f / (float)count;
  }
  if (s + t == 2) return 1.0f;
  double sum = 0.0;
  if (P.mis_mode == 2) {
    // Textbook walk (extension): strategy (s', t') of the same path is
    // reached by moving the connection; each vertex's reverse pdf is taken
    // once, with the connection's own pdfs at the two vertices next to it.
    // Alternatives the subpath lengths cannot produce (max_depth + 1
    // vertices per subpath) or that do not exist (s' = 0; t' = 0 without an
    // area light) are left out, so a path's weights sum to one.
    const int k = (int)(s + t), dmax = (int)P.pt.max_depth + 1;
    double ri = 1.0;
    for (int i = (int)s - 1; i >= 1; --i) {  // alternative (i, k - i)
      const float2 v = pd(0, (uint32_t)i);
      const double rev = i == (int)s - 1 ? pdf_cc : (i == (int)s - 2 ? rev_z2 : fmaxf(v.y, 0.0f));
      ri *= rev / (double)fmaxf(v.x, 1e-30f);
      if (k - i <= dmax) sum += powb(fabs(ri), P.mis_beta);
    }
    ri = 1.0;
    for (int i = (int)t - 1; i >= 0; --i) {  // alternative (k - i, i)
      const float2 v = pd(1, (uint32_t)i);
      const double rev = i == (int)t - 1 ? pdf_cl : (i == (int)t - 2 ? rev_y2 : fmaxf(v.y, 0.0f));
      ri *= rev / (double)fmaxf(v.x, 1e-30f);
      if ((i >= 1 || area_light) && k - i <= dmax) sum += powb(fabs(ri), P.mis_beta);
    }
    return (float)(1.0 / (1.0 + sum));
  }
  // Only surface vertices can be MIS-delta, and the reference never marks
  // them delta (:268), so every "delta" test below is false except where the
  // reference forces an unimplemented s' = 0 / t' = 0 strategy.
  {
    double ri = 1.0;
    if (s >= 1) {
      const float2 v = pd(0, s - 1);
      ri *= (double)pdf_cc / (double)fmaxf(v.x, 1e-30f);
      bool prev_delta = s >= 2 ? false : true;
      if (!prev_delta) sum += powb(fabs(ri), P.mis_beta);
    }
    for (uint32_t i = s - 1; s >= 1 && i >= 1; --i) {
      const float2 v = pd(0, i);
      ri *= (double)fmaxf(v.y, 0.0f) / (double)fmaxf(v.x, 1e-30f);
      sum += powb(fabs(ri), P.mis_beta);
    }
  }
  {
    // with an area light the t' = 0 alternative counts (extension; the
    // reference skips it "for delta lights", :735-742)
    double ri = 1.0;
    if (t >= 1) {
      const float2 v = pd(1, t - 1);
      ri *= (double)pdf_cl / (double)fmaxf(v.x, 1e-30f);
      bool prev_delta = t >= 2 ? false : !area_light;
      if (!prev_delta) sum += powb(fabs(ri), P.mis_beta);
    }
    const int lo = area_light ? 0 : 1;
    for (int i = (int)t - 1; i >= lo; --i) {
      const float2 v = pd(1, (uint32_t)i);
      ri *= (double)fmaxf(v.y, 0.0f) / (double)fmaxf(v.x, 1e-30f);
      sum += powb(fabs(ri), P.mis_beta);
    }
  }
  return (float)(1.0 / (1.0 + sum));
}

__device__ __forceinline__ void splat_add(unsigned long long* film, uint32_t idx, f3 c) {
  float vals[3] = {c.x, c.y, c.z};
  for (int k = 0; k < 3; ++k) {
    float s = vals[k] * kSplatScale;
    long long q = 0;
    if (s == s) q = (long long)fminf(fmaxf(rintf(s), -4.0e18f), 4.0e18f);
    if (q != 0) atomicAdd(film + 3 * (size_t)idx + k, (unsigned long long)q);
  }
}

constexpr uint32_t kOwnPixel = 0xffffffffu;  // target of a contribution to the path's own pixel

// Every connection strategy of path p in the reference order (t outer, s
// inner; :1281-1334) whose unweighted contribution is positive, as
// emit(shadow ray, weighted contribution, target, ray): the light-tracing
// strategies (s = 1, :1051-1155) target the pixel the light vertex projects
// to (a splat), the others (:844-1043) kOwnPixel.  The contribution counts
// only if the shadow ray is unoccluded; the caller decides when to trace it.
// The t = 0 strategy (extension) has no shadow ray (ray = false); only the
// last camera vertex can lie on an emitter, so it is a path's first strategy
// if it exists.
template <class Pdfs, class Emit>
__device__ __forceinline__ void bdpt_strategies(const ShadeView& v, const BdptParams& P, const Verts& Vt, uint32_t p,
                                                const Pdfs& pdfs, Emit emit) {
  const Cam& cam = P.pt.cam;
  const uint32_t W = P.pt.width, H = P.pt.height;
  uint2 cnt = Vt.count[p];
  uint32_t smax = cnt.x, tmax = cnt.y, limit = P.pt.max_depth + 2;
  // the path's light end lies on an area light: a t' = 0 alternative exists
  const bool area_light = tmax >= 1 && v.lights[load_v(Vt, 1, 0, p).id].type == AR_LIGHT_QUAD;
  for (uint32_t t = 0; t <= tmax; ++t) {
    for (uint32_t s = 0; s <= smax; ++s) {
      if (s + t < 2 || s + t > limit || s == 0) continue;
      if (t == 0) {  // the camera subpath ends on an emitter (:810-818, extension)
        if (s < 2) continue;
        PV z = load_v(Vt, 0, s - 1, p);
        if (!z.emit) continue;
        const DMaterial mz = material_at(v, z.id, z.tu, z.tv);
        f3 c = z.beta * mk(mz.emission[0], mz.emission[1], mz.emission[2]);
        if (maxc(c) <= 0.0f) continue;
        float pdf_cc = 0.0f;
        if (has_light_dist(v)) {
          const DLight& L = v.lights[mz.light_id];
          pdf_cc = L.pmf / L.area;
        }
        const float rev_z2 = s >= 3 ? pdfs(0, s - 2).y : 0.0f;
        float w = mis_weight(P, pdfs, smax, tmax, s, 0, pdf_cc, 0.0f, true, rev_z2, 0.0f);
        emit(Ray{}, c * w, kOwnPixel, false);
        continue;
      }
      if (s == 1) {  // light tracing to the camera (:1051-1155)
        if (t < 2) continue;
        PV y = load_v(Vt, 1, t - 1, p);
        WeSample cs;
        if (!cam_sample_we(cam, y.p, W, H, cs)) continue;
        f3 to_cam = cs.wi;
        f3 off = y.n * 1e-4f * signum(dot(to_cam, y.n));
        Ray sr = make_ray(y.p + off, to_cam, 0.0f, cs.dist * (1.0f - 1e-4f));
        PV yp = load_v(Vt, 1, t - 2, p);
        Onb on = onb_from_normal(y.n);
        f3 wo_l = to_local(on, normalize(yp.p - y.p));
        f3 wi_l = to_local(on, to_cam);
        const DMaterial m = material_at(v, y.id, y.tu, y.tv);
        f3 f = bsdf_eval(m, wi_l, wo_l);
        if (is_zero(f)) continue;
        float cos_y = fabsf(dot(y.n, to_cam));
        float g = cos_y / (cs.dist * cs.dist);
        f3 c = y.beta * f * g * cs.we;
        // the textbook walk takes the camera's cosine, which the reference's
        // We = 1 / (A cos^4) leaves out (:1104-1108)
        if (P.mis_mode == 2) c = c * dot(-to_cam, cam.fwd);
        if (maxc(c) <= 0.0f) continue;
        float pdf_cl, rev_y2 = 0.0f;
        if (P.mis_mode == 2) {
          // the camera's own pdf of the direction to y, in area measure at y
          const Ray cr = make_ray(cam.origin, -to_cam, 0.0f, kF32Max);
          pdf_cl = pdf_sa_to_area(cam_pdf_we(cam, cr.d), cam.origin, y.p, y.n);
          rev_y2 = pdf_sa_to_area(bsdf_pdf(m, wo_l, wi_l), y.p, yp.p, yp.n);
        } else {
          float pd = bsdf_pdf(m, wi_l, wo_l);
          PV c0 = load_v(Vt, 0, 0, p);
          pdf_cl = pdf_sa_to_area(pd, y.p, cam.origin, c0.n);
        }
        float w = mis_weight(P, pdfs, smax, tmax, 1, t, 0.0f, pdf_cl, area_light, 0.0f, rev_y2);
        c = c * w;
        uint32_t ix = (uint32_t)cs.px, iy = (uint32_t)cs.py;
        if (ix > W - 1) ix = W - 1;
        if (iy > H - 1) iy = H - 1;
        emit(sr, c, iy * W + ix, true);
        continue;
      }
      // general / NEE (:844-1043)
      PV z = load_v(Vt, 0, s - 1, p);
      PV y = load_v(Vt, 1, t - 1, p);
      f3 d = y.p - z.p;
      float d2 = dot(d, d);
      if (d2 < 1e-10f) continue;
      float dist = sqrtf(d2);
      f3 wi_z = d / dist;
      f3 wi_y = -wi_z;
      f3 oz = z.n * 1e-4f * signum(dot(wi_z, z.n));
      f3 oy = y.n * 1e-4f * signum(dot(wi_y, y.n));
      f3 so = z.p + oz, st = y.p + oy;
      f3 sd = st - so;
      float sdist = length(sd);
      if (sdist < 1e-6f) continue;
      Ray sr = make_ray(so, sd / sdist, 0.0f, sdist * (1.0f - 1e-4f));
      PV zp = load_v(Vt, 0, s - 2, p);
      Onb onz = onb_from_normal(z.n);
      f3 wo_z_l = to_local(onz, normalize(zp.p - z.p));
      const DMaterial mz = material_at(v, z.id, z.tu, z.tv);
      f3 f_z = bsdf_eval(mz, to_local(onz, wi_z), wo_z_l);
      if (is_zero(f_z)) continue;
      f3 f_y = mk(1.0f, 1.0f, 1.0f);
      const DLight* Lt = nullptr;
      Onb ony = onb_from_normal(y.n);
      f3 wo_y_l = splat(0.0f);
      if (t == 1) {
        Lt = &v.lights[y.id];
        if (Lt->type == AR_LIGHT_QUAD && !(dot(y.n, wi_y) > 0.0f)) f_y = splat(0.0f);
      } else {
        PV yp = load_v(Vt, 1, t - 2, p);
        wo_y_l = to_local(ony, normalize(yp.p - y.p));
        f_y = bsdf_eval(material_at(v, y.id, y.tu, y.tv), to_local(ony, wi_y), wo_y_l);
      }
      if (is_zero(f_y)) continue;
      float g = geometry_term(z.p, z.n, y.p, y.n);
      f3 c;
      if (t == 1 && Lt->type == AR_LIGHT_POINT) {  // :947-967
        LightSample ls;
        if (!light_sample(*Lt, z.p, 0.0f, 0.0f, ls)) continue;
        f3 fd = bsdf_eval(mz, to_local(onz, ls.wi), wo_z_l);
        float cos_z = fabsf(dot(z.n, ls.wi));
        c = z.beta * fd * cos_z * ls.li;
        if (P.mis_mode == 2) c = c / Lt->pmf;  // textbook: the light pick's pmf
      } else {
        c = z.beta * f_z * g * f_y * y.beta;
      }
      if (maxc(c) <= 0.0f) continue;
      float pdf_cc, rev_y2 = 0.0f, rev_z2 = 0.0f;
      if (t == 1) {
        f3 dyz = normalize(z.p - y.p);
        pdf_cc = pdf_sa_to_area(light_pdf_emission_dir(*Lt, dyz), y.p, z.p, z.n);
      } else {
        f3 wi_y2 = normalize(z.p - y.p);
        const DMaterial my = material_at(v, y.id, y.tu, y.tv);
        float pd = bsdf_pdf(my, to_local(ony, wi_y2), wo_y_l);
        pdf_cc = pdf_sa_to_area(pd, y.p, z.p, z.n);
        if (P.mis_mode == 2) {
          const PV yp = load_v(Vt, 1, t - 2, p);
          rev_y2 = pdf_sa_to_area(bsdf_pdf(my, wo_y_l, to_local(ony, wi_y2)), y.p, yp.p, yp.n);
        }
      }
      f3 wi_z2 = normalize(y.p - z.p);
      float pdz = bsdf_pdf(mz, to_local(onz, wi_z2), wo_z_l);
      float pdf_cl = pdf_sa_to_area(pdz, z.p, y.p, y.n);
      if (P.mis_mode == 2 && s >= 3) rev_z2 = pdf_sa_to_area(bsdf_pdf(mz, wo_z_l, to_local(onz, wi_z2)), z.p, zp.p, zp.n);
      float w = mis_weight(P, pdfs, smax, tmax, s, t, pdf_cc, pdf_cl, area_light, rev_z2, rev_y2);
      emit(sr, c * w, kOwnPixel, true);
    }
  }
}

// Connection strategies traced where they are enumerated, one lane per path
// (the megakernel schedule).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bdpt_connect(ShadeView v, BdptParams P, Verts Vt, float4* __restrict__ Lout,
                                                        unsigned long long* __restrict__ splat_film, int2* __restrict__ spill,
                                                        unsigned long long* __restrict__ stats) {
  __shared__ int2 lds[kTraceLdsStack * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK>(lds, spill, gtid, nthreads);
  uint32_t nrays = 0;
  const uint32_t n = P.pt.npix * P.pt.k;
  for (uint32_t p = gtid; p < n; p += nthreads) {
    f3 acc = splat(0.0f);
    bdpt_strategies(v, P, Vt, p, PdfsGlobal{Vt, p}, [&](const Ray& sr, f3 c, uint32_t target, bool ray) {
      WorldHit hh;
      if (ray) {
        nrays++;
        if (trace_scene_wave<true, true>(v.sv, sr, hh, stk)) return;
      }
      if (target == kOwnPixel)
        acc = acc + c;
      else
        splat_add(splat_film, target, c);
    });
    Lout[p] = make_float4(acc.x, acc.y, acc.z, 0.0f);
  }
  wave_add_u64(&stats[1], nrays);
}

// ---- wavefront connection: enumerate, trace the shadow rays together, gather
// Pass i takes paths [p0, p0 + m) of the chunk, p0 = i * pass.  Path p's
// strategies occupy slots (p - p0) * K + j, j < cnt, in enumeration order;
// the queue lists the used slots (appended per wave) for the any-hit stage,
// which writes occl[slot].  A pass past the chunk's paths (the last chunk
// can be shorter than the one the launches were recorded for) has m = 0.
struct ConnBufs {
  float4* o;        // shadow ray origin, t_min           [m * K]
  float4* d;        // direction, t_max
  float4* val;      // weighted contribution, target bits
  uint8_t* occl;    // 1: occluded
  uint32_t* queue;  // used slots                          [m * K]
  uint32_t* cnt;    // strategies per path                 [m]
  uint32_t* qn;     // queue length
  uint32_t K, pass;
};

struct PassRange {
  uint32_t p0, m;
};
__device__ __forceinline__ PassRange pass_range(const BdptParams& P, const ConnBufs& C, uint32_t pi) {
  const uint32_t n = P.pt.npix * P.pt.k, p0 = pi * C.pass;
  return PassRange{p0, n > p0 ? min(C.pass, n - p0) : 0u};
}

__global__ void k_bdpt_conn_reset(ConnBufs C) {
  if (threadIdx.x == 0) *C.qn = 0u;
}

// With LDS, each lane first copies its path's (pdf_fwd, pdf_rev) pairs into
// the workgroup's dynamic LDS (2 * V * BLOCK float2, [side][vertex][lane]),
// so the MIS walks of its strategies read them there instead of from HBM.
template <int BLOCK, bool LDS>
__global__ __launch_bounds__(BLOCK) void k_bdpt_conn_prep(ShadeView v, const BdptParams* __restrict__ Pd, Verts Vt,
                                                          ConnBufs C, uint32_t pi, unsigned long long* __restrict__ stats) {
  extern __shared__ float2 pdf_lds[];
  const BdptParams P = *Pd;
  const PassRange R = pass_range(P, C, pi);
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t lane = __lane_id();
  uint32_t nrays = 0;
  // whole waves iterate together so the wave-aggregated append sees every lane
  const uint32_t rounds = (R.m + nthreads - 1) / nthreads;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t i = r * nthreads + gtid;
    uint32_t j = 0, j0 = 0;  // slots used; leading slots without a shadow ray (t = 0)
    if (i < R.m) {
      const uint32_t p = R.p0 + i;
      const size_t base = (size_t)i * C.K;
      auto emit = [&](const Ray& sr, f3 c, uint32_t target, bool ray) {
        const size_t slot = base + j++;
        C.val[slot] = make_float4(c.x, c.y, c.z, __uint_as_float(target));
        if (ray) {
          C.o[slot] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmin);
          C.d[slot] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.tmax);
        } else {
          C.occl[slot] = 0;  // counts unconditionally; not queued
          j0 = j;
        }
      };
      if (LDS) {
        // the walks read vertices below the subpath lengths only
        const uint2 cnt = Vt.count[p];
        float2* mine = pdf_lds + threadIdx.x;
        for (uint32_t k = 0; k < cnt.x; ++k) {
          const size_t a = Vt.at(0, k, p);
          mine[k * BLOCK] = make_float2(Vt.pf[a].w, Vt.nr[a].w);
        }
        for (uint32_t k = 0; k < cnt.y; ++k) {
          const size_t a = Vt.at(1, k, p);
          mine[(Vt.V + k) * BLOCK] = make_float2(Vt.pf[a].w, Vt.nr[a].w);
        }
        bdpt_strategies(v, P, Vt, p, PdfsLds{mine, Vt.V, (uint32_t)BLOCK}, emit);
      } else {
        bdpt_strategies(v, P, Vt, p, PdfsGlobal{Vt, p}, emit);
      }
      C.cnt[i] = j;
    }
    const uint32_t nq = j - j0;  // slots j0 .. j-1 need their shadow ray traced
    nrays += nq;
    uint32_t x = nq;  // inclusive scan of the wave's counts
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    const uint32_t total = __shfl(x, 63);
    uint32_t start = 0;
    if (lane == 63 && total) start = atomicAdd(C.qn, total);
    start = __shfl(start, 63) + x - nq;
    for (uint32_t k = 0; k < nq; ++k) C.queue[start + k] = i * C.K + j0 + k;
  }
  wave_add_u64(&stats[1], nrays);
}

template <int BLOCK, int S, uint32_t SLOTS>
__global__ __launch_bounds__(BLOCK, trace_min_waves(S)) void k_bdpt_conn_vis(SceneView sv, ConnBufs C,
                                                                             int2* __restrict__ spill,
                                                                             uint32_t* __restrict__ work,
                                                                             uint32_t refill_min, uint32_t search_min) {
  __shared__ int2 lds[S * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK, S>(lds, spill, gtid, nthreads);
  const uint32_t n = *C.qn;
  walk_items<true, true, SLOTS>(
      sv, stk, n, gtid / 64, nthreads / 64, work, refill_min, search_min,
      [&](uint64_t i) {
        const uint32_t slot = C.queue[i];
        return make_ray4(C.o[slot], C.d[slot]);
      },
      [&](uint64_t i, const Walker<true, true>& w) { C.occl[C.queue[i]] = w.bpos != kNoHit ? 1 : 0; });
}

// Unoccluded contributions in enumeration order: the path's own radiance is
// summed exactly as the megakernel schedule sums it; splats are fixed-point.
__global__ void k_bdpt_conn_gather(const BdptParams* __restrict__ Pd, ConnBufs C, uint32_t pi, float4* __restrict__ Lout,
                                   unsigned long long* __restrict__ splat_film) {
  const PassRange R = pass_range(*Pd, C, pi);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < R.m; i += gridDim.x * blockDim.x) {
    const uint32_t n = C.cnt[i];
    f3 acc = splat(0.0f);
    for (uint32_t k = 0; k < n; ++k) {
      const size_t slot = (size_t)i * C.K + k;
      if (C.occl[slot]) continue;
      const float4 c = C.val[slot];
      const uint32_t target = __float_as_uint(c.w);
      if (target == kOwnPixel)
        acc = acc + mk(c.x, c.y, c.z);
      else
        splat_add(splat_film, target, mk(c.x, c.y, c.z));
    }
    Lout[R.p0 + i] = make_float4(acc.x, acc.y, acc.z, 0.0f);
  }
}

// film = acc / spp, then + splat * (1/spp) where the splat is non-zero (:1341-1348)
__global__ void k_bdpt_resolve(const float4* __restrict__ acc, const long long* __restrict__ splat, float* out,
                               uint32_t npix, float spp) {
  float inv = 1.0f / spp;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    float4 a = acc[px];
    float m[3] = {a.x / spp, a.y / spp, a.z / spp};
    float sp[3];
    bool nz = false;
    for (int k = 0; k < 3; ++k) {
      sp[k] = (float)((double)splat[3 * (size_t)px + k] * (1.0 / (double)kSplatScale));
      nz |= sp[k] != 0.0f;
    }
    for (int k = 0; k < 3; ++k) out[3 * (size_t)px + k] = nz ? m[k] + sp[k] * inv : m[k];
  }
}

__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

// The chunk's per-sample radiance added to the film in sample order.
__global__ void k_film_accumulate_b(const float4* __restrict__ L, float4* __restrict__ acc,
                                    const BdptParams* __restrict__ Pd) {
  const uint32_t npix = Pd->pt.npix, k = Pd->pt.k;
  const bool first = Pd->pt.s0 == 0;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    float4 a = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : acc[px];
    for (uint32_t s = 0; s < k; ++s) {
      float4 l = L[(size_t)s * npix + px];
      a.x = a.x + l.x;
      a.y = a.y + l.y;
      a.z = a.z + l.z;
    }
    acc[px] = a;
  }
}

static uint32_t bdpt_grid(int block) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) == hipSuccess) cus = pr.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return (uint32_t)(cus * kTraceBlocksPerCU / 8 * 8);
}

// The chunk's parameters, written ahead of its (possibly graph-replayed)
// launches, which read them from this block.
static_assert(sizeof(BdptParams) <= kParamsBytes, "chunk parameters exceed their device block");
__global__ void k_bdpt_params(BdptParams P, BdptParams* dst) {
  if (threadIdx.x == 0) *dst = P;
}

__global__ void k_bdpt_zero_stats(unsigned long long* st) {
  if (threadIdx.x < 2) st[threadIdx.x] = 0ull;
}

class BdptJob : public JobRun {
 public:
  BdptParams P{};
  ShadeView v{};
  Verts Vt{};
  ConnBufs C{};
  ar_trace_config tc{}, ta{};
  unsigned long long* splat = nullptr;
  uint32_t k = 1, s0 = 0, G = 0, eg = 0, Gs = 0, Ge = 0, Gw = 0, pass = 0, cap = 0, spill_entries = 0;
  uint32_t phase = 0;          // next piece of the wavefront chunk: 0 camera, 1 light subpaths, 2 + i pass i
  double step_share = 1.0;     // share of the chunk the last step issued
  bool wavefront = true;
  static constexpr int kSubBlock = 512;  // subpath generate / shade stages
  static constexpr size_t kPdfLdsMax = 64 * 1024;  // LDS for the MIS walks' pdf pairs (depth <= 31 at 128 lanes)

  ar_status begin() override {
    constexpr int BLOCK = kTraceBlock;
    const ar_render_config* cfg = &spec.cfg;
    const uint32_t w = spec.width, h = spec.height, npix = w * h;
    v = shade_view(scene);
    P.pt.width = w;
    P.pt.height = h;
    P.pt.spp = cfg->spp;
    P.pt.max_depth = cfg->max_depth;
    P.pt.rr_depth = cfg->rr_depth;
    P.pt.npix = npix;
    P.pt.seed = cfg->seed;
    P.pt.cam = cam_from(spec.camera);
    P.pt.variants = 1;
    P.V = cfg->max_depth + 1;
    P.mis_mode = cfg->mis_mode;
    P.mis_beta = cfg->mis_beta;
    k = chunk_samples(spec);
    G = bdpt_grid(BLOCK);
    cap = npix * k;
    wavefront = cfg->schedule != AR_SCHED_MEGAKERNEL;
    // subpath extension stages: the scene's closest-hit launch; any-hit
    // stage of the wavefront connection: the scene's any-hit launch
    tc = scene->tcfg;
    ta = scene->tcfg_any;
    uint64_t spill_words = stack_spill_entries(scene) * (uint64_t)G * BLOCK;
    if (wavefront) {
      Gw = (uint32_t)(device_cus() * (kTraceBlocksPerCU * kTraceBlock / kSubBlock) / 8 * 8);
      const uint32_t Gt = (uint32_t)(device_cus() * tc.blocks_per_cu / 8 * 8);
      Ge = 0;
#define AR_EXT_GRID(BB, SS)                                                                         \
  if (tc.workgroup_size == BB && tc.lds_stack == SS)                                                \
    Ge = clamp_resident(Gt,                                                                         \
                        tc.park_slots == 4 ? (const void*)&k_bdpt_sub_extend<BB, SS, 4>             \
                                           : (const void*)&k_bdpt_sub_extend<BB, SS, 3>,            \
                        BB);
      AR_TRACE_VARIANTS(AR_EXT_GRID)
#undef AR_EXT_GRID
      if (Ge == 0) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
      {
        const uint32_t need = scene->stack_depth;
        const uint64_t e_t = need > tc.lds_stack ? need - tc.lds_stack : 0;
        spill_words = std::max<uint64_t>(spill_words, e_t * (uint64_t)Ge * tc.workgroup_size);
      }
      const uint32_t Ga = (uint32_t)(device_cus() * ta.blocks_per_cu / 8 * 8);
      Gs = 0;
#define AR_VIS_GRID(BB, SS)                                                                         \
  if (ta.workgroup_size == BB && ta.lds_stack == SS)                                                \
    Gs = clamp_resident(Ga,                                                                         \
                        ta.park_slots == 4 ? (const void*)&k_bdpt_conn_vis<BB, SS, 4>               \
                                           : (const void*)&k_bdpt_conn_vis<BB, SS, 3>,              \
                        BB);
      AR_TRACE_VARIANTS(AR_VIS_GRID)
#undef AR_VIS_GRID
      if (Gs == 0) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
      const uint32_t need = scene->stack_depth;
      const uint64_t e_t = need > ta.lds_stack ? need - ta.lds_stack : 0;
      spill_words = std::max<uint64_t>(spill_words, e_t * (uint64_t)Gs * ta.workgroup_size);
    }
    spill_entries = (uint32_t)((spill_words + G * BLOCK - 1) / (G * BLOCK));
    ar_status e = ws->reserve_paths(cap, npix, spill_entries, G * BLOCK, stream);
    if (e != AR_OK) return e;
    size_t nvert = 2ull * P.V * cap;
    C.K = (uint32_t)bdpt_conn_slots(cfg->max_depth);
    pass = wavefront ? (uint32_t)bdpt_conn_pass(cfg->max_depth, cap, spec.conn_pass) : 0;
    C.pass = pass;
    const size_t slots = (size_t)pass * C.K;
    auto al = [](size_t n) { return (n + 255) / 256 * 256; };
    size_t need = al(nvert * 3 * sizeof(float4)) + al(nvert * sizeof(float2)) + al((size_t)cap * sizeof(uint2)) +
                  al((size_t)npix * 3 * sizeof(long long));
    if (wavefront)
      need += 3 * al(slots * sizeof(float4)) + al(slots) + al(slots * sizeof(uint32_t)) +
              al((size_t)pass * sizeof(uint32_t)) + 256;
    if (ws->aux.bytes < need) {
      AR_HIP(hipStreamSynchronize(stream));
      AR_HIP(ws->aux.alloc(need));
    }
    char* cur = (char*)ws->aux.p;
    auto take = [&](size_t n) {
      char* q = cur;
      cur += al(n);
      return q;
    };
    Vt.pf = (float4*)take(nvert * 3 * sizeof(float4));
    Vt.nr = Vt.pf + nvert;
    Vt.bm = Vt.nr + nvert;
    Vt.uv = (float2*)take(nvert * sizeof(float2));
    Vt.count = (uint2*)take((size_t)cap * sizeof(uint2));
    splat = (unsigned long long*)take((size_t)npix * 3 * sizeof(long long));
    if (wavefront) {
      C.o = (float4*)take(slots * sizeof(float4));
      C.d = (float4*)take(slots * sizeof(float4));
      C.val = (float4*)take(slots * sizeof(float4));
      C.occl = (uint8_t*)take(slots);
      C.queue = (uint32_t*)take(slots * sizeof(uint32_t));
      C.cnt = (uint32_t*)take((size_t)pass * sizeof(uint32_t));
      C.qn = (uint32_t*)take(256);
    }
    Vt.V = P.V;
    Vt.stride = cap;
    eg = std::min<uint32_t>((npix + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bdpt_zero_stats, dim3(1), dim3(64), 0, stream, ws->B.stats);
    hipLaunchKernelGGL(k_zero_u64, dim3(eg), dim3(256), 0, stream, splat, (size_t)npix * 3);
    s0 = 0;
    return AR_OK;
  }

  // A chunk of the wavefront schedule is issued in pieces, so that a server
  // can put other work between them: the camera subpaths, the light
  // subpaths (each bounce loop captured once per job shape in a HIP graph
  // and replayed, the chunk's s0 and k being read from the device block),
  // then each connection pass, the last followed by the film.  Passes past
  // the chunk's paths are not issued (they would read an empty range).
  ar_status run_subpaths(const BdptParams* Pd, uint32_t side) {
    const WaveBufs& B = ws->B;
    int2* const spill = ws->spill;
    auto record = [&]() {
      if (side == 0)
        hipLaunchKernelGGL((k_bdpt_cam_gen<kSubBlock>), dim3(Gw), dim3(kSubBlock), 0, stream, Pd, Vt, B);
      else
        hipLaunchKernelGGL((k_bdpt_light_gen<kSubBlock>), dim3(Gw), dim3(kSubBlock), 0, stream, v, Pd, Vt, B);
      for (uint32_t d = 0; d < P.pt.max_depth; ++d) {
        const int qi = d & 1;
#define AR_SUB_EXT(BB, SS)                                                                                   \
  if (tc.workgroup_size == BB && tc.lds_stack == SS) {                                                       \
    if (tc.park_slots == 4)                                                                                  \
      hipLaunchKernelGGL((k_bdpt_sub_extend<BB, SS, 4>), dim3(Ge), dim3(BB), 0, stream, v.sv, B, qi, spill,  \
                         tc.balance ? B.work : nullptr, tc.refill_min, tc.search_min);                       \
    else                                                                                                     \
      hipLaunchKernelGGL((k_bdpt_sub_extend<BB, SS, 3>), dim3(Ge), dim3(BB), 0, stream, v.sv, B, qi, spill,  \
                         tc.balance ? B.work : nullptr, tc.refill_min, tc.search_min);                       \
  }
        AR_TRACE_VARIANTS(AR_SUB_EXT)
#undef AR_SUB_EXT
        hipLaunchKernelGGL((k_bdpt_sub_shade<kSubBlock>), dim3(Gw), dim3(kSubBlock), 0, stream, v, Pd, Vt, B, qi,
                           side);
        hipLaunchKernelGGL(k_bdpt_sub_reset, dim3(1), dim3(64), 0, stream, B, qi);
      }
      hipLaunchKernelGGL(k_bdpt_sub_reset, dim3(1), dim3(64), 0, stream, B, -1);
    };
    const uint64_t key[GraphCache::kKeyWords] = {(uint64_t)(uintptr_t)scene,
                                                 (uint64_t)(uintptr_t)ws->buf.p,
                                                 (uint64_t)(uintptr_t)ws->aux.p,
                                                 launch_key(tc, ta),
                                                 launch_key_thresholds(tc, ta),
                                                 ((uint64_t)Ge << 32) | Gs,
                                                 ((uint64_t)Gw << 32) | G,
                                                 ((uint64_t)cap << 32) | P.pt.npix,
                                                 ((uint64_t)pass << 32) | C.K,
                                                 ((uint64_t)P.pt.max_depth << 32) | spill_entries,
                                                 view_hash(v),
                                                 side};
    return ws->bdpt_graph[side].launch(key, stream, side ? "BDPT light subpaths" : "BDPT camera subpaths", record);
  }

  // Connection pass pi of the chunk: enumerate the strategies and store the
  // weighted contributions, trace their shadow rays (any hit), gather the
  // unoccluded ones in enumeration order.
  void run_conn_pass(const BdptParams* Pd, uint32_t pi) {
    const WaveBufs& B = ws->B;
    int2* const spill = ws->spill;
    // the MIS walks' pdf pairs in LDS while they fit a workgroup's share
    const size_t pdf_lds_bytes = 2ull * Vt.V * kTraceBlock * sizeof(float2) <= kPdfLdsMax
                                     ? 2ull * Vt.V * kTraceBlock * sizeof(float2)
                                     : 0;
    const uint32_t Gg = std::min<uint32_t>((pass + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bdpt_conn_reset, dim3(1), dim3(64), 0, stream, C);
    if (pdf_lds_bytes)
      hipLaunchKernelGGL((k_bdpt_conn_prep<kTraceBlock, true>), dim3(G), dim3(kTraceBlock), pdf_lds_bytes, stream, v,
                         Pd, Vt, C, pi, B.stats);
    else
      hipLaunchKernelGGL((k_bdpt_conn_prep<kTraceBlock, false>), dim3(G), dim3(kTraceBlock), 0, stream, v, Pd, Vt, C,
                         pi, B.stats);
#define AR_VIS(BB, SS)                                                                                      \
  if (ta.workgroup_size == BB && ta.lds_stack == SS) {                                                      \
    if (ta.park_slots == 4)                                                                                 \
      hipLaunchKernelGGL((k_bdpt_conn_vis<BB, SS, 4>), dim3(Gs), dim3(BB), 0, stream, v.sv, C, spill,       \
                         ta.balance ? B.work : nullptr, ta.refill_min, ta.search_min);                      \
    else                                                                                                    \
      hipLaunchKernelGGL((k_bdpt_conn_vis<BB, SS, 3>), dim3(Gs), dim3(BB), 0, stream, v.sv, C, spill,       \
                         ta.balance ? B.work : nullptr, ta.refill_min, ta.search_min);                      \
  }
    AR_TRACE_VARIANTS(AR_VIS)
#undef AR_VIS
    hipLaunchKernelGGL(k_bdpt_conn_gather, dim3(Gg), dim3(256), 0, stream, Pd, C, pi, B.L, splat);
  }

  ar_status step(bool* done) override {
    constexpr int BLOCK = kTraceBlock;
    BdptParams* Pd = reinterpret_cast<BdptParams*>(ws->params);
    step_share = 1.0;
    if (s0 < P.pt.spp && !wavefront) {
      P.pt.s0 = s0;
      P.pt.k = std::min(k, P.pt.spp - s0);
      hipLaunchKernelGGL(k_bdpt_params, dim3(1), dim3(64), 0, stream, P, Pd);
      hipLaunchKernelGGL((k_bdpt_subpaths<BLOCK>), dim3(G), dim3(BLOCK), 0, stream, v, P, Vt, ws->spill,
                         ws->B.stats);
      hipLaunchKernelGGL((k_bdpt_connect<BLOCK>), dim3(G), dim3(BLOCK), 0, stream, v, P, Vt, ws->B.L, splat,
                         ws->spill, ws->B.stats);
      hipLaunchKernelGGL(k_film_accumulate_b, dim3(eg), dim3(256), 0, stream, ws->B.L, ws->acc, Pd);
      AR_HIP(hipGetLastError());
      s0 += P.pt.k;
    } else if (s0 < P.pt.spp) {
      // subpaths take about a third of a chunk, the passes share the rest
      const uint32_t paths = P.pt.npix * std::min(k, P.pt.spp - s0);
      const uint32_t npass = (paths + pass - 1) / pass;
      if (phase == 0) {
        P.pt.s0 = s0;
        P.pt.k = std::min(k, P.pt.spp - s0);
        hipLaunchKernelGGL(k_bdpt_params, dim3(1), dim3(64), 0, stream, P, Pd);
        ar_status e = run_subpaths(Pd, 0);
        if (e != AR_OK) return e;
        step_share = 1.0 / 6.0;
        phase = 1;
      } else if (phase == 1) {
        ar_status e = run_subpaths(Pd, 1);
        if (e != AR_OK) return e;
        step_share = 1.0 / 6.0;
        phase = 2;
      } else {
        const uint32_t pi = phase - 2;
        run_conn_pass(Pd, pi);
        step_share = 2.0 / 3.0 / npass;
        if (pi + 1 < npass) {
          phase++;
        } else {
          hipLaunchKernelGGL(k_film_accumulate_b, dim3(eg), dim3(256), 0, stream, ws->B.L, ws->acc, Pd);
          phase = 0;
          s0 += P.pt.k;
        }
      }
      AR_HIP(hipGetLastError());
    }
    *done = s0 >= P.pt.spp;
    return AR_OK;
  }

  double last_step_share() const override { return step_share; }

  ar_status finish() override {
    hipLaunchKernelGGL(k_bdpt_resolve, dim3(eg), dim3(256), 0, stream, ws->acc, (const long long*)splat, out_dev,
                       P.pt.npix, (float)P.pt.spp);
    AR_HIP(hipGetLastError());
    return AR_OK;
  }

  ar_status read_stats(ar_render_stats* st) override {
    unsigned long long hs[2] = {0, 0};
    AR_HIP(hipMemcpy(hs, ws->B.stats, 16, hipMemcpyDeviceToHost));
    st->paths = (uint64_t)P.pt.npix * P.pt.spp;
    st->rays = hs[0] + hs[1];
    st->shadow_rays = hs[1];
    return AR_OK;
  }
};

std::unique_ptr<JobRun> make_bdpt_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st,
                                      float* out_dev) {
  std::unique_ptr<BdptJob> j(new BdptJob());
  j->scene = s;
  j->spec = spec;
  j->ws = ws;
  j->stream = st;
  j->out_dev = out_dev;
  return j;
}

}  // namespace ar
