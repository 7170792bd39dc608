// Path tracer on gfx950 behind ar_render (PathTracer::render,
// agent-ray src/integrators/path_tracer.rs:159-191).
//
// Samples are processed in chunks of k samples per pixel; path id within a
// chunk = s_local * npix + pixel, so neighbouring lanes trace neighbouring
// pixels.  Every path writes its radiance to its own slot; a reduction adds
// the slots of each pixel in sample order, which reproduces the reference's
// per-pixel accumulation order exactly (film.add_sample(accum / spp)) and
// makes the film independent of chunking and schedule.
//
// Schedules:
//   megakernel  one thread per path, whole bounce loop inline.
//   wavefront   per bounce: extend (closest hit over the active queue) ->
//               shade (emission, next-event sample, BSDF sample, RR; pushes
//               survivors and shadow requests with wave-aggregated atomics)
//               -> shadow (any hit; adds the contribution if unoccluded).
//               Stage kernels are persistent (fixed grid, grid-stride over a
//               device-side count), so the host never reads queue sizes and
//               the whole bounce loop of a chunk is captured once as a HIP
//               graph and replayed per chunk.
#include "ar_kernels.h"
#include "ar_pt.h"
#include "ar_job.h"
#include "ar_wf.h"

#include <cstring>

namespace ar {



// ---------------------------------------------------------------------------
// Megakernel
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pt_mega(ShadeView v, PtParams P, float4* __restrict__ Lbuf,
                                                   int2* __restrict__ spill, unsigned long long* __restrict__ stats) {
  __shared__ int2 lds[kTraceLdsStack * BLOCK];
  const uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  const uint32_t nthreads = gridDim.x * BLOCK;
  Stack stk = make_stack<BLOCK>(lds, spill, gtid, nthreads);
  const uint32_t npaths = P.npix * P.k * P.variants;
  uint32_t nrays = 0;
  for (uint32_t p = gtid; p < npaths; p += nthreads) {
    PathState st;
    Rng rng;
    f3 L = splat(0.0f);
    if (gen_path(P, p, st, rng)) L = pt_li<true>(v, P, st, rng, stk, nrays);
    Lbuf[p] = make_float4(L.x, L.y, L.z, 0.0f);
  }
  wave_add_u64(stats, nrays);
}

// ---------------------------------------------------------------------------
// Wavefront stages
// ---------------------------------------------------------------------------

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_wf_generate(const PtParams* __restrict__ Pd, WaveBufs B) {
  const PtParams P = *Pd;
  const uint32_t npaths = P.npix * P.k * P.variants;
  const uint32_t nthreads = gridDim.x * BLOCK;
  const uint32_t rounds = (npaths + nthreads - 1) / nthreads;
  const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
  __shared__ uint32_t qsh[2][1][BLOCK / 64];
  uint32_t* const ctr[1] = {&B.cnt[0]};
  for (uint32_t r = 0; r < rounds; ++r) {
    uint32_t p = r * nthreads + gtid;
    bool live = false;
    if (p < npaths) {
      PathState s;
      Rng rng;
      live = gen_path(P, p, s, rng) && P.max_depth > 0;
      if (!live) {
        s.L = splat(0.0f);
        s.beta = splat(0.0f);
        s.prev_p = splat(0.0f);
        s.prev_pdf = 0.0f;
        s.depth = 0;
        s.dim = 0;
        s.ray.o = s.ray.d = splat(0.0f);
        s.ray.tmin = s.ray.tmax = 0.0f;
      }
      store_state(B, p, s);
    }
    const bool fl[1] = {live};
    uint32_t slot[1];
    block_push<BLOCK, 1>(qsh, r & 1, ctr, fl, slot);
    if (live) B.q[0][slot[0]] = p;
  }
}

// The chunk's parameters, written to device memory ahead of the (possibly
// graph-replayed) bounce loop, whose kernels read them there.
static_assert(sizeof(PtParams) <= kParamsBytes, "chunk parameters exceed their device block");
__global__ void k_wf_params(PtParams P, PtParams* dst) {
  if (threadIdx.x == 0) *dst = P;
}

// Queue counters start at zero for a chunk (a kernel, not a memset, so the
// chunk can be captured in a graph).
__global__ void k_wf_begin(WaveBufs B) {
  if (threadIdx.x < 3) B.cnt[threadIdx.x] = 0;
}


// At least 4 waves per SIMD (<= 128 VGPRs; 28 B of scratch per lane): two
// 512-thread workgroups per CU instead of one at the unconstrained 138
// VGPRs.  The room's 1920x1080x64 spp render measured 516 -> 509 ms;
// 256-thread workgroups at 3 waves per SIMD measured 523 ms.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void k_wf_shade(ShadeView v, const PtParams* __restrict__ Pd, WaveBufs B, int qi) {
  const PtParams P = *Pd;
  const uint32_t n = B.cnt[qi];
  const uint32_t* q = B.q[qi];
  uint32_t* qn = B.q[qi ^ 1];
  const uint32_t nthreads = gridDim.x * BLOCK;
  // Iterate whole waves together so the wave-aggregated pushes see every lane.
  uint32_t rounds = (n + nthreads - 1) / nthreads;
  uint32_t gtid = xcd_block() * BLOCK + threadIdx.x;
  __shared__ uint32_t qsh[2][2][BLOCK / 64];
  uint32_t* const ctr[2] = {&B.cnt[2], &B.cnt[qi ^ 1]};
  for (uint32_t r = 0; r < rounds; ++r) {
    uint32_t i = r * nthreads + gtid;
    bool active = i < n;
    bool cont = false;
    ShadowReq sh;
    sh.valid = false;
    uint32_t p = 0;
    PathState s;
    if (active) {
      p = q[i];
      load_state(B, p, s);
      float4 hv = B.hit[i];
      WorldHit h;
      h.t = hv.x;
      h.pos = __float_as_uint(hv.y);
      h.u = hv.z;
      h.v = hv.w;
      h.inst = B.hit_inst[i];
      Rng rng;
      rng.init(P.seed, s.pixel, s.sample, 0, s.dim);
      cont = pt_shade(v, P, s, h, rng, sh);
      s.dim = rng.dim;
      store_state(B, p, s);
    }
    const bool fl[2] = {active && sh.valid, active && cont};
    uint32_t slot[2];
    block_push<BLOCK, 2>(qsh, r & 1, ctr, fl, slot);
    const uint32_t si = slot[0];
    if (active && sh.valid) {
      B.sh_path[si] = p;
      B.sh_o[si] = make_float4(sh.ray.o.x, sh.ray.o.y, sh.ray.o.z, sh.ray.tmin);
      B.sh_d[si] = make_float4(sh.ray.d.x, sh.ray.d.y, sh.ray.d.z, sh.ray.tmax);
      B.sh_c[si] = make_float4(sh.c.x, sh.c.y, sh.c.z, 0.0f);
    }
    if (active && cont) qn[slot[1]] = p;
  }
}

__global__ void k_zero_stats(unsigned long long* st) {
  if (threadIdx.x < 2) st[threadIdx.x] = 0ull;
}

// End of bounce: the consumed queue and the shadow queue become empty.
__global__ void k_wf_reset(WaveBufs B, int qi) {
  if (threadIdx.x == 0) {
    B.cnt[qi] = 0;
    B.cnt[2] = 0;
  }
}

// ---------------------------------------------------------------------------
// Film: acc[pixel] += L[s * npix + pixel] for s in sample order.
// ---------------------------------------------------------------------------
__global__ void k_film_accumulate(const float4* __restrict__ L, float4* __restrict__ acc, uint32_t npix, uint32_t k,
                                  uint32_t first) {
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

__global__ void k_film_resolve(const float4* __restrict__ acc, float* __restrict__ out, uint32_t npix, float spp) {
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    float4 a = acc[px];
    out[3 * (size_t)px] = a.x / spp;
    out[3 * (size_t)px + 1] = a.y / spp;
    out[3 * (size_t)px + 2] = a.z / spp;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
ShadeView shade_view(const ar_scene* s) {
  ShadeView v;
  v.sv = s->view;
  v.objects = s->d_objects.as<DObject>();
  v.materials = s->d_materials.as<DMaterial>();
  v.lights = s->d_lights.as<DLight>();
  v.idx = s->d_idx.as<uint32_t>();
  v.nrm = s->d_nrm.as<float>();
  v.slots = s->d_slots.as<uint32_t>();
  v.objmats = s->d_objmats.as<uint32_t>();
  v.stri = s->d_stri.as<float4>();
  v.suv = s->has_uv ? s->d_suv.as<float4>() : nullptr;
  v.texels = s->d_tex.as<float4>();
  v.tex_info = s->d_texinfo.as<uint4>();
  v.n_tex = (uint32_t)s->tex_info.size();
  v.n_lights = s->n_lights;
  v.dist_mode = s->has_dist ? s->dist_mode : AR_LIGHTDIST_NONE;
  return v;
}

static uint32_t persistent_grid(int block, int per_cu) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return (uint32_t)(cus * per_cu / 8 * 8);
}

uint32_t stack_spill_entries(const ar_scene* s) {
  uint32_t need = s->stack_depth;
  return need > (uint32_t)kTraceLdsStack ? need - kTraceLdsStack : 0;
}

ar_status Workspace::reserve_paths(uint32_t cap, uint32_t npix, uint32_t spill_entries, uint32_t grid_threads,
                                   hipStream_t st) {
  size_t f4 = sizeof(float4);
  size_t bytes = 0;
  auto add = [&](size_t& off, size_t n) {
    off = bytes;
    bytes += (n + 255) / 256 * 256;
  };
  size_t o_ro, o_rd, o_b, o_L, o_pp, o_meta, o_q0, o_q1, o_hit, o_hi, o_shp, o_sho, o_shd, o_shc, o_cnt, o_st, o_work,
      o_acc, o_spill;
  // the walker counters first: their place must not move when the same
  // buffer is laid out again for another capacity (they stay zero between
  // launches and are cleared only when the buffer is allocated)
  add(o_work, kWalkWorkBytes);
  add(o_ro, cap * f4);
  add(o_rd, cap * f4);
  add(o_b, cap * f4);
  add(o_L, cap * f4);
  add(o_pp, cap * f4);
  add(o_meta, cap * sizeof(uint4));
  add(o_q0, cap * 4);
  add(o_q1, cap * 4);
  add(o_hit, cap * f4);
  add(o_hi, cap * 4);
  add(o_shp, cap * 4);
  add(o_sho, cap * f4);
  add(o_shd, cap * f4);
  add(o_shc, cap * f4);
  add(o_cnt, 64);
  add(o_st, 64);
  add(o_acc, (size_t)npix * f4 * 3);
  add(o_spill, (size_t)spill_entries * grid_threads * sizeof(int2));
  size_t o_par;
  add(o_par, kParamsBytes);
  if (buf.bytes < bytes) {
    // the buffer's last user is this stream (callers order other streams'
    // uses before it): drain it, not the whole device
    AR_HIP(hipStreamSynchronize(st));
    AR_HIP(buf.alloc(bytes));
    AR_HIP(clear_walk_work((char*)buf.p + o_work, st));
    pt_graph.reset();
    bdpt_graph[0].reset();
    bdpt_graph[1].reset();
  }
  char* b = (char*)buf.p;
  B.ray_o = (float4*)(b + o_ro);
  B.ray_d = (float4*)(b + o_rd);
  B.beta = (float4*)(b + o_b);
  B.L = (float4*)(b + o_L);
  B.prev_p = (float4*)(b + o_pp);
  B.meta = (uint4*)(b + o_meta);
  B.q[0] = (uint32_t*)(b + o_q0);
  B.q[1] = (uint32_t*)(b + o_q1);
  B.hit = (float4*)(b + o_hit);
  B.hit_inst = (uint32_t*)(b + o_hi);
  B.sh_path = (uint32_t*)(b + o_shp);
  B.sh_o = (float4*)(b + o_sho);
  B.sh_d = (float4*)(b + o_shd);
  B.sh_c = (float4*)(b + o_shc);
  B.cnt = (uint32_t*)(b + o_cnt);
  B.stats = (unsigned long long*)(b + o_st);
  B.work = (uint32_t*)(b + o_work);
  acc = (float4*)(b + o_acc);
  spill = (int2*)(b + o_spill);
  params = b + o_par;
  capacity = cap;
  return AR_OK;
}

void Workspace::trim() {
  pt_graph.reset();
  bdpt_graph[0].reset();
  bdpt_graph[1].reset();
  buf.release();
  out.release();
  aux.release();
  pois.release();
  B = WaveBufs{};
  acc = nullptr;
  spill = nullptr;
  params = nullptr;
  capacity = 0;
}

uint64_t view_hash(const ShadeView& v) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) {
    for (int i = 0; i < 8; ++i) h = (h ^ ((x >> (8 * i)) & 255u)) * 1099511628211ull;
  };
  const SceneView& sv = v.sv;
  for (const void* p : {(const void*)sv.nodes, (const void*)sv.qnodes, (const void*)sv.tris, (const void*)sv.qtris,
                        (const void*)sv.insts, (const void*)sv.qinsts,
                        (const void*)v.objects, (const void*)v.materials, (const void*)v.lights, (const void*)v.idx,
                        (const void*)v.nrm, (const void*)v.slots, (const void*)v.objmats, (const void*)v.stri,
                        (const void*)v.suv, (const void*)v.texels, (const void*)v.tex_info})
    mix((uint64_t)(uintptr_t)p);
  mix(v.n_tex);
  mix(((uint64_t)(uint32_t)sv.tlas_root << 32) | (uint32_t)sv.qtlas_root);
  for (int k = 0; k < 6; ++k) mix(bits_u(sv.tlas_box[k]));
  mix(((uint64_t)sv.n_inst << 32) | v.n_lights);
  mix(v.dist_mode);
  return h;
}

// The bounce loop of one chunk: begin, generate, then per bounce extend ->
// shade -> shadow -> reset.  Every kernel argument is fixed for a workspace
// (buffers, scene views, launch configuration); the chunk's parameters are
// written to device memory first (k_wf_params), so with use_graph the loop
// is captured once in a HIP graph and replayed for every chunk with one
// launch instead of 2 + 4 * max_depth.
static ar_status run_chunk_wavefront(const ar_scene* s, Workspace& ws, const ShadeView& v, const PtParams& P,
                                     hipStream_t st, bool use_graph) {
  // generate / shade: 512-thread workgroups (8 waves share one queue atomic),
  // as many threads per CU as the traversal stages' default
  constexpr int BLOCK = 512;
  const uint32_t G = persistent_grid(BLOCK, kTraceBlocksPerCU * kTraceBlock / BLOCK);
  const WaveBufs B = ws.B;
  const PtParams* Pd = reinterpret_cast<const PtParams*>(ws.params);
  // Traversal stages use the scene's (autotuned) launch configurations:
  // closest hit for extension rays, any hit for shadow rays.
  const ar_trace_config tc = s->tcfg, ta = s->tcfg_any;
  const uint32_t Gt = (uint32_t)(device_cus() * tc.blocks_per_cu / 8 * 8);
  const uint32_t Ga = (uint32_t)(device_cus() * ta.blocks_per_cu / 8 * 8);
  uint32_t Ge = 0, Gs = 0;  // resident-clamped grids, computed outside any capture
#define AR_GRIDS(BB, SS)                                                                                  \
  if (tc.workgroup_size == BB && tc.lds_stack == SS)                                                      \
    Ge = clamp_resident(Gt,                                                                               \
                        tc.park_slots == 4 ? (const void*)&k_wf_extend<BB, SS, 4>                         \
                                           : (const void*)&k_wf_extend<BB, SS, 3>,                        \
                        BB);                                                                              \
  if (ta.workgroup_size == BB && ta.lds_stack == SS)                                                      \
    Gs = clamp_resident(Ga,                                                                               \
                        ta.park_slots == 4 ? (const void*)&k_wf_shadow<BB, SS, 4>                         \
                                           : (const void*)&k_wf_shadow<BB, SS, 3>,                        \
                        BB);
  AR_TRACE_VARIANTS(AR_GRIDS)
#undef AR_GRIDS
  if (Ge == 0 || Gs == 0) return fail(AR_ERR_INVALID_ARG, "unsupported traversal configuration");
  uint32_t* const work_c = tc.balance ? B.work : nullptr;
  uint32_t* const work_a = ta.balance ? B.work : nullptr;
  hipLaunchKernelGGL(k_wf_params, dim3(1), dim3(64), 0, st, P, reinterpret_cast<PtParams*>(ws.params));
  auto record = [&]() {
    hipLaunchKernelGGL(k_wf_begin, dim3(1), dim3(64), 0, st, B);
    hipLaunchKernelGGL((k_wf_generate<BLOCK>), dim3(G), dim3(BLOCK), 0, st, Pd, B);
    for (uint32_t d = 0; d < P.max_depth; ++d) {
      int qi = d & 1;
#define AR_EXT(BB, SS)                                                                                       \
  if (tc.workgroup_size == BB && tc.lds_stack == SS) {                                                       \
    if (tc.park_slots == 4)                                                                                  \
      hipLaunchKernelGGL((k_wf_extend<BB, SS, 4>), dim3(Ge), dim3(BB), 0, st, v.sv, B, qi, ws.spill,         \
                         work_c, tc.refill_min, tc.search_min);                                              \
    else                                                                                                     \
      hipLaunchKernelGGL((k_wf_extend<BB, SS, 3>), dim3(Ge), dim3(BB), 0, st, v.sv, B, qi, ws.spill,         \
                         work_c, tc.refill_min, tc.search_min);                                              \
  }
      AR_TRACE_VARIANTS(AR_EXT)
#undef AR_EXT
      hipLaunchKernelGGL((k_wf_shade<BLOCK>), dim3(G), dim3(BLOCK), 0, st, v, Pd, B, qi);
#define AR_SHD(BB, SS)                                                                                       \
  if (ta.workgroup_size == BB && ta.lds_stack == SS) {                                                       \
    if (ta.park_slots == 4)                                                                                  \
      hipLaunchKernelGGL((k_wf_shadow<BB, SS, 4>), dim3(Gs), dim3(BB), 0, st, v.sv, B, ws.spill,             \
                         work_a, ta.refill_min, ta.search_min);                                              \
    else                                                                                                     \
      hipLaunchKernelGGL((k_wf_shadow<BB, SS, 3>), dim3(Gs), dim3(BB), 0, st, v.sv, B, ws.spill,             \
                         work_a, ta.refill_min, ta.search_min);                                              \
  }
      AR_TRACE_VARIANTS(AR_SHD)
#undef AR_SHD
      hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, st, B, qi);
    }
  };
  if (!use_graph) {
    record();
    AR_HIP(hipGetLastError());
    return AR_OK;
  }
  // what the graph depends on besides the chunk parameters
  const uint64_t key[GraphCache::kKeyWords] = {(uint64_t)(uintptr_t)s,
                                               (uint64_t)(uintptr_t)ws.buf.p,
                                               launch_key(tc, ta),
                                               launch_key_thresholds(tc, ta),
                                               ((uint64_t)Ge << 32) | Gs,
                                               ((uint64_t)G << 32) | P.max_depth,
                                               view_hash(v),
                                               (uint64_t)ws.capacity};
  return ws.pt_graph.launch(key, st, "bounce loop", record);
}

// ---------------------------------------------------------------------------
// Gradient-domain path tracer (GradientDomainPathTracer::render,
// gradient_domain_pt.rs:346-415): per pixel and sample, a base path and the
// x+1 / y+1 offset paths (variants 1, 2) that replay the base path's random
// numbers.  Accumulated in sample order: primal += base, dx += right - base,
// dy += down - base; an offset outside the image equals the base.
// ---------------------------------------------------------------------------
__global__ void k_gdpt_accumulate(const float4* __restrict__ L, float4* __restrict__ acc, uint32_t w, uint32_t h,
                                  uint32_t k, uint32_t first) {
  const uint32_t npix = w * h;
  const size_t per = (size_t)npix * k;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    uint32_t x = px % w, y = px / w;
    float4 a0 = first ? make_float4(0, 0, 0, 0) : acc[px];
    float4 a1 = first ? make_float4(0, 0, 0, 0) : acc[npix + px];
    float4 a2 = first ? make_float4(0, 0, 0, 0) : acc[2 * (size_t)npix + px];
    for (uint32_t s = 0; s < k; ++s) {
      size_t i = (size_t)s * npix + px;
      float4 b = L[i];
      float4 r = x + 1 < w ? L[per + i] : b;
      float4 d = y + 1 < h ? L[2 * per + i] : b;
      a0.x = a0.x + b.x;
      a0.y = a0.y + b.y;
      a0.z = a0.z + b.z;
      a1.x = a1.x + (r.x - b.x);
      a1.y = a1.y + (r.y - b.y);
      a1.z = a1.z + (r.z - b.z);
      a2.x = a2.x + (d.x - b.x);
      a2.y = a2.y + (d.y - b.y);
      a2.z = a2.z + (d.z - b.z);
    }
    acc[px] = a0;
    acc[npix + px] = a1;
    acc[2 * (size_t)npix + px] = a2;
  }
}

// films (interleaved RGB): primal, dx, dy = accum * (1/spp)
__global__ void k_gdpt_resolve(const float4* __restrict__ acc, float* __restrict__ films, uint32_t npix, float inv_spp) {
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x)
    for (int f = 0; f < 3; ++f) {
      float4 a = acc[(size_t)f * npix + px];
      float* o = films + (size_t)f * npix * 3 + 3 * (size_t)px;
      o[0] = a.x * inv_spp;
      o[1] = a.y * inv_spp;
      o[2] = a.z * inv_spp;
    }
}

// ---------------------------------------------------------------------------
// Job interface: PT (variants = 1) and GDPT (variants = 3).
// ---------------------------------------------------------------------------
class PtJob : public JobRun {
 public:
  PtParams P{};
  ShadeView v{};
  bool gdpt = false, mega = false;
  uint32_t k = 1, s0 = 0, G = 0, eg = 0;
  float* films = nullptr;

  ar_status begin() override {
    constexpr int BLOCK = kTraceBlock;
    const ar_render_config* cfg = &spec.cfg;
    const uint32_t w = spec.width, h = spec.height, npix = w * h;
    gdpt = cfg->integrator == AR_INTEGRATOR_GDPT;
    mega = cfg->schedule == AR_SCHED_MEGAKERNEL;
    v = shade_view(scene);
    P.width = w;
    P.height = h;
    P.spp = cfg->spp;
    P.max_depth = cfg->max_depth;
    P.rr_depth = cfg->rr_depth;
    P.npix = npix;
    P.seed = cfg->seed;
    P.cam = cam_from(spec.camera);
    P.variants = gdpt ? 3 : 1;
    k = chunk_samples(spec);
    G = persistent_grid(BLOCK, kTraceBlocksPerCU);
    eg = std::min<uint32_t>((npix + 255) / 256, 4096);
    // Spill area: covers the megakernel's (128, 16) launch and the traversal
    // stages' tuned launches (closest hit, any hit), whichever runs.
    uint32_t need = scene->stack_depth;
    uint64_t spill_words = stack_spill_entries(scene) * (uint64_t)G * BLOCK;
    for (const ar_trace_config* tc : {&scene->tcfg, &scene->tcfg_any}) {
      uint64_t gt = (uint64_t)(device_cus() * tc->blocks_per_cu / 8 * 8) * tc->workgroup_size;
      uint64_t e_t = need > tc->lds_stack ? need - tc->lds_stack : 0;
      spill_words = std::max<uint64_t>(spill_words, e_t * gt);
    }
    ar_status e = ws->reserve_paths(P.variants * npix * k, npix, (uint32_t)((spill_words + G * BLOCK - 1) / (G * BLOCK)),
                                    G * BLOCK, stream);
    if (e != AR_OK) return e;
    if (gdpt) {
      films = aux_dev;
      size_t film_bytes = (size_t)npix * 3 * sizeof(float);
      if (!films) {
        if (ws->aux.bytes < 3 * film_bytes) {
          AR_HIP(hipStreamSynchronize(stream));
          AR_HIP(ws->aux.alloc(3 * film_bytes));
        }
        films = ws->aux.as<float>();
      }
    }
    hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, stream, ws->B.stats);
    s0 = 0;
    return AR_OK;
  }

  ar_status step(bool* done) override {
    constexpr int BLOCK = kTraceBlock;
    if (s0 < P.spp) {
      P.s0 = s0;
      P.k = std::min(k, P.spp - s0);
      if (mega) {
        hipLaunchKernelGGL((k_pt_mega<BLOCK>), dim3(G), dim3(BLOCK), 0, stream, v, P, ws->B.L, ws->spill,
                           ws->B.stats);
      } else {
        ar_status e = run_chunk_wavefront(scene, *ws, v, P, stream, true);
        if (e != AR_OK) return e;
      }
      if (gdpt)
        hipLaunchKernelGGL(k_gdpt_accumulate, dim3(eg), dim3(256), 0, stream, ws->B.L, ws->acc, P.width, P.height,
                           P.k, s0 == 0 ? 1u : 0u);
      else
        hipLaunchKernelGGL(k_film_accumulate, dim3(eg), dim3(256), 0, stream, ws->B.L, ws->acc, P.npix, P.k,
                           s0 == 0 ? 1u : 0u);
      AR_HIP(hipGetLastError());
      s0 += P.k;
    }
    *done = s0 >= P.spp;
    return AR_OK;
  }

  ar_status finish() override {
    if (!gdpt) {
      hipLaunchKernelGGL(k_film_resolve, dim3(eg), dim3(256), 0, stream, ws->acc, out_dev, P.npix, (float)P.spp);
      AR_HIP(hipGetLastError());
      return AR_OK;
    }
    hipLaunchKernelGGL(k_gdpt_resolve, dim3(eg), dim3(256), 0, stream, ws->acc, films, P.npix, 1.0f / (float)P.spp);
    AR_HIP(hipGetLastError());
    const ar_render_config* cfg = &spec.cfg;
    return poisson_solve_device(films, films + (size_t)P.npix * 3, films + (size_t)P.npix * 6, P.width, P.height,
                                cfg->alpha, cfg->poisson_iterations, cfg->poisson_sign, out_dev, nullptr, ws->pois,
                                stream);
  }

  ar_status read_stats(ar_render_stats* st) override {
    unsigned long long hs[2] = {0, 0};
    AR_HIP(hipMemcpy(hs, ws->B.stats, 16, hipMemcpyDeviceToHost));
    st->paths = (uint64_t)P.variants * P.npix * P.spp;
    st->rays = hs[0] + hs[1];
    st->shadow_rays = mega ? 0 : hs[1];
    return AR_OK;
  }
};

std::unique_ptr<JobRun> make_pt_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st,
                                    float* out_dev, float* aux_dev) {
  std::unique_ptr<PtJob> j(new PtJob());
  j->scene = s;
  j->spec = spec;
  j->ws = ws;
  j->stream = st;
  j->out_dev = out_dev;
  j->aux_dev = aux_dev;
  return j;
}

}  // namespace ar
