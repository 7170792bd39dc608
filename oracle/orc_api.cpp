// Oracle (test infrastructure only): flat C entry points used by the Python
// test harness (oracle/oracle.py) and by bench.py's CPU-baseline leg.
#include <chrono>
#include <cstdio>

#include "orc_render.h"

using namespace orc;

namespace {
V3 v3(const float* p) { return V3(p[0], p[1], p[2]); }
Camera cam_from(const float* c) {
  Camera k;
  k.origin = v3(c);
  k.llc = v3(c + 3);
  k.horiz = v3(c + 6);
  k.vert = v3(c + 9);
  k.forward = v3(c + 12);
  k.area = c[15];
  return k;
}
Ray ray_from(const float* r) {
  Ray k;
  k.o = v3(r);
  k.tmin = r[3];
  k.d = v3(r + 4);
  k.tmax = r[7];
  return k;
}
}  // namespace

extern "C" {

void* orc_scene_new() { return new Scene(); }
void orc_scene_free(void* s) { delete (Scene*)s; }

// params: albedo[3], roughness, eta, emission[3]
int orc_scene_add_material(void* s, uint32_t type, const float* p) {
  Scene* sc = (Scene*)s;
  Material m;
  m.type = type;
  m.albedo = v3(p);
  m.roughness = p[3];
  m.eta = p[4];
  m.emission = v3(p + 5);
  sc->materials.push_back(m);
  return (int)sc->materials.size() - 1;
}

// ImageTexture::load from decoded 8-bit RGB (texture.rs:74-93).
int orc_scene_add_texture(void* s, const uint8_t* rgb, uint32_t w, uint32_t h) {
  Scene* sc = (Scene*)s;
  sc->textures.emplace_back();
  sc->textures.back().load_rgb8(rgb, w, h);
  return (int)sc->textures.size() - 1;
}

// Lambertian::with_texture (lambertian.rs:28-33): albedo from texture `tex`.
void orc_scene_set_material_texture(void* s, uint32_t mat, int tex) { ((Scene*)s)->materials[mat].tex = tex; }

// TriangleMesh.tex_coords of mesh object `obj` (mesh.rs:10).
void orc_scene_set_mesh_uv(void* s, uint32_t obj, const float* uv, uint32_t nv) {
  ((Scene*)s)->objects[obj]->mesh->uv.assign(uv, uv + 2 * (size_t)nv);
}

// Texture sampling for the tests: n (u, v) pairs -> n linear RGB.
void orc_texture_sample(void* s, uint32_t tex, const float* uv, uint32_t n, float* out) {
  const Texture& t = ((Scene*)s)->textures[tex];
  for (uint32_t i = 0; i < n; ++i) {
    V3 c = t.sample(uv[2 * i], uv[2 * i + 1]);
    out[3 * i] = c.x, out[3 * i + 1] = c.y, out[3 * i + 2] = c.z;
  }
}

float orc_srgb_to_linear(float x) { return srgb_to_linear(x); }

int orc_scene_add_mesh(void* s, const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt,
                       const float* nrm, const uint32_t* slots, uint32_t nslots, const uint32_t* mats,
                       uint32_t nmats, const float* xf16) {
  Scene* sc = (Scene*)s;
  auto mesh = std::make_shared<Mesh>();
  mesh->pos.resize(nv);
  for (uint32_t i = 0; i < nv; ++i) mesh->pos[i] = v3(pos + 3 * i);
  if (nrm) {
    mesh->nrm.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) mesh->nrm[i] = v3(nrm + 3 * i);
  }
  mesh->idx.assign(idx, idx + 3 * (size_t)nt);
  mesh->slots.assign(slots, slots + nslots);
  std::unique_ptr<Object> o(new Object());
  o->mesh = mesh;
  if (xf16) {
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r) o->xf.c[c][r] = xf16[4 * c + r];
  } else {
    o->xf = M4::identity();
  }
  o->mats.assign(mats, mats + nmats);
  sc->objects.push_back(std::move(o));
  return (int)sc->objects.size() - 1;
}

// params: pos[3], intensity[3], eu[3], ev[3].  A quad light also adds its
// emissive geometry (two triangles, normal = cross(eu, ev)) as an object.
int orc_scene_add_light(void* s, uint32_t type, const float* p) {
  Scene* sc = (Scene*)s;
  Light L;
  L.type = type;
  L.pos = v3(p);
  L.intensity = v3(p + 3);
  L.eu = v3(p + 6);
  L.ev = v3(p + 9);
  int id = (int)sc->lights.size();
  if (type == LIGHT_QUAD) {
    V3 c = cross(L.eu, L.ev);
    L.area = length(c);
    L.n = normalize(c);
    Material m;
    m.type = MAT_LAMBERT;
    m.albedo = V3();
    m.emission = L.intensity;
    m.light_id = id;
    sc->materials.push_back(m);
    uint32_t mid = (uint32_t)sc->materials.size() - 1;
    V3 a = L.pos, b = L.pos + L.eu, cc = (L.pos + L.eu) + L.ev, d = L.pos + L.ev;
    float pos[12] = {a.x, a.y, a.z, b.x, b.y, b.z, cc.x, cc.y, cc.z, d.x, d.y, d.z};
    uint32_t idx[6] = {0, 1, 2, 0, 2, 3};
    uint32_t slot = 0;
    orc_scene_add_mesh(s, pos, 4, idx, 2, nullptr, &slot, 1, &mid, 1, nullptr);
  }
  sc->lights.push_back(L);
  return id;
}

int orc_scene_commit(void* s, int dist_mode) {
  ((Scene*)s)->commit(dist_mode);
  return 0;
}

// BVH dump for topology comparison: returns node count; when buffers are
// non-null writes boxes (6 floats/node), meta (left_or_start, right_or_end,
// count) and the primitive order.  obj = -1 selects the TLAS.
uint32_t orc_bvh_dump(void* s, int obj, float* boxes, uint32_t* meta, uint32_t* prims) {
  Scene* sc = (Scene*)s;
  const BVH& b = obj < 0 ? sc->tlas.bvh : sc->objects[obj]->blas->bvh;
  if (boxes) {
    for (size_t i = 0; i < b.nodes.size(); ++i) {
      const BVHNode& n = b.nodes[i];
      float* o = boxes + 6 * i;
      o[0] = n.box.lo.x; o[1] = n.box.lo.y; o[2] = n.box.lo.z;
      o[3] = n.box.hi.x; o[4] = n.box.hi.y; o[5] = n.box.hi.z;
      meta[3 * i] = n.left_or_start;
      meta[3 * i + 1] = n.right_or_end;
      meta[3 * i + 2] = n.count;
    }
    for (size_t i = 0; i < b.prims.size(); ++i) prims[i] = b.prims[i];
  }
  return (uint32_t)b.nodes.size();
}

// rays: 8 floats each (o.xyz, tmin, d.xyz, tmax), directions as given.
// Closest hit: t, prim (UINT32_MAX on miss), inst.  counters[0..1] += node
// visits and triangle tests.  Returns wall seconds.
// Light selection (lights.rs:115-235) of a committed scene: pmf and cdf
// (n entries each, NULL to skip); returns the light count (0 without a
// distribution).
uint32_t orc_light_dist(void* s, float* pmf, float* cdf) {
  const Scene* sc = (const Scene*)s;
  if (!sc->has_dist) return 0;
  const uint32_t n = (uint32_t)sc->dist.pmf.size();
  for (uint32_t i = 0; i < n; ++i) {
    if (pmf) pmf[i] = sc->dist.pmf[i];
    if (cdf) cdf[i] = sc->dist.cdf[i];
  }
  return n;
}

// LightDistribution::sample_index (lights.rs:147-150, 219-226).
uint32_t orc_light_sample_index(void* s, float u, float* pmf) {
  uint32_t idx = 0;
  float p = 0.0f;
  ((const Scene*)s)->dist.sample_index(u, &idx, &p);
  if (pmf) *pmf = p;
  return idx;
}

double orc_trace(void* s, const float* rays, uint64_t n, float* t, uint32_t* prim, uint32_t* inst,
                 uint64_t* counters, int nthreads) {
  Scene* sc = (Scene*)s;
  auto t0 = std::chrono::steady_clock::now();
  const uint64_t chunk = 4096;
  uint32_t nchunks = (uint32_t)((n + chunk - 1) / chunk);
  std::atomic<uint64_t> cn(0), ct(0);
  parallel_rows(nchunks, nthreads, [&](uint32_t ch) {
    TraceCounters c;
    uint64_t b = (uint64_t)ch * chunk, e = std::min<uint64_t>(n, b + chunk);
    for (uint64_t i = b; i < e; ++i) {
      Ray r = ray_from(rays + 8 * i);
      RayHit h;
      if (sc->tlas.intersect(r, &h, &c)) {
        t[i] = h.t;
        prim[i] = h.prim;
        if (inst) inst[i] = h.inst;
      } else {
        t[i] = INFINITY;
        prim[i] = 0xffffffffu;
        if (inst) inst[i] = 0xffffffffu;
      }
    }
    cn += c.nodes;
    ct += c.tris;
  });
  if (counters) {
    counters[0] += cn.load();
    counters[1] += ct.load();
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_occluded(void* s, const float* rays, uint64_t n, uint32_t* occ, uint64_t* counters,
                    int nthreads) {
  Scene* sc = (Scene*)s;
  auto t0 = std::chrono::steady_clock::now();
  const uint64_t chunk = 4096;
  uint32_t nchunks = (uint32_t)((n + chunk - 1) / chunk);
  std::atomic<uint64_t> cn(0), ct(0);
  parallel_rows(nchunks, nthreads, [&](uint32_t ch) {
    TraceCounters c;
    uint64_t b = (uint64_t)ch * chunk, e = std::min<uint64_t>(n, b + chunk);
    for (uint64_t i = b; i < e; ++i) occ[i] = sc->occluded(ray_from(rays + 8 * i), &c) ? 1u : 0u;
    cn += c.nodes;
    ct += c.tris;
  });
  if (counters) {
    counters[0] += cn.load();
    counters[1] += ct.load();
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Ray::new (geometry.rs:12-20)
void orc_ray_new(const float* o, const float* d, float tmin, float tmax, float* out8) {
  Ray r = Ray::make(v3(o), v3(d), tmin, tmax);
  out8[0] = r.o.x; out8[1] = r.o.y; out8[2] = r.o.z; out8[3] = r.tmin;
  out8[4] = r.d.x; out8[5] = r.d.y; out8[6] = r.d.z; out8[7] = r.tmax;
}

// Triangle::hit for known-answer tests: returns 1 on hit with (t, u, v, n).
int orc_triangle_hit(const float* v0, const float* v1, const float* v2, const float* ray8, float* out6) {
  Ray r = ray_from(ray8);
  HitInfo h;
  V3 a = v3(v0);
  if (!triangle_hit(a, v3(v1) - a, v3(v2) - a, r, &h)) return 0;
  out6[0] = h.t; out6[1] = h.u; out6[2] = h.v;
  out6[3] = h.n.x; out6[4] = h.n.y; out6[5] = h.n.z;
  return 1;
}

int orc_aabb_hit(const float* lo, const float* hi, const float* ray8) {
  AABB b{v3(lo), v3(hi)};
  return b.intersect(ray_from(ray8)) ? 1 : 0;
}

void orc_camera_lookat(const float* from, const float* at, const float* up, float vfov, float aspect,
                       float* cam16) {
  Camera c = Camera::lookat(v3(from), v3(at), v3(up), vfov, aspect);
  V3 vals[5] = {c.origin, c.llc, c.horiz, c.vert, c.forward};
  for (int i = 0; i < 5; ++i) {
    cam16[3 * i] = vals[i].x;
    cam16[3 * i + 1] = vals[i].y;
    cam16[3 * i + 2] = vals[i].z;
  }
  cam16[15] = c.area;
}

void orc_camera_ray(const float* cam16, float u, float v, float* ray8) {
  Ray r = cam_from(cam16).generate_ray(u, v);
  ray8[0] = r.o.x; ray8[1] = r.o.y; ray8[2] = r.o.z; ray8[3] = r.tmin;
  ray8[4] = r.d.x; ray8[5] = r.d.y; ray8[6] = r.d.z; ray8[7] = r.tmax;
}

// sample_we for closed-form tests: out = px, py, wi.xyz, dist, we, pdf.
int orc_camera_sample_we(const float* cam16, const float* p, uint32_t w, uint32_t h, float* out8) {
  Camera::WeSample s;
  if (!cam_from(cam16).sample_we(v3(p), w, h, &s)) return 0;
  out8[0] = s.px; out8[1] = s.py; out8[2] = s.wi.x; out8[3] = s.wi.y; out8[4] = s.wi.z;
  out8[5] = s.dist; out8[6] = s.we; out8[7] = s.pdf;
  return 1;
}

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

float orc_counter_sample(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t stream, uint32_t dim) {
  CounterSampler cs(seed, pixel, sample, stream);
  cs.dim = dim;
  return cs.next_1d();
}

void orc_ref_sampler(uint32_t x, uint32_t y, uint32_t n, float* out) {
  RefSampler r(RefSampler::pixel_seed(x, y));
  for (uint32_t i = 0; i < n; ++i) out[i] = r.next_1d();
}

// BSDF hooks for closed-form / chi-square tests.  params as in add_material.
static Material mat_from(uint32_t type, const float* p) {
  Material m;
  m.type = type;
  m.albedo = v3(p);
  m.roughness = p[3];
  m.eta = p[4];
  return m;
}
void orc_bsdf_eval(uint32_t type, const float* p, const float* wi, const float* wo, float* out3) {
  V3 f = bsdf_eval(mat_from(type, p), v3(wi), v3(wo));
  out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
float orc_bsdf_pdf(uint32_t type, const float* p, const float* wi, const float* wo) {
  return bsdf_pdf(mat_from(type, p), v3(wi), v3(wo));
}
// out7: wi.xyz, pdf, f.xyz
int orc_bsdf_sample(uint32_t type, const float* p, const float* wo, float u_sel, float ux, float uy,
                    float* out7) {
  BsdfSample bs;
  if (!bsdf_sample(mat_from(type, p), v3(wo), u_sel, ux, uy, &bs)) return 0;
  out7[0] = bs.wi.x; out7[1] = bs.wi.y; out7[2] = bs.wi.z; out7[3] = bs.pdf;
  out7[4] = bs.f.x; out7[5] = bs.f.y; out7[6] = bs.f.z;
  return 1;
}

// Path tracer render (path_tracer.rs:160-191).  out_mean = accum/spp per
// pixel (RGB), out_m2 (optional) = mean of squared per-sample radiance.
// rng_mode 0: per-pixel reference stream; 1: counter sampler (seed keyed).
// Returns wall seconds; counters: node visits, tri tests, rays traced.
// pt_mode / only_len: test-only estimator switches (PtConfig).
double orc_render_pt_ex(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp,
                        uint32_t max_depth, uint32_t rr_depth, int rng_mode, uint64_t seed, int nthreads,
                        uint32_t row_begin, uint32_t row_end, uint32_t col_begin, uint32_t col_end, int pt_mode,
                        uint32_t only_len, float* out_mean, float* out_m2, uint64_t* counters) {
  Scene* sc = (Scene*)s;
  Camera cam = cam_from(cam16);
  PtConfig cfg;
  cfg.spp = spp;
  cfg.max_depth = max_depth;
  cfg.rr_depth = rr_depth;
  cfg.pt_mode = pt_mode;
  cfg.only_len = only_len;
  if (row_end == 0 || row_end > h) row_end = h;
  if (col_end == 0 || col_end > w) col_end = w;
  auto t0 = std::chrono::steady_clock::now();
  std::atomic<uint64_t> cn(0), ct(0);
  parallel_rows(row_end - row_begin, nthreads, [&](uint32_t yy) {
    uint32_t y = row_begin + yy;
    TraceCounters c;
    for (uint32_t x = col_begin; x < col_end; ++x) {
      RefSampler ref(RefSampler::pixel_seed(x, y));
      V3 acc, acc2;
      for (uint32_t si = 0; si < spp; ++si) {
        CounterSampler cs(seed, y * w + x, si, 0);
        Sampler& smp = rng_mode == RNG_REF ? (Sampler&)ref : (Sampler&)cs;
        float jx = smp.next_1d(), jy = smp.next_1d();
        float u = ((float)x + jx) / (float)w;
        float v = 1.0f - ((float)y + jy) / (float)h;
        V3 L = pt_li(*sc, cam.generate_ray(u, v), smp, cfg, &c);
        acc = acc + L;
        acc2 = acc2 + L * L;
      }
      size_t o = 3 * ((size_t)y * w + x);
      V3 mean = acc / (float)spp;
      out_mean[o] = mean.x; out_mean[o + 1] = mean.y; out_mean[o + 2] = mean.z;
      if (out_m2) {
        V3 m2 = acc2 / (float)spp;
        out_m2[o] = m2.x; out_m2[o + 1] = m2.y; out_m2[o + 2] = m2.z;
      }
    }
    cn += c.nodes;
    ct += c.tris;
  });
  if (counters) {
    counters[0] += cn.load();
    counters[1] += ct.load();
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_render_pt(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                     uint32_t rr_depth, int rng_mode, uint64_t seed, int nthreads, uint32_t row_begin,
                     uint32_t row_end, float* out_mean, float* out_m2, uint64_t* counters) {
  return orc_render_pt_ex(s, cam16, w, h, spp, max_depth, rr_depth, rng_mode, seed, nthreads, row_begin, row_end, 0,
                          0, 0, 0, out_mean, out_m2, counters);
}

}  // extern "C"
