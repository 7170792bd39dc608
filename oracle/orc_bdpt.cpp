// Oracle (test infrastructure only): bidirectional path tracer restated
// from the reference's src/integrators/bidirectional_path_tracer.rs.
//
// Reference behaviour kept as is (bug-compatible): s = number of camera
// vertices and t = number of light vertices; t = 0 and s = 0 strategies are
// not implemented and are excluded from the MIS sums; the point-light t = 1
// contribution omits 1/sel_pmf (:947-967); surface vertices are never delta
// (:268); the camera vertex "normal" is the first ray direction (:350); light
// tracing (s = 1) multiplies by We = 1 / (A cos^4) without the cosine at the
// camera (:1104-1108, cameras.rs:229), which brightens it by 1 / cos(theta).
// Extension (quad lights, which the reference does not have): the t = 1
// connection to a quad light uses the light subpath's sampled emission point
// with the general formula (f_y = one-sided emission test), because the
// reference's re-sampling of the light at u = (0, 0) only makes sense for a
// point light.  The t = 0 strategy the reference leaves as a TODO for area
// lights (:810-818) is built: a camera subpath vertex on the front face of an
// emitter contributes beta * Le, weighted by the reference's own walk
// (mis_weight) with the pdfs the TODO names: the emitter vertex's reverse pdf
// is the light-origin pdf (sel_pmf / area, area measure) and the vertex
// before it gets the emission pdf toward it (cos / pi, converted to area).
// For area lights the light-side walk also counts the t' = 0 alternative
// (:735-742 skip it "for delta lights").
//
// Test-only switches (pins of this restatement, not reference behaviour):
// one strategy at a time (only_s, only_t), unweighted contributions, and
// mis_mode 2, a textbook power-heuristic walk (each vertex's reverse pdf
// taken once, the pdfs at the connection recomputed, the light-tracing
// strategy's camera pdf in area measure and its camera cosine, 1/sel_pmf on
// point-light t = 1)
// whose weights sum to one over the strategies that can make a path, so
// that its image equals the path tracer's in expectation.
#include <chrono>
#include <mutex>

#include "orc_render.h"

namespace orc {

enum VType { V_CAMERA = 0, V_LIGHT = 1, V_SURFACE = 2 };

struct PVertex {
  VType type;
  V3 p, n;
  Onb onb;
  Material mat;  // the surface's material, textured albedo resolved
  V3 beta;
  float pdf_fwd = 0, pdf_rev = 0;
  bool is_delta = false;
  int light = -1;     // light vertex: its light
  int emit = -1;      // camera-side surface vertex on the front face of an emitter: the emitter's light
  bool connectible() const { return type != V_SURFACE || !is_delta; }  // :281-292
};

// :305-314
static float pdf_sa_to_area(float pdf, V3 from, V3 to, V3 n_to) {
  V3 d = to - from;
  float d2 = length_sq(d);
  if (d2 < 1e-12f) return 0.0f;
  V3 dir = d / std::sqrt(d2);
  float c = std::fabs(dot(n_to, dir));
  return pdf * c / d2;
}

// :318-328
static float geometry_term(V3 pa, V3 na, V3 pb, V3 nb) {
  V3 d = pb - pa;
  float d2 = length_sq(d);
  if (d2 < 1e-12f) return 0.0f;
  V3 dir = d / std::sqrt(d2);
  float ca = std::fabs(dot(na, dir));
  float cb = std::fabs(dot(nb, -dir));
  return ca * cb / d2;
}

struct BdptCfg {
  uint32_t max_depth, rr_depth;
  int mis_mode;  // 0 power, 1 uniform, 2 textbook power walk (test only)
  float beta;
  int only_s = -1, only_t = -1;  // test only: evaluate this strategy alone (-1: all)
  bool unweighted = false;       // test only: MIS weight 1
};

// Shared bounce loop of both subpaths (:368-434, :498-556).
static void extend_subpath(const Scene& sc, Ray ray, V3 throughput, float prev_pdf_dir, const BdptCfg& cfg,
                           Sampler& smp, std::vector<PVertex>& v, bool camera_side, TraceCounters* c) {
  for (uint32_t depth = 0; depth < cfg.max_depth; ++depth) {
    float t;
    ShadingPoint hit;
    if (!sc.intersect(ray, &t, &hit, c)) break;
    PVertex sv;
    sv.type = V_SURFACE;
    sv.p = hit.p;
    sv.n = hit.n;
    sv.onb = hit.onb;
    sv.mat = *hit.mat;
    sv.beta = throughput;
    sv.pdf_fwd = pdf_sa_to_area(prev_pdf_dir, v.back().p, hit.p, hit.n);
    // t = 0 (extension): a camera subpath reaching an emitter's front face
    // records it with the reverse pdfs of the light side generating it: the
    // emitter point (light-origin pdf, area measure) and, from there, the
    // vertex before it (emission pdf, converted to area).  Only with a light
    // distribution (otherwise no light subpath exists).
    if (camera_side && hit.mat->emissive() && dot(hit.ng_world, -ray.d) > 0.0f) {
      sv.emit = hit.mat->light_id;
      if (sc.has_dist) {
        const Light& L = sc.lights[sv.emit];
        sv.pdf_rev = sc.dist.pmf[sv.emit] / L.area;
        PVertex& pv = v.back();
        V3 to_prev = normalize(pv.p - hit.p);
        pv.pdf_rev = pdf_sa_to_area(L.pdf_emission_dir(to_prev), hit.p, pv.p, pv.n);
      }
    }
    v.push_back(sv);
    V3 wo = -ray.d;
    V3 wo_l = hit.onb.to_local(wo);
    float u_sel = smp.next_1d();
    float ux = smp.next_1d(), uy = smp.next_1d();
    BsdfSample bs;
    if (!bsdf_sample(*hit.mat, wo_l, u_sel, ux, uy, &bs)) break;
    if (bs.pdf == 0.0f || is_zero(bs.f)) break;
    float cos_wi = std::fabs(bs.wi.z);
    throughput = throughput * (bs.f * cos_wi / bs.pdf);
    prev_pdf_dir = bs.pdf;
    float pdf_rev_dir = bsdf_pdf(*hit.mat, wo_l, bs.wi);
    size_t pi = v.size() - 2;
    v[pi].pdf_rev = pdf_sa_to_area(pdf_rev_dir, hit.p, v[pi].p, v[pi].n);
    if (depth + 1 >= cfg.rr_depth) {
      float surv = std::fmin(max_element(throughput), 0.95f);
      if (smp.next_1d() > surv || surv <= 0.0f) break;
      throughput = throughput / surv;
    }
    V3 wi_w = hit.onb.to_world(bs.wi);
    V3 off = hit.n * 1e-4f * rs_signum(dot(wi_w, hit.n));
    ray = Ray::make(hit.p + off, wi_w, 1e-4f, kF32Max);
  }
}

// :338-437
static void camera_subpath(const Scene& sc, const Camera& cam, Ray ray, const BdptCfg& cfg, Sampler& smp,
                           std::vector<PVertex>& v, TraceCounters* c) {
  PVertex cv;
  cv.type = V_CAMERA;
  cv.p = cam.origin;
  cv.n = ray.d;  // :350 first ray direction as the camera "normal"
  cv.onb = Onb::from_normal(ray.d);
  cv.beta = V3(1, 1, 1);
  cv.pdf_fwd = 1.0f;
  cv.pdf_rev = 0.0f;
  cv.is_delta = true;
  float cam_pdf_dir = cam.pdf_we(ray);
  v.push_back(cv);
  extend_subpath(sc, ray, V3(1, 1, 1), cam_pdf_dir, cfg, smp, v, true, c);
}

// :443-559
static void light_subpath(const Scene& sc, const BdptCfg& cfg, Sampler& smp, std::vector<PVertex>& v,
                          TraceCounters* c) {
  if (!sc.has_dist) return;
  float u_sel = smp.next_1d();
  uint32_t li;
  float sel_pmf;
  sc.dist.sample_index(u_sel, &li, &sel_pmf);
  const Light& L = sc.lights[li];
  float upx = smp.next_1d(), upy = smp.next_1d();
  float udx = smp.next_1d(), udy = smp.next_1d();
  EmissionSample em;
  if (!L.sample_emission(upx, upy, udx, udy, &em)) return;
  PVertex lv;
  lv.type = V_LIGHT;
  lv.p = em.ray.o;
  lv.n = em.n_light;
  lv.onb = Onb::from_normal(em.n_light);
  lv.pdf_fwd = em.pdf_pos * sel_pmf;
  lv.pdf_rev = 0.0f;
  lv.is_delta = L.is_delta();
  lv.light = (int)li;
  lv.beta = em.le / (em.pdf_pos * sel_pmf);
  v.push_back(lv);
  if (cfg.max_depth == 0) return;
  V3 throughput = em.le / (em.pdf_pos * sel_pmf * em.pdf_dir);
  float cos_emit = std::fabs(dot(em.n_light, em.ray.d));
  throughput = throughput * cos_emit;
  extend_subpath(sc, em.ray, throughput, em.pdf_dir, cfg, smp, v, false, c);
}

static inline double powb(double x, float beta) { return beta == 2.0f ? x * x : std::pow(x, (double)beta); }

// What the walks need of the strategy being weighted besides the vertices'
// own pdfs (all area measure).
struct Conn {
  float pdf_cc = 0.0f;  // z[s-1] generated from the light side (:987-1006)
  float pdf_cl = 0.0f;  // y[t-1] generated from the camera side (:1011-1021)
  float rev_z2 = 0.0f;  // textbook walk: z[s-2] generated from z[s-1] arriving from the connection
  float rev_y2 = 0.0f;  // textbook walk: y[t-2] generated from y[t-1] arriving from the connection
  bool area_light = false;  // the path's light end is on an area light: a t' = 0 alternative exists
};

static inline bool mis_delta(const PVertex& v) { return v.type == V_SURFACE && v.is_delta; }  // :654-660

// :594-770
static float mis_weight(const std::vector<PVertex>& cv, const std::vector<PVertex>& lv, int s, int t,
                        const Conn& cn, const BdptCfg& cfg) {
  if (cfg.unweighted) return 1.0f;
  if (cfg.mis_mode == 1) {  // :605-636
    size_t len = (size_t)(s + t), nc = cv.size(), nl = lv.size();
    uint32_t count = 0;
    for (size_t t2 = 0; t2 <= std::min(len, nl); ++t2) {
      size_t s2 = len - t2;
      if (s2 > nc || s2 + t2 < 2) continue;
      if (s2 == 0) continue;
      count++;
    }
    return count == 0 ? 1.0f : 1.0f / (float)count;
  }
  if (s + t == 2) return 1.0f;
  double sum = 0.0;
  auto fwd = [](const PVertex& v) { return (double)std::fmax(v.pdf_fwd, 1e-30f); };
  if (cfg.mis_mode == 2) {
    // Textbook walk: strategy (s', t') of the same path is reached by moving
    // the connection; each vertex's reverse pdf is taken once, with the
    // connection's own pdfs at the two vertices next to it.  Alternatives
    // that the subpath lengths cannot produce (max_depth + 1 vertices per
    // subpath) or that are not implemented (s' = 0; t' = 0 without an area
    // light) are left out, so the weights of a path sum to one.
    const int k = s + t, dmax = (int)cfg.max_depth + 1;
    double ri = 1.0;
    for (int i = s - 1; i >= 1; --i) {  // alternative (i, k - i)
      const double rev = i == s - 1 ? cn.pdf_cc : (i == s - 2 ? cn.rev_z2 : std::fmax(cv[i].pdf_rev, 0.0f));
      ri *= rev / fwd(cv[i]);
      if (!mis_delta(cv[i]) && !mis_delta(cv[i - 1]) && k - i <= dmax) sum += powb(std::fabs(ri), cfg.beta);
    }
    ri = 1.0;
    for (int i = t - 1; i >= 0; --i) {  // alternative (k - i, i)
      const double rev = i == t - 1 ? cn.pdf_cl : (i == t - 2 ? cn.rev_y2 : std::fmax(lv[i].pdf_rev, 0.0f));
      ri *= rev / fwd(lv[i]);
      const bool prev_ok = i >= 1 ? !mis_delta(lv[i - 1]) : cn.area_light;
      if (!mis_delta(lv[i]) && prev_ok && k - i <= dmax) sum += powb(std::fabs(ri), cfg.beta);
    }
    return (float)(1.0 / (1.0 + sum));
  }
  {  // camera side (:678-717), as written: vertex s-1 enters the first step and the loop
    double ri = 1.0;
    if (s >= 1) {
      const PVertex& v = cv[s - 1];
      ri *= (double)cn.pdf_cc / fwd(v);
      bool prev_delta = s >= 2 ? mis_delta(cv[s - 2]) : true;
      if (!mis_delta(v) && !prev_delta) sum += powb(std::fabs(ri), cfg.beta);
    }
    for (int i = s - 1; i >= 1; --i) {
      const PVertex& v = cv[i];
      ri *= (double)std::fmax(v.pdf_rev, 0.0f) / fwd(v);
      bool prev_delta = mis_delta(cv[i - 1]);
      if (!mis_delta(v) && !prev_delta) sum += powb(std::fabs(ri), cfg.beta);
    }
  }
  {  // light side (:720-764); with an area light the t' = 0 alternative counts (extension)
    double ri = 1.0;
    if (t >= 1) {
      const PVertex& v = lv[t - 1];
      ri *= (double)cn.pdf_cl / fwd(v);
      bool prev_delta = t >= 2 ? mis_delta(lv[t - 2]) : !cn.area_light;
      if (!mis_delta(v) && !prev_delta) sum += powb(std::fabs(ri), cfg.beta);
    }
    const int lo = cn.area_light ? 0 : 1;
    for (int i = t - 1; i >= lo; --i) {
      const PVertex& v = lv[i];
      ri *= (double)std::fmax(v.pdf_rev, 0.0f) / fwd(v);
      bool prev_delta = i >= 1 ? mis_delta(lv[i - 1]) : !cn.area_light;
      if (!mis_delta(v) && !prev_delta) sum += powb(std::fabs(ri), cfg.beta);
    }
  }
  return (float)(1.0 / (1.0 + sum));
}

static bool path_area_light(const Scene& sc, const std::vector<PVertex>& lv, int t) {
  return t >= 1 && lv[0].light >= 0 && sc.lights[lv[0].light].type == LIGHT_QUAD;
}

// t = 0 (extension, the TODO at :810-818): the camera subpath's vertex s-1
// lies on the front face of an emitter.
static bool connect_t0(const Scene& sc, const std::vector<PVertex>& cv, const std::vector<PVertex>& lv, int s,
                       const BdptCfg& cfg, V3* out) {
  if (s < 2 || s > (int)cv.size()) return false;
  const PVertex& z = cv[s - 1];
  if (z.emit < 0) return false;
  V3 contrib = z.beta * z.mat.emission;
  if (max_element(contrib) <= 0.0f) return false;
  Conn cn;
  cn.pdf_cc = sc.has_dist ? sc.dist.pmf[z.emit] / sc.lights[z.emit].area : 0.0f;
  cn.rev_z2 = s >= 3 ? cv[s - 2].pdf_rev : 0.0f;
  cn.area_light = true;
  *out = contrib * mis_weight(cv, lv, s, 0, cn, cfg);
  return true;
}

// :782-1044 for s >= 2, t >= 1.
static bool connect(const Scene& sc, const std::vector<PVertex>& cv, const std::vector<PVertex>& lv, int s, int t,
                    const BdptCfg& cfg, V3* out, TraceCounters* c) {
  if (s < 2 || s > (int)cv.size() || t < 1 || t > (int)lv.size()) return false;
  const PVertex& z = cv[s - 1];
  const PVertex& y = lv[t - 1];
  if (!z.connectible() || !y.connectible()) return false;
  V3 d = y.p - z.p;
  float d2 = length_sq(d);
  if (d2 < 1e-10f) return false;
  float dist = std::sqrt(d2);
  V3 wi_z = d / dist;
  V3 wi_y = -wi_z;
  V3 oz = z.n * 1e-4f * rs_signum(dot(wi_z, z.n));
  V3 oy = y.n * 1e-4f * rs_signum(dot(wi_y, y.n));
  V3 so = z.p + oz, st = y.p + oy;
  V3 sd = st - so;
  float sdist = length(sd);
  if (sdist < 1e-6f) return false;
  Ray shadow = Ray::make(so, sd / sdist, 0.0f, sdist * (1.0f - 1e-4f));
  if (sc.occluded(shadow, c)) return false;
  V3 wo_z = normalize(cv[s - 2].p - z.p);
  V3 wo_z_l = z.onb.to_local(wo_z);
  V3 f_z = bsdf_eval(z.mat, z.onb.to_local(wi_z), wo_z_l);
  if (is_zero(f_z)) return false;
  const Light* Lt = (t == 1) ? &sc.lights[y.light] : nullptr;
  V3 f_y(1, 1, 1);
  if (t == 1) {
    if (Lt->type == LIGHT_QUAD && !(dot(y.n, wi_y) > 0.0f)) f_y = V3();  // one-sided emitter (extension)
  } else {
    V3 wo_y = normalize(lv[t - 2].p - y.p);
    f_y = bsdf_eval(y.mat, y.onb.to_local(wi_y), y.onb.to_local(wo_y));
  }
  if (is_zero(f_y)) return false;
  float g = geometry_term(z.p, z.n, y.p, y.n);
  V3 contrib;
  if (t == 1 && Lt->type == LIGHT_POINT) {  // :947-967
    LightSample ls;
    if (!Lt->sample(z.p, 0.0f, 0.0f, &ls)) return false;
    V3 f_direct = bsdf_eval(z.mat, z.onb.to_local(ls.wi), wo_z_l);
    float cos_z = std::fabs(dot(z.n, ls.wi));
    contrib = z.beta * f_direct * cos_z * ls.li;
    if (cfg.mis_mode == 2) contrib = contrib / sc.dist.pmf[y.light];  // textbook: the light pick's pmf
  } else {
    contrib = z.beta * f_z * g * f_y * y.beta;
  }
  if (max_element(contrib) <= 0.0f) return false;
  Conn cn;
  if (t == 1) {  // :987-994
    V3 dyz = normalize(z.p - y.p);
    cn.pdf_cc = pdf_sa_to_area(Lt->pdf_emission_dir(dyz), y.p, z.p, z.n);
  } else {  // :995-1006
    V3 wo_y = normalize(lv[t - 2].p - y.p);
    V3 wi_y2 = normalize(z.p - y.p);
    float pd = bsdf_pdf(y.mat, y.onb.to_local(wi_y2), y.onb.to_local(wo_y));
    cn.pdf_cc = pdf_sa_to_area(pd, y.p, z.p, z.n);
    if (cfg.mis_mode == 2)
      cn.rev_y2 = pdf_sa_to_area(bsdf_pdf(y.mat, y.onb.to_local(wo_y), y.onb.to_local(wi_y2)), y.p, lv[t - 2].p,
                                 lv[t - 2].n);
  }
  V3 wi_z2 = normalize(y.p - z.p);  // :1011-1021
  float pdz = bsdf_pdf(z.mat, z.onb.to_local(wi_z2), wo_z_l);
  cn.pdf_cl = pdf_sa_to_area(pdz, z.p, y.p, y.n);
  if (cfg.mis_mode == 2 && s >= 3)
    cn.rev_z2 = pdf_sa_to_area(bsdf_pdf(z.mat, wo_z_l, z.onb.to_local(wi_z2)), z.p, cv[s - 2].p, cv[s - 2].n);
  cn.area_light = path_area_light(sc, lv, t);
  *out = contrib * mis_weight(cv, lv, s, t, cn, cfg);
  return true;
}

// :1051-1155
static bool connect_s1(const Scene& sc, const Camera& cam, const std::vector<PVertex>& cv,
                       const std::vector<PVertex>& lv, int t, uint32_t w, uint32_t h, const BdptCfg& cfg, V3* out,
                       float* px, float* py, TraceCounters* c) {
  if (t < 2 || (int)lv.size() < t) return false;
  const PVertex& y = lv[t - 1];
  if (!y.connectible()) return false;
  Camera::WeSample cs;
  if (!cam.sample_we(y.p, w, h, &cs)) return false;
  V3 to_cam = cs.wi;
  float dist = cs.dist;
  V3 off = y.n * 1e-4f * rs_signum(dot(to_cam, y.n));
  Ray shadow = Ray::make(y.p + off, to_cam, 0.0f, dist * (1.0f - 1e-4f));
  if (sc.occluded(shadow, c)) return false;
  V3 wo_y = normalize(lv[t - 2].p - y.p);
  V3 wo_y_l = y.onb.to_local(wo_y);
  V3 wi_y_l = y.onb.to_local(to_cam);
  V3 f_y = bsdf_eval(y.mat, wi_y_l, wo_y_l);
  if (is_zero(f_y)) return false;
  float cos_y = std::fabs(dot(y.n, to_cam));
  float g = cos_y / (dist * dist);
  V3 contrib = y.beta * f_y * g * cs.we;
  // The reference's We = 1 / (A cos^4) is used without the cosine at the
  // camera (:1104-1108), so light tracing is bright by 1 / cos(theta); the
  // textbook walk takes the camera's cosine.
  if (cfg.mis_mode == 2) contrib = contrib * dot(-to_cam, cam.forward);
  if (max_element(contrib) <= 0.0f) return false;
  Conn cn;
  cn.pdf_cc = 0.0f;  // :1120
  if (cfg.mis_mode == 2) {
    // textbook: the camera's own pdf of the direction to y, in area measure at y
    Ray r = Ray::make(cam.origin, -to_cam, 0.0f, kF32Max);
    cn.pdf_cl = pdf_sa_to_area(cam.pdf_we(r), cam.origin, y.p, y.n);
    cn.rev_y2 = pdf_sa_to_area(bsdf_pdf(y.mat, wo_y_l, wi_y_l), y.p, lv[t - 2].p, lv[t - 2].n);
  } else {
    float pd = bsdf_pdf(y.mat, wi_y_l, wo_y_l);
    cn.pdf_cl = pdf_sa_to_area(pd, y.p, cam.origin, cv[0].n);  // :1124-1137, camera "normal" = first ray direction
  }
  cn.area_light = path_area_light(sc, lv, t);
  *out = contrib * mis_weight(cv, lv, 1, t, cn, cfg);
  *px = cs.px;
  *py = cs.py;
  return true;
}

}  // namespace orc

using namespace orc;

extern "C" {

// BidirectionalPathTracer::render (:1221-1357).  out = film (accum/spp plus
// splats * (1/spp)).  out_splat (optional) receives the splat film alone.
// only_s / only_t (test only): evaluate strategy (only_s, only_t) alone
// (-1: all); unweighted (test only): MIS weight 1.
double orc_render_bdpt_ex(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                          uint32_t rr_depth, int mis_mode, float mis_beta, int rng_mode, uint64_t seed, int nthreads,
                          int only_s, int only_t, int unweighted, uint32_t x0, uint32_t x1, uint32_t y0,
                          uint32_t y1, float* out, float* out_splat) {
  Scene* sc = (Scene*)s;
  Camera cam;
  const float* c = cam16;
  cam.origin = V3(c[0], c[1], c[2]);
  cam.llc = V3(c[3], c[4], c[5]);
  cam.horiz = V3(c[6], c[7], c[8]);
  cam.vert = V3(c[9], c[10], c[11]);
  cam.forward = V3(c[12], c[13], c[14]);
  cam.area = c[15];
  BdptCfg cfg{max_depth, rr_depth, mis_mode, mis_beta};
  cfg.only_s = only_s;
  cfg.only_t = only_t;
  cfg.unweighted = unweighted != 0;
  size_t npix = (size_t)w * h;
  // Light-tracing splats land on any pixel from any thread; the reference
  // adds them with float CAS loops in whatever order they arrive (film.rs:59-73,
  // nondeterministic).  They are summed here as 64-bit fixed point with 2^-24
  // resolution (each contribution rounded to nearest), the representation the
  // library's deterministic film uses, so the sum is order-independent.
  std::vector<uint64_t> splat(3 * npix, 0);
  auto fixed = [](float c) -> uint64_t {
    const float v = c * 16777216.0f;
    if (!(v == v)) return 0;
    return (uint64_t)(int64_t)std::fmin(std::fmax(std::rint(v), -4.0e18f), 4.0e18f);
  };
  std::mutex mu;
  auto t0 = std::chrono::steady_clock::now();
  if (x1 == 0 || x1 > w) x1 = w;
  if (y1 == 0 || y1 > h) y1 = h;
  parallel_rows(y1 - y0, nthreads, [&](uint32_t yy) {
    const uint32_t yr = y0 + yy;
    std::vector<std::pair<size_t, V3>> local;
    std::vector<PVertex> cvv, lvv;
    for (uint32_t x = x0; x < x1; ++x) {
      RefSampler ref(RefSampler::pixel_seed(x, yr));
      V3 acc;
      for (uint32_t si = 0; si < spp; ++si) {
        CounterSampler cs(seed, yr * w + x, si, 0);
        Sampler& smp = rng_mode == RNG_REF ? (Sampler&)ref : (Sampler&)cs;
        float jx = smp.next_1d(), jy = smp.next_1d();
        float u = ((float)x + jx) / (float)w;
        float v = 1.0f - ((float)yr + jy) / (float)h;
        Ray ray = cam.generate_ray(u, v);
        cvv.clear();
        lvv.clear();
        camera_subpath(*sc, cam, ray, cfg, smp, cvv, nullptr);
        light_subpath(*sc, cfg, smp, lvv, nullptr);
        const int smax = (int)cvv.size(), tmax = (int)lvv.size(), limit = (int)max_depth + 2;
        // The reference adds every strategy's contribution to one running
        // pixel sum (:1322); here a sample's contributions are summed first
        // and the sample then added to the pixel (the same value up to float
        // association; the library reduces per path the same way).
        V3 ls;
        for (int t = 0; t <= tmax; ++t)
          for (int sv = 0; sv <= smax; ++sv) {
            if (sv + t < 2 || sv + t > limit || sv == 0) continue;
            if (cfg.only_s >= 0 && (sv != cfg.only_s || t != cfg.only_t)) continue;
            V3 contrib;
            if (t == 0) {  // extension (:810-818)
              if (connect_t0(*sc, cvv, lvv, sv, cfg, &contrib)) ls = ls + contrib;
              continue;
            }
            if (sv == 1) {
              float px, py;
              if (connect_s1(*sc, cam, cvv, lvv, t, w, h, cfg, &contrib, &px, &py, nullptr)) {
                size_t ix = std::min<size_t>((size_t)px, w - 1), iy = std::min<size_t>((size_t)py, h - 1);
                local.push_back({iy * w + ix, contrib});
              }
              continue;
            }
            if (connect(*sc, cvv, lvv, sv, t, cfg, &contrib, nullptr)) ls = ls + contrib;
          }
        acc = acc + ls;
      }
      V3 m = acc / (float)spp;
      size_t o = 3 * ((size_t)yr * w + x);
      out[o] = m.x;
      out[o + 1] = m.y;
      out[o + 2] = m.z;
    }
    std::lock_guard<std::mutex> lk(mu);
    for (auto& e : local) {
      splat[3 * e.first] += fixed(e.second.x);
      splat[3 * e.first + 1] += fixed(e.second.y);
      splat[3 * e.first + 2] += fixed(e.second.z);
    }
  });
  float inv = 1.0f / (float)spp;  // :1341-1348
  for (size_t i = 0; i < npix; ++i) {
    auto val = [&](size_t k) { return (float)((double)(int64_t)splat[k] * (1.0 / 16777216.0)); };
    V3 sp(val(3 * i), val(3 * i + 1), val(3 * i + 2));
    if (out_splat) {
      out_splat[3 * i] = sp.x;
      out_splat[3 * i + 1] = sp.y;
      out_splat[3 * i + 2] = sp.z;
    }
    if (!is_zero(sp)) {
      V3 add = sp * inv;
      out[3 * i] += add.x;
      out[3 * i + 1] += add.y;
      out[3 * i + 2] += add.z;
    }
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_render_bdpt(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                       uint32_t rr_depth, int mis_mode, float mis_beta, int rng_mode, uint64_t seed, int nthreads,
                       float* out, float* out_splat) {
  return orc_render_bdpt_ex(s, cam16, w, h, spp, max_depth, rr_depth, mis_mode, mis_beta, rng_mode, seed, nthreads, -1,
                            -1, 0, 0, 0, 0, 0, out, out_splat);
}

}  // extern "C"
