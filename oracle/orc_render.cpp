// Oracle (test infrastructure only): scene, camera, BSDFs, lights, samplers
// and the path tracer, restated from the reference renderer.
#include "orc_render.h"

#include <algorithm>

namespace orc {

// ---------------------------------------------------------------------------
// surfaces.rs helpers
// ---------------------------------------------------------------------------
static inline bool same_hemisphere(V3 a, V3 b) { return a.z * b.z > 0.0f; }  // surfaces.rs:42
static inline float abs_cos(V3 v) { return std::fabs(v.z); }

// sin / cos of an fp32 angle in fp64, rounded once: the library's device
// code evaluates them the same way, so both sample the same directions
// (the reference's f32 sin/cos, Rust's libm, are statistically equivalent).
static inline float sin_f32(float x) { return (float)std::sin((double)x); }
static inline float cos_f32(float x) { return (float)std::cos((double)x); }

// surfaces.rs:50-57
static V3 cosine_hemisphere_sample(float ux, float uy) {
  float r = std::sqrt(ux);
  float phi = (2.0f * kPi) * uy;
  float x = r * cos_f32(phi);
  float y = r * sin_f32(phi);
  float z = std::sqrt(std::max((1.0f - x * x) - y * y, 0.0f));
  return V3(x, y, z);
}

// Rust powi(5) as x^2, x^4, x^4*x (LLVM's lowering of powi; unpinned).
static inline float powi5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x;
}

// surfaces.rs:79-83
static V3 fresnel_schlick(float cos_i, V3 f0) {
  float c = std::fabs(cos_i);
  float t = powi5(std::min(std::max(1.0f - c, 0.0f), 1.0f));
  return f0 + (V3(1, 1, 1) - f0) * t;
}

// surfaces.rs:86-101
static float fresnel_dielectric(float cos_theta_i, float eta) {
  float cos_i = std::min(std::max(cos_theta_i, -1.0f), 1.0f);
  if (cos_i < 0.0f) {
    cos_i = -cos_i;
    eta = 1.0f / eta;
  }
  float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return 1.0f;
  float cos_t = std::sqrt(1.0f - sin2_t);
  float r_par = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  float r_per = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  return 0.5f * (r_par * r_par + r_per * r_per);
}

// surfaces.rs:109-111
static inline V3 reflect(V3 v, V3 n) { return (-v) + (n * 2.0f) * dot(v, n); }

// surfaces.rs:115-123
static bool refract(V3 v, V3 n, float eta, V3* out) {
  float cos_i = dot(n, v);
  float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return false;
  float cos_t = std::sqrt(1.0f - sin2_t);
  *out = (-v) * eta + n * (eta * cos_i - cos_t);
  return true;
}

// microfacet.rs:14-65
static float ggx_d(float alpha, V3 h) {
  float a2 = alpha * alpha;
  float cos2_h = h.z * h.z;
  float denom = cos2_h * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * denom * denom);
}
static float ggx_lambda(float alpha, V3 v) {
  float cos2 = v.z * v.z;
  if (cos2 <= 0.0f) return INFINITY;
  float a2 = alpha * alpha;
  float tan2 = (1.0f - cos2) / cos2;
  return (std::sqrt(1.0f + a2 * tan2) - 1.0f) * 0.5f;
}
static float ggx_g2(float alpha, V3 wi, V3 wo) {
  return 1.0f / (1.0f + ggx_lambda(alpha, wi) + ggx_lambda(alpha, wo));
}
static V3 ggx_sample_wh(float alpha, V3 wo, float ux, float uy) {
  float a2 = alpha * alpha;
  float cos2_h = (1.0f - ux) / ((a2 - 1.0f) * ux + 1.0f);
  float cos_h = std::sqrt(std::max(cos2_h, 0.0f));
  float sin_h = std::sqrt(std::max(1.0f - cos2_h, 0.0f));
  float phi = (2.0f * kPi) * uy;
  V3 h(sin_h * cos_f32(phi), sin_h * sin_f32(phi), cos_h);
  if (h.z * wo.z < 0.0f) h = -h;
  return h;
}

static float clamp_rough(float r) { return std::min(std::max(r, 1e-3f), 1.0f); }

// --- Lambertian (lambertian.rs:34-65)
static V3 lambert_eval(const Material& m, V3 wi, V3 wo) {
  if (!same_hemisphere(wi, wo)) return V3();
  return m.albedo * kInvPi;
}
static bool lambert_sample(const Material& m, V3 wo, float ux, float uy, BsdfSample* o) {
  V3 wi = cosine_hemisphere_sample(ux, uy);
  if (wo.z < 0.0f) wi.z = -wi.z;
  float pdf = abs_cos(wi) * kInvPi;
  if (pdf == 0.0f) return false;
  o->wi = wi;
  o->pdf = pdf;
  o->f = m.albedo * kInvPi;
  return true;
}
static float lambert_pdf(V3 wi, V3 wo) {
  if (!same_hemisphere(wi, wo)) return 0.0f;
  return abs_cos(wi) * kInvPi;
}

// --- Conductor (microfacet.rs:102-153)
static V3 conductor_eval(const Material& m, V3 wi, V3 wo) {
  float a = clamp_rough(m.roughness);
  if (!same_hemisphere(wi, wo)) return V3();
  float cos_wi = abs_cos(wi), cos_wo = abs_cos(wo);
  if (cos_wi == 0.0f || cos_wo == 0.0f) return V3();
  V3 h = normalize(wi + wo);
  V3 f = fresnel_schlick(std::fabs(dot(wi, h)), m.albedo);
  float d = ggx_d(a, h);
  float g = ggx_g2(a, wi, wo);
  return f * (d * g / (4.0f * cos_wi * cos_wo));
}
static float conductor_pdf(const Material& m, V3 wi, V3 wo) {
  float a = clamp_rough(m.roughness);
  if (!same_hemisphere(wi, wo)) return 0.0f;
  V3 h = normalize(wi + wo);
  float wh_dot_wi = std::fabs(dot(wi, h));
  if (wh_dot_wi == 0.0f) return 0.0f;
  return ggx_d(a, h) * abs_cos(h) / (4.0f * wh_dot_wi);
}
static bool conductor_sample(const Material& m, V3 wo, float ux, float uy, BsdfSample* o) {
  if (wo.z == 0.0f) return false;
  V3 h = ggx_sample_wh(clamp_rough(m.roughness), wo, ux, uy);
  V3 wi = reflect(wo, h);
  if (!same_hemisphere(wi, wo)) return false;
  float pdf = conductor_pdf(m, wi, wo);
  if (pdf == 0.0f) return false;
  o->wi = wi;
  o->pdf = pdf;
  o->f = conductor_eval(m, wi, wo);
  return true;
}

// --- Dielectric (microfacet.rs:184-307)
static V3 dielectric_eval(const Material& m, V3 wi, V3 wo) {
  float a = clamp_rough(m.roughness);
  bool reflect_side = same_hemisphere(wi, wo);
  float cos_wo = abs_cos(wo), cos_wi = abs_cos(wi);
  if (cos_wo == 0.0f || cos_wi == 0.0f) return V3();
  if (reflect_side) {
    V3 h = normalize(wi + wo);
    float f = fresnel_dielectric(std::fabs(dot(wi, h)), m.eta);
    float d = ggx_d(a, h);
    float g = ggx_g2(a, wi, wo);
    return V3::splat(f * d * g / (4.0f * cos_wi * cos_wo));
  }
  float eta = wo.z > 0.0f ? m.eta : 1.0f / m.eta;
  V3 h = -normalize(wo + wi * eta);
  if (h.z * wo.z < 0.0f) h = -h;
  float wo_dot_h = dot(wo, h), wi_dot_h = dot(wi, h);
  if (wo_dot_h == 0.0f || wi_dot_h == 0.0f) return V3();
  float f = 1.0f - fresnel_dielectric(std::fabs(wi_dot_h), m.eta);
  float d = ggx_d(a, h);
  float g = ggx_g2(a, wi, wo);
  float x = wo_dot_h + eta * wi_dot_h;
  float denom = x * x;
  if (denom == 0.0f) return V3();
  float eta2 = eta * eta;
  float value = f * d * g * eta2 * std::fabs(wo_dot_h) * std::fabs(wi_dot_h) / (cos_wi * cos_wo * denom);
  return V3::splat(value);
}
static float dielectric_pdf(const Material& m, V3 wi, V3 wo) {
  float a = clamp_rough(m.roughness);
  float cos_wo = abs_cos(wo), cos_wi = abs_cos(wi);
  if (cos_wo == 0.0f || cos_wi == 0.0f) return 0.0f;
  if (same_hemisphere(wi, wo)) {
    V3 h = normalize(wi + wo);
    float wh_dot_wi = std::fabs(dot(wi, h));
    if (wh_dot_wi == 0.0f) return 0.0f;
    float f = fresnel_dielectric(std::fabs(dot(wo, h)), m.eta);
    return f * ggx_d(a, h) * abs_cos(h) / (4.0f * wh_dot_wi);
  }
  float eta = wo.z > 0.0f ? m.eta : 1.0f / m.eta;
  V3 h = -normalize(wo + wi * eta);
  if (h.z * wo.z < 0.0f) h = -h;
  float wo_dot_h = dot(wo, h), wi_dot_h = dot(wi, h);
  if (wo_dot_h == 0.0f || wi_dot_h == 0.0f) return 0.0f;
  float x = wo_dot_h + eta * wi_dot_h;
  float denom = x * x;
  if (denom == 0.0f) return 0.0f;
  float f = 1.0f - fresnel_dielectric(std::fabs(wo_dot_h), m.eta);
  return f * ggx_d(a, h) * abs_cos(h) * eta * eta * std::fabs(wi_dot_h) / denom;
}
static bool dielectric_sample(const Material& m, V3 wo, float u_sel, float ux, float uy, BsdfSample* o) {
  if (wo.z == 0.0f) return false;
  V3 h = ggx_sample_wh(clamp_rough(m.roughness), wo, ux, uy);
  float f = fresnel_dielectric(std::fabs(dot(wo, h)), m.eta);
  V3 wi;
  if (u_sel < f) {
    wi = reflect(wo, h);
    if (!same_hemisphere(wi, wo)) return false;
  } else {
    float eta = wo.z > 0.0f ? 1.0f / m.eta : m.eta;
    if (!refract(wo, h, eta, &wi)) return false;
    if (same_hemisphere(wi, wo)) return false;
  }
  float pdf = dielectric_pdf(m, wi, wo);
  if (pdf == 0.0f) return false;
  o->wi = wi;
  o->pdf = pdf;
  o->f = dielectric_eval(m, wi, wo);
  return true;
}

V3 bsdf_eval(const Material& m, V3 wi, V3 wo) {
  switch (m.type) {
    case MAT_CONDUCTOR: return conductor_eval(m, wi, wo);
    case MAT_DIELECTRIC: return dielectric_eval(m, wi, wo);
    default: return lambert_eval(m, wi, wo);
  }
}
bool bsdf_sample(const Material& m, V3 wo, float u_sel, float ux, float uy, BsdfSample* out) {
  switch (m.type) {
    case MAT_CONDUCTOR: return conductor_sample(m, wo, ux, uy, out);
    case MAT_DIELECTRIC: return dielectric_sample(m, wo, u_sel, ux, uy, out);
    default: return lambert_sample(m, wo, ux, uy, out);
  }
}
float bsdf_pdf(const Material& m, V3 wi, V3 wo) {
  switch (m.type) {
    case MAT_CONDUCTOR: return conductor_pdf(m, wi, wo);
    case MAT_DIELECTRIC: return dielectric_pdf(m, wi, wo);
    default: return lambert_pdf(wi, wo);
  }
}

// ---------------------------------------------------------------------------
// Lights
// ---------------------------------------------------------------------------
// point.rs:41-46
static V3 uniform_sphere_sample(float ux, float uy) {
  float cos_t = 1.0f - 2.0f * ux;
  float sin_t = std::sqrt(std::max(1.0f - cos_t * cos_t, 0.0f));
  float phi = (2.0f * kPi) * uy;
  return V3(sin_t * cos_f32(phi), sin_t * sin_f32(phi), cos_t);
}

bool Light::sample(V3 p, float ux, float uy, LightSample* out) const {
  if (type == LIGHT_POINT) {  // point.rs:49-60
    V3 off = pos - p;
    float dist2 = length_sq(off);
    float dist = std::sqrt(dist2);
    if (dist < 1e-6f) return false;
    out->wi = off / dist;
    out->dist = dist;
    out->li = intensity / dist2;
    out->pdf = 1.0f;
    return true;
  }
  // Quad area light (extension): uniform point on the quad, one-sided
  // emission along n, solid-angle pdf dist^2 / (area * cos_l).
  V3 x = (pos + eu * ux) + ev * uy;
  V3 off = x - p;
  float dist2 = length_sq(off);
  float dist = std::sqrt(dist2);
  if (dist < 1e-6f) return false;
  V3 wi = off / dist;
  float cos_l = -dot(n, wi);
  if (cos_l <= 0.0f) return false;
  out->wi = wi;
  out->dist = dist;
  out->li = intensity;
  out->pdf = dist2 / (area * cos_l);
  out->cos_l = cos_l;
  return true;
}

float Light::pdf_li(V3 p, V3 x, float dist) const {
  if (type == LIGHT_POINT) return 0.0f;
  V3 wi = (x - p) / dist;
  float cos_l = -dot(n, wi);
  if (cos_l <= 0.0f) return 0.0f;
  return (dist * dist) / (area * cos_l);
}

V3 Light::power() const {
  if (type == LIGHT_POINT) return intensity * (4.0f * kPi);  // point.rs:63-65
  return intensity * (kPi * area);                           // Lambertian emitter: pi*A*L
}

bool Light::sample_emission(float upx, float upy, float udx, float udy, EmissionSample* out) const {
  if (type == LIGHT_POINT) {  // point.rs:75-90
    V3 dir = uniform_sphere_sample(udx, udy);
    out->ray = Ray::make(pos, dir, 1e-4f, kF32Max);
    out->le = intensity;
    out->pdf_pos = 1.0f;
    out->pdf_dir = 1.0f / (4.0f * kPi);
    out->n_light = dir;
    return true;
  }
  // Quad (extension): uniform position, cosine-weighted direction about n.
  V3 x = (pos + eu * upx) + ev * upy;
  V3 l = cosine_hemisphere_sample(udx, udy);
  Onb o = Onb::from_normal(n);
  V3 dir = o.to_world(l);
  float cos_t = l.z;
  if (cos_t <= 0.0f) return false;
  out->ray = Ray::make(x, dir, 1e-4f, kF32Max);
  out->le = intensity;
  out->pdf_pos = 1.0f / area;
  out->pdf_dir = cos_t * kInvPi;
  out->n_light = n;
  return true;
}

float Light::pdf_emission_dir(V3 dir) const {
  if (type == LIGHT_POINT) return 1.0f / (4.0f * kPi);  // point.rs:92-95
  float c = dot(n, dir);
  return c > 0.0f ? c * kInvPi : 0.0f;
}

void LightDist::build(const std::vector<Light>& lights, bool use_power) {
  size_t n = lights.size();
  power = use_power;
  pmf.assign(n, 0.0f);
  cdf.assign(n, 0.0f);
  if (n == 0) return;
  if (!use_power) {  // lights.rs:146-159 (uniform)
    for (size_t i = 0; i < n; ++i) pmf[i] = 1.0f / (float)n;
    return;
  }
  std::vector<float> w(n);  // lights.rs:181-215
  float total = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    V3 p = lights[i].power();
    w[i] = std::fmax(0.2126f * p.x + 0.7152f * p.y + 0.0722f * p.z, 0.0f);
  }
  for (size_t i = 0; i < n; ++i) total = total + w[i];
  if (total == 0.0f) {
    float u = 1.0f / (float)n;
    for (size_t i = 0; i < n; ++i) {
      pmf[i] = u;
      cdf[i] = (float)(i + 1) * u;
    }
    return;
  }
  float acc = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    pmf[i] = w[i] / total;
    acc += pmf[i];
    cdf[i] = acc;
  }
  cdf[n - 1] = 1.0f;
}

void LightDist::sample_index(float u, uint32_t* idx, float* p) const {
  size_t n = pmf.size();
  if (!power) {  // lights.rs:147-150
    size_t i = (size_t)(u * (float)n);
    if (i > n - 1) i = n - 1;
    *idx = (uint32_t)i;
    *p = 1.0f / (float)n;
    return;
  }
  size_t i = 0;  // lights.rs:219-226: partition_point(c < u)
  while (i < n && cdf[i] < u) ++i;
  if (i > n - 1) i = n - 1;
  *idx = (uint32_t)i;
  *p = pmf[i];
}

// ---------------------------------------------------------------------------
// Scene (scene.rs)
// ---------------------------------------------------------------------------
void Scene::commit(int dist_mode) {
  tlas.inst.clear();
  for (auto& o : objects) {
    o->blas.reset(new Blas());
    o->blas->mesh = o->mesh;
    o->blas->build();
    o->nmat = normal_matrix(o->xf);
    Instance in;
    in.blas = o->blas.get();
    in.xf = o->xf;
    tlas.inst.push_back(in);
  }
  tlas.build();
  has_dist = dist_mode != 0 && !lights.empty();
  if (has_dist) dist.build(lights, dist_mode == 2);
}

// scene.rs:142-162
float srgb_to_linear(float x) {
  if (x <= 0.04045f) return x / 12.92f;
  return std::pow((x + 0.055f) / 1.055f, 2.4f);
}

void Texture::load_rgb8(const uint8_t* rgb, uint32_t width, uint32_t height) {
  w = width;
  h = height;
  data.resize((size_t)w * h);
  for (size_t i = 0; i < data.size(); ++i)
    data[i] = V3(srgb_to_linear((float)rgb[3 * i] / 255.0f), srgb_to_linear((float)rgb[3 * i + 1] / 255.0f),
                 srgb_to_linear((float)rgb[3 * i + 2] / 255.0f));
}

static int rem_euclid(int a, int b) {
  int r = a % b;
  return r < 0 ? r + b : r;
}

V3 Texture::fetch(int x, int y) const {
  uint32_t xx = (uint32_t)rem_euclid(x, (int)w), yy = (uint32_t)rem_euclid(y, (int)h);
  return data[(size_t)yy * w + xx];
}

// f32::fract (x - trunc(x)), then rem_euclid(1.0): a negative fraction + 1.
static float wrap01(float x) {
  float f = x - std::trunc(x);
  float r = std::fmod(f, 1.0f);
  return r < 0.0f ? r + 1.0f : r;
}

V3 Texture::sample(float uvx, float uvy) const {
  float fw = (float)w, fh = (float)h;
  float u = wrap01(uvx) * fw - 0.5f;
  float v = wrap01(uvy) * fh - 0.5f;
  int x0 = (int)std::floor(u), y0 = (int)std::floor(v);
  float tx = u - std::floor(u), ty = v - std::floor(v);
  V3 c00 = fetch(x0, y0), c10 = fetch(x0 + 1, y0), c01 = fetch(x0, y0 + 1), c11 = fetch(x0 + 1, y0 + 1);
  return ((c00 * ((1.0f - tx) * (1.0f - ty)) + c10 * (tx * (1.0f - ty))) + c01 * ((1.0f - tx) * ty)) + c11 * (tx * ty);
}

bool Scene::intersect(const Ray& r, float* tout, ShadingPoint* sp, TraceCounters* c) const {
  RayHit h;
  if (!tlas.intersect(r, &h, c)) return false;
  const Object& obj = *objects[h.inst];
  const Mesh& m = *obj.mesh;
  V3 p = r.at(h.t);
  V3 n_obj;
  if (!m.nrm.empty()) {  // scene.rs:75-87
    uint32_t i0 = m.idx[3 * h.prim], i1 = m.idx[3 * h.prim + 1], i2 = m.idx[3 * h.prim + 2];
    float w = 1.0f - h.u - h.v;
    n_obj = normalize((m.nrm[i0] * w + m.nrm[i1] * h.u) + m.nrm[i2] * h.v);
  } else {
    n_obj = h.n;
  }
  V3 n = normalize(mul(obj.nmat, n_obj));
  V3 wo = -r.d;
  V3 nf = dot(n, wo) < 0.0f ? -n : n;
  sp->p = p;
  sp->n = nf;
  sp->ng_world = n;
  sp->onb = Onb::from_normal(nf);
  uint32_t slot = m.slots.size() == 1 ? m.slots[0] : m.slots[h.prim];  // scene.rs:41-48
  size_t nm = obj.mats.size();
  if (!m.uv.empty()) {  // scene.rs:55-64: interpolated tex_coords, else the barycentrics
    uint32_t i0 = m.idx[3 * h.prim], i1 = m.idx[3 * h.prim + 1], i2 = m.idx[3 * h.prim + 2];
    float w = 1.0f - h.u - h.v;
    sp->tu = (m.uv[2 * i0] * w + m.uv[2 * i1] * h.u) + m.uv[2 * i2] * h.v;
    sp->tv = (m.uv[2 * i0 + 1] * w + m.uv[2 * i1 + 1] * h.u) + m.uv[2 * i2 + 1] * h.v;
  } else {
    sp->tu = h.u;
    sp->tv = h.v;
  }
  sp->mtl = materials[obj.mats[std::min<size_t>(slot, nm - 1)]];
  if (sp->mtl.tex >= 0) sp->mtl.albedo = textures[sp->mtl.tex].sample(sp->tu, sp->tv);  // lambertian.rs:35-56
  sp->mat = &sp->mtl;
  sp->inst = h.inst;
  sp->prim = h.prim;
  *tout = h.t;
  return true;
}

// ---------------------------------------------------------------------------
// Camera (cameras.rs:91-248)
// ---------------------------------------------------------------------------
Camera Camera::lookat(V3 from, V3 at, V3 up, float vfov, float aspect) {
  const float rads_per_deg = kPi / 180.0f;
  float theta = vfov * rads_per_deg;
  float h = std::tan(theta / 2.0f);
  float vh = 2.0f * h;
  float vw = aspect * vh;
  V3 w = normalize(from - at);
  V3 u = normalize(cross(up, w));
  V3 v = cross(w, u);
  Camera c;
  c.origin = from;
  c.horiz = u * vw;
  c.vert = v * vh;
  c.llc = ((from - c.horiz / 2.0f) - c.vert / 2.0f) - w;
  c.forward = -w;
  c.area = vw * vh;
  return c;
}

bool Camera::sample_we(V3 p, uint32_t width, uint32_t height, WeSample* out) const {
  V3 to_cam = origin - p;
  float dist2 = length_sq(to_cam);
  float dist = std::sqrt(dist2);
  if (dist < 1e-6f) return false;
  V3 wi = to_cam / dist;
  V3 dir = -wi;
  float cos_theta = dot(dir, forward);
  if (cos_theta <= 1e-6f) return false;
  V3 right = normalize(horiz), up = normalize(vert);
  float vw = length(horiz), vh = length(vert);
  V3 ds = dir / cos_theta;
  float u = dot(ds, right) / vw + 0.5f;
  float v = dot(ds, up) / vh + 0.5f;
  if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return false;
  float px = u * (float)width;
  float py = (1.0f - v) * (float)height;
  if (px < 0.0f || px >= (float)width || py < 0.0f || py >= (float)height) return false;
  float cos2 = cos_theta * cos_theta;
  float cos4 = cos2 * cos2;
  float cos3 = cos2 * cos_theta;
  out->px = px;
  out->py = py;
  out->wi = wi;
  out->dist = dist;
  out->we = 1.0f / (area * cos4);
  out->pdf = dist2 / (area * cos3);
  return true;
}

float Camera::pdf_we(const Ray& r) const {
  float c = dot(r.d, forward);
  if (c <= 0.0f) return 0.0f;
  return 1.0f / (area * c * c * c);
}

// ---------------------------------------------------------------------------
// Samplers
// ---------------------------------------------------------------------------
RefSampler::RefSampler(uint64_t state) {
  const uint64_t PHI = 0x9e3779b97f4a7c15ull;
  for (int i = 0; i < 4; ++i) {
    state += PHI;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    s[i] = z ^ (z >> 31);
  }
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t RefSampler::next_u64() {  // Xoshiro256++
  uint64_t result = rotl(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

float CounterSampler::next_1d() {
  uint32_t ctr[4] = {dim >> 2, sample, pixel, stream};
  uint32_t o[4];
  philox4x32_10(ctr, key, o);
  uint32_t w = o[dim & 3];
  dim++;
  return (float)(w >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Path tracer (path_tracer.rs:59-156) + emission on hit (extension: the
// reference's non-delta MIS branch is a skeleton that never adds Le).
// ---------------------------------------------------------------------------
float power_heuristic(float a, float b) {  // path_tracer.rs:201-205
  float a2 = a * a, b2 = b * b;
  return a2 / (a2 + b2);
}

V3 pt_li(const Scene& sc, Ray ray, Sampler& smp, const PtConfig& cfg, TraceCounters* c) {
  V3 radiance, beta(1, 1, 1);
  float prev_pdf = 0.0f;
  V3 prev_p;
  for (uint32_t depth = 0; depth < cfg.max_depth; ++depth) {
    float t;
    ShadingPoint hit;
    if (!sc.intersect(ray, &t, &hit, c)) break;
    V3 wo = -ray.d;
    // Emission (extension): camera hit adds Le; BSDF-sampled hit adds Le with
    // the power-heuristic weight against light sampling.  Emitters are black.
    if (hit.mat->emissive()) {
      // path of depth + 2 vertices: camera, depth surfaces, the emitter
      const bool keep = cfg.only_len == 0 || cfg.only_len == depth + 2;
      if (dot(hit.ng_world, wo) > 0.0f && keep) {
        if (depth == 0 || !sc.has_dist || cfg.pt_mode == 2) {
          radiance = radiance + beta * hit.mat->emission;
        } else if (cfg.pt_mode == 0) {
          const Light& L = sc.lights[hit.mat->light_id];
          float pdf_l = sc.dist.pmf[hit.mat->light_id] * L.pdf_li(prev_p, hit.p, length(hit.p - prev_p));
          float w = power_heuristic(prev_pdf, pdf_l);
          radiance = radiance + beta * hit.mat->emission * w;
        }
      }
      break;
    }
    V3 wo_l = hit.onb.to_local(wo);
    if (sc.has_dist && cfg.pt_mode != 2) {  // path_tracer.rs:79-117
      float u_sel = smp.next_1d();
      uint32_t li;
      float sel_pmf;
      sc.dist.sample_index(u_sel, &li, &sel_pmf);
      const Light& L = sc.lights[li];
      float ux = smp.next_1d(), uy = smp.next_1d();
      LightSample ls;
      if (L.sample(hit.p, ux, uy, &ls)) {
        V3 off = hit.n * 1e-4f * rs_signum(dot(ls.wi, hit.n));
        float t_end = ls.dist;  // point light: path_tracer.rs:88-92
        // Quad (extension): distance from the offset origin to the emitter
        // plane, so the ray stops a relative 1e-4 short of the emitter's own
        // triangles.  With t_end = dist the origin offset and the shortening
        // cancel wherever the light is ~1 unit away, and occlusion by the
        // emitter is then decided by the last bit.
        if (!L.is_delta()) t_end = ls.dist + dot(off, L.n) / ls.cos_l;
        Ray sr = Ray::make(hit.p + off, ls.wi, 0.0f, t_end * (1.0f - 1e-4f));
        // path of depth + 3 vertices: camera, depth + 1 surfaces, the light
        const bool keep = cfg.only_len == 0 || cfg.only_len == depth + 3;
        if (keep && !sc.occluded(sr, c)) {
          V3 wi_l = hit.onb.to_local(ls.wi);
          float cos_wi = std::fabs(wi_l.z);
          if (cos_wi > 0.0f) {
            V3 f = bsdf_eval(*hit.mat, wi_l, wo_l);
            if (L.is_delta()) {
              radiance = radiance + beta * f * cos_wi * ls.li / sel_pmf;
            } else {
              float pdf_light = ls.pdf * sel_pmf;
              float pdf_bsdf = bsdf_pdf(*hit.mat, wi_l, wo_l);
              float w = cfg.pt_mode == 1 ? 1.0f : power_heuristic(pdf_light, pdf_bsdf);
              radiance = radiance + beta * f * cos_wi * ls.li * w / pdf_light;
            }
          }
        }
      }
    }
    float u_sel = smp.next_1d();
    float ux = smp.next_1d(), uy = smp.next_1d();
    BsdfSample bs;
    if (!bsdf_sample(*hit.mat, wo_l, u_sel, ux, uy, &bs)) break;
    if (bs.pdf == 0.0f || is_zero(bs.f)) break;
    float cos_wi = std::fabs(bs.wi.z);
    beta = beta * (bs.f * cos_wi / bs.pdf);
    prev_pdf = bs.pdf;
    prev_p = hit.p;
    if (depth + 1 >= cfg.rr_depth) {  // path_tracer.rs:139-145
      float survival = std::fmin(max_element(beta), 0.95f);
      if (smp.next_1d() > survival) break;
      beta = beta / survival;
    }
    V3 wi_w = hit.onb.to_world(bs.wi);
    V3 off = hit.n * 1e-4f * rs_signum(dot(wi_w, hit.n));
    ray = Ray::make(hit.p + off, wi_w, 1e-4f, kF32Max);
  }
  return radiance;
}

void parallel_rows(uint32_t h, int nthreads, const std::function<void(uint32_t)>& fn) {
  if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
  if (nthreads <= 1) {
    for (uint32_t y = 0; y < h; ++y) fn(y);
    return;
  }
  std::atomic<uint32_t> next(0);
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i)
    th.emplace_back([&]() {
      for (;;) {
        uint32_t y = next.fetch_add(1);
        if (y >= h) break;
        fn(y);
      }
    });
  for (auto& t : th) t.join();
}

}  // namespace orc
