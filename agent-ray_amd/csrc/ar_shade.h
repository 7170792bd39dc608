// Device shading: hit -> shading point (agent-ray src/scene.rs:142-162), the
// closed BSDF set (src/surfaces.rs, surfaces/lambertian.rs,
// surfaces/microfacet.rs), lights (src/lights.rs, lights/point.rs, plus the
// quad area light extension) and the pinhole camera (src/cameras.rs).
// Arithmetic follows the reference's expression order so that every
// integrator built on these functions computes the same per-path values.
#pragma once
#include "ar_internal.h"

namespace ar {

struct ShadeView {
  SceneView sv;
  const DObject* objects;
  const DMaterial* materials;
  const DLight* lights;
  const uint32_t* idx;
  const float* nrm;
  const uint32_t* slots;
  const uint32_t* objmats;
  const float4* stri;  // per Tri record: vertex normals, material, object (3 float4)
  const float4* suv;   // per Tri record: vertex tex_coords (2 float4), or NULL when no mesh has them
  const float4* texels;     // linear texels of every texture
  const uint4* tex_info;    // per texture: (first texel, width, height, -)
  uint32_t n_tex;           // textures in the table
  uint32_t n_lights;
  uint32_t dist_mode;  // 0 none, 1 uniform, 2 power
};

struct Onb {
  f3 n, t, b;
};
AR_HD Onb onb_from_normal(f3 nn) {  // geometry.rs:218-232
  Onb o;
  o.n = normalize(nn);
  f3 a = fabsf(o.n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  o.b = normalize(cross(o.n, a));
  o.t = cross(o.b, o.n);
  return o;
}
AR_HD f3 to_local(const Onb& o, f3 v) { return mk(dot(v, o.t), dot(v, o.b), dot(v, o.n)); }
AR_HD f3 to_world(const Onb& o, f3 v) { return (o.t * v.x + o.b * v.y) + o.n * v.z; }

struct ShadingPoint {
  f3 p, n, ng;  // n: shading normal flipped toward wo; ng: unflipped world normal
  Onb onb;
  uint32_t mat;
  float tu, tv;  // tex_uv (scene.rs:55-64), set when the material is textured
};

// f32::fract(x).rem_euclid(1.0): the fractional part, moved into [0, 1).
__device__ __forceinline__ float wrap01(float x) {
  const float r = fmodf(x - truncf(x), 1.0f);
  return r < 0.0f ? r + 1.0f : r;
}

// ImageTexture::sample (texture.rs:97-135): bilinear over the four nearest
// texels, repeat wrap, in the reference's operation order.
__device__ __forceinline__ f3 texture_sample(const ShadeView& v, int32_t tex, float uvx, float uvy) {
  const uint4 ti = v.tex_info[tex];
  const int w = (int)ti.y, h = (int)ti.z;
  const float u = wrap01(uvx) * (float)w - 0.5f;
  const float vv = wrap01(uvy) * (float)h - 0.5f;
  const float fu = floorf(u), fv = floorf(vv);
  const int x0 = (int)fu, y0 = (int)fv;
  const float tx = u - fu, ty = vv - fv;
  auto fetch = [&](int x, int y) {
    int xx = x % w, yy = y % h;  // rem_euclid
    xx = xx < 0 ? xx + w : xx;
    yy = yy < 0 ? yy + h : yy;
    const float4 c = v.texels[ti.x + (uint32_t)yy * (uint32_t)w + (uint32_t)xx];
    return mk(c.x, c.y, c.z);
  };
  const f3 c00 = fetch(x0, y0), c10 = fetch(x0 + 1, y0), c01 = fetch(x0, y0 + 1), c11 = fetch(x0 + 1, y0 + 1);
  return ((c00 * ((1.0f - tx) * (1.0f - ty)) + c10 * (tx * (1.0f - ty))) + c01 * ((1.0f - tx) * ty)) + c11 * (tx * ty);
}

// The material of a surface point: a textured Lambertian's albedo sampled
// at tex_uv (lambertian.rs:35-56), every other material as stored.
__device__ __forceinline__ DMaterial material_at(const ShadeView& v, uint32_t mat, float tu, float tv) {
  DMaterial m = v.materials[mat];
  if (m.tex >= 0 && (uint32_t)m.tex < v.n_tex) {
    const f3 a = texture_sample(v, m.tex, tu, tv);
    m.albedo[0] = a.x;
    m.albedo[1] = a.y;
    m.albedo[2] = a.z;
  }
  return m;
}

// Scene::intersect's shading part (scene.rs:142-162) for a closest hit.
__device__ __forceinline__ void make_shading(const ShadeView& v, const Ray& r, float t, uint32_t pos, uint32_t inst,
                                             float bu, float bv, ShadingPoint& sp) {
  // The Tri record's shading record holds its vertex normals, material and
  // object, resolved at commit from the object's tables, so those loads do
  // not wait on each other (the instance record is not needed: a Tri record
  // belongs to one object).
  const Tri* tr = v.sv.tris + pos;
  const float4* srec = v.stri + 3 * (size_t)pos;
  const float4 sa = srec[0], sb = srec[1], sc = srec[2];
  const DObject& ob = v.objects[__float_as_uint(sc.w)];
  sp.p = r.o + r.d * t;
  f3 n_obj;
  if (__float_as_uint(sb.w) != 0u) {  // scene.rs:75-84
    f3 n0 = mk(sa.x, sa.y, sa.z);
    f3 n1 = mk(sb.x, sb.y, sb.z);
    f3 n2 = mk(sc.x, sc.y, sc.z);
    float w = 1.0f - bu - bv;
    n_obj = normalize((n0 * w + n1 * bu) + n2 * bv);
  } else {
    float4 e1 = tr->e1, e2 = tr->e2;
    n_obj = normalize(cross(mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z)));  // geometry.rs:203
  }
  const float* M = ob.nmat;  // Mat3 * Vec3: (c0*x + c1*y) + c2*z
  f3 nw = mk((M[0] * n_obj.x + M[3] * n_obj.y) + M[6] * n_obj.z, (M[1] * n_obj.x + M[4] * n_obj.y) + M[7] * n_obj.z,
             (M[2] * n_obj.x + M[5] * n_obj.y) + M[8] * n_obj.z);
  f3 n = normalize(nw);
  f3 wo = -r.d;
  sp.ng = n;
  sp.n = dot(n, wo) < 0.0f ? -n : n;
  sp.onb = onb_from_normal(sp.n);
  sp.mat = __float_as_uint(sa.w);  // scene.rs:41-48
  sp.tu = bu;
  sp.tv = bv;
  if (v.suv && v.materials[sp.mat].tex >= 0 && (uint32_t)v.materials[sp.mat].tex < v.n_tex) {  // scene.rs:55-64
    const float4 ua = v.suv[2 * (size_t)pos], ub = v.suv[2 * (size_t)pos + 1];
    if (ub.z != 0.0f) {
      const float w = 1.0f - bu - bv;
      sp.tu = (ua.x * w + ua.z * bu) + ub.x * bv;
      sp.tv = (ua.y * w + ua.w * bu) + ub.y * bv;
    }
  }
}

// ---------------------------------------------------------------------------
// BSDFs (local frame, +Z = shading normal)
// ---------------------------------------------------------------------------
AR_HD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }

// sin and cos of an fp32 angle evaluated in fp64 and rounded once to fp32.
// Two accurate fp64 implementations (the device's and a CPU's) then give
// the same fp32 bits except where their results straddle an fp32 rounding
// boundary (about one angle in 2^29), so a sampled direction does not
// depend on whose fp32 sin/cos ran.  The reference's f32 sin/cos (Rust's
// libm, surfaces.rs:50-57, microfacet.rs:53-65, point.rs:41-46) fixes no
// particular last bit either.
__device__ __forceinline__ void sincos_f32(float x, float* s, float* c) {
  double sd, cd;
  sincos((double)x, &sd, &cd);
  *s = (float)sd;
  *c = (float)cd;
}

__device__ __forceinline__ f3 cosine_hemisphere(float ux, float uy) {  // surfaces.rs:50-57
  float r = sqrtf(ux);
  float phi = (2.0f * kPi) * uy;
  float s, c;
  sincos_f32(phi, &s, &c);
  float x = r * c, y = r * s;
  float z = sqrtf(fmaxf((1.0f - x * x) - y * y, 0.0f));
  return mk(x, y, z);
}
AR_HD float powi5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x;
}
AR_HD f3 fresnel_schlick(float cos_i, f3 f0) {  // surfaces.rs:79-83
  float c = fabsf(cos_i);
  float t = powi5(fminf(fmaxf(1.0f - c, 0.0f), 1.0f));
  return f0 + (mk(1.0f, 1.0f, 1.0f) - f0) * t;
}
AR_HD float fresnel_dielectric(float cos_theta_i, float eta) {  // surfaces.rs:86-101
  float cos_i = fminf(fmaxf(cos_theta_i, -1.0f), 1.0f);
  if (cos_i < 0.0f) {
    cos_i = -cos_i;
    eta = 1.0f / eta;
  }
  float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return 1.0f;
  float cos_t = sqrtf(1.0f - sin2_t);
  float r_par = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  float r_per = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  return 0.5f * (r_par * r_par + r_per * r_per);
}
AR_HD f3 reflect(f3 v, f3 n) { return (-v) + (n * 2.0f) * dot(v, n); }
AR_HD bool refract(f3 v, f3 n, float eta, f3* out) {
  float cos_i = dot(n, v);
  float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return false;
  float cos_t = sqrtf(1.0f - sin2_t);
  *out = (-v) * eta + n * (eta * cos_i - cos_t);
  return true;
}
AR_HD float ggx_d(float a, f3 h) {  // microfacet.rs:14-19
  float a2 = a * a;
  float c2 = h.z * h.z;
  float den = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * den * den);
}
AR_HD float ggx_lambda(float a, f3 v) {
  float c2 = v.z * v.z;
  if (c2 <= 0.0f) return INFINITY;
  float a2 = a * a;
  float tan2 = (1.0f - c2) / c2;
  return (sqrtf(1.0f + a2 * tan2) - 1.0f) * 0.5f;
}
AR_HD float ggx_g2(float a, f3 wi, f3 wo) { return 1.0f / (1.0f + ggx_lambda(a, wi) + ggx_lambda(a, wo)); }
__device__ __forceinline__ f3 ggx_sample_wh(float a, f3 wo, float ux, float uy) {  // microfacet.rs:53-65
  float a2 = a * a;
  float c2 = (1.0f - ux) / ((a2 - 1.0f) * ux + 1.0f);
  float ch = sqrtf(fmaxf(c2, 0.0f));
  float sh = sqrtf(fmaxf(1.0f - c2, 0.0f));
  float phi = (2.0f * kPi) * uy;
  float s, c;
  sincos_f32(phi, &s, &c);
  f3 h = mk(sh * c, sh * s, ch);
  if (h.z * wo.z < 0.0f) h = -h;
  return h;
}

__device__ __forceinline__ f3 bsdf_eval(const DMaterial& m, f3 wi, f3 wo) {
  float a = m.roughness;
  if (m.type == AR_MAT_LAMBERTIAN) {
    if (!same_hemi(wi, wo)) return splat(0.0f);
    return mk(m.albedo[0], m.albedo[1], m.albedo[2]) * kInvPi;
  }
  float cwo = fabsf(wo.z), cwi = fabsf(wi.z);
  if (m.type == AR_MAT_CONDUCTOR) {
    if (!same_hemi(wi, wo)) return splat(0.0f);
    if (cwi == 0.0f || cwo == 0.0f) return splat(0.0f);
    f3 h = normalize(wi + wo);
    f3 f = fresnel_schlick(fabsf(dot(wi, h)), mk(m.albedo[0], m.albedo[1], m.albedo[2]));
    float d = ggx_d(a, h);
    float g = ggx_g2(a, wi, wo);
    return f * (d * g / (4.0f * cwi * cwo));
  }
  // dielectric (microfacet.rs:185-229)
  if (cwo == 0.0f || cwi == 0.0f) return splat(0.0f);
  if (same_hemi(wi, wo)) {
    f3 h = normalize(wi + wo);
    float f = fresnel_dielectric(fabsf(dot(wi, h)), m.eta);
    float d = ggx_d(a, h);
    float g = ggx_g2(a, wi, wo);
    return splat(f * d * g / (4.0f * cwi * cwo));
  }
  float eta = wo.z > 0.0f ? m.eta : 1.0f / m.eta;
  f3 h = -normalize(wo + wi * eta);
  if (h.z * wo.z < 0.0f) h = -h;
  float woh = dot(wo, h), wih = dot(wi, h);
  if (woh == 0.0f || wih == 0.0f) return splat(0.0f);
  float f = 1.0f - fresnel_dielectric(fabsf(wih), m.eta);
  float d = ggx_d(a, h);
  float g = ggx_g2(a, wi, wo);
  float x = woh + eta * wih;
  float den = x * x;
  if (den == 0.0f) return splat(0.0f);
  float eta2 = eta * eta;
  return splat(f * d * g * eta2 * fabsf(woh) * fabsf(wih) / (cwi * cwo * den));
}

__device__ __forceinline__ float bsdf_pdf(const DMaterial& m, f3 wi, f3 wo) {
  float a = m.roughness;
  if (m.type == AR_MAT_LAMBERTIAN) {
    if (!same_hemi(wi, wo)) return 0.0f;
    return fabsf(wi.z) * kInvPi;
  }
  if (m.type == AR_MAT_CONDUCTOR) {
    if (!same_hemi(wi, wo)) return 0.0f;
    f3 h = normalize(wi + wo);
    float whi = fabsf(dot(wi, h));
    if (whi == 0.0f) return 0.0f;
    return ggx_d(a, h) * fabsf(h.z) / (4.0f * whi);
  }
  float cwo = fabsf(wo.z), cwi = fabsf(wi.z);
  if (cwo == 0.0f || cwi == 0.0f) return 0.0f;
  if (same_hemi(wi, wo)) {
    f3 h = normalize(wi + wo);
    float whi = fabsf(dot(wi, h));
    if (whi == 0.0f) return 0.0f;
    float f = fresnel_dielectric(fabsf(dot(wo, h)), m.eta);
    return f * ggx_d(a, h) * fabsf(h.z) / (4.0f * whi);
  }
  float eta = wo.z > 0.0f ? m.eta : 1.0f / m.eta;
  f3 h = -normalize(wo + wi * eta);
  if (h.z * wo.z < 0.0f) h = -h;
  float woh = dot(wo, h), wih = dot(wi, h);
  if (woh == 0.0f || wih == 0.0f) return 0.0f;
  float x = woh + eta * wih;
  float den = x * x;
  if (den == 0.0f) return 0.0f;
  float f = 1.0f - fresnel_dielectric(fabsf(woh), m.eta);
  return f * ggx_d(a, h) * fabsf(h.z) * eta * eta * fabsf(wih) / den;
}

struct BsdfSample {
  f3 wi;
  float pdf;
  f3 f;
};

__device__ __forceinline__ bool bsdf_sample(const DMaterial& m, f3 wo, float u_sel, float ux, float uy,
                                            BsdfSample& o) {
  if (m.type == AR_MAT_LAMBERTIAN) {  // lambertian.rs:42-57
    f3 wi = cosine_hemisphere(ux, uy);
    if (wo.z < 0.0f) wi.z = -wi.z;
    float pdf = fabsf(wi.z) * kInvPi;
    if (pdf == 0.0f) return false;
    o.wi = wi;
    o.pdf = pdf;
    o.f = mk(m.albedo[0], m.albedo[1], m.albedo[2]) * kInvPi;
    return true;
  }
  if (wo.z == 0.0f) return false;
  f3 h = ggx_sample_wh(m.roughness, wo, ux, uy);
  f3 wi;
  if (m.type == AR_MAT_CONDUCTOR) {  // microfacet.rs:121-139
    wi = reflect(wo, h);
    if (!same_hemi(wi, wo)) return false;
  } else {  // microfacet.rs:231-270
    float f = fresnel_dielectric(fabsf(dot(wo, h)), m.eta);
    if (u_sel < f) {
      wi = reflect(wo, h);
      if (!same_hemi(wi, wo)) return false;
    } else {
      float eta = wo.z > 0.0f ? 1.0f / m.eta : m.eta;
      if (!refract(wo, h, eta, &wi)) return false;
      if (same_hemi(wi, wo)) return false;
    }
  }
  float pdf = bsdf_pdf(m, wi, wo);
  if (pdf == 0.0f) return false;
  o.wi = wi;
  o.pdf = pdf;
  o.f = bsdf_eval(m, wi, wo);
  return true;
}

// ---------------------------------------------------------------------------
// Lights
// ---------------------------------------------------------------------------
__device__ __forceinline__ void light_select(const ShadeView& v, float u, uint32_t* idx, float* pmf) {
  uint32_t n = v.n_lights;
  if (v.dist_mode == AR_LIGHTDIST_UNIFORM) {  // lights.rs:147-150
    float fi = u * (float)n;
    uint32_t i = (uint32_t)fi;
    if (i > n - 1) i = n - 1;
    *idx = i;
    *pmf = 1.0f / (float)n;
    return;
  }
  uint32_t i = 0;  // lights.rs:219-226
  while (i < n && v.lights[i].cdf < u) ++i;
  if (i > n - 1) i = n - 1;
  *idx = i;
  *pmf = v.lights[i].pmf;
}

struct LightSample {
  f3 wi;
  float dist;
  f3 li;
  float pdf;
  float cos_l;  // quad: cosine at the light (0 for a point light)
};

__device__ __forceinline__ bool light_sample(const DLight& L, f3 p, float ux, float uy, LightSample& o) {
  if (L.type == AR_LIGHT_POINT) {  // point.rs:49-60
    f3 off = mk(L.pos[0], L.pos[1], L.pos[2]) - p;
    float d2 = dot(off, off);
    float d = sqrtf(d2);
    if (d < 1e-6f) return false;
    o.wi = off / d;
    o.dist = d;
    o.li = mk(L.intensity[0], L.intensity[1], L.intensity[2]) / d2;
    o.pdf = 1.0f;
    o.cos_l = 0.0f;
    return true;
  }
  f3 x = (mk(L.pos[0], L.pos[1], L.pos[2]) + mk(L.eu[0], L.eu[1], L.eu[2]) * ux) + mk(L.ev[0], L.ev[1], L.ev[2]) * uy;
  f3 off = x - p;
  float d2 = dot(off, off);
  float d = sqrtf(d2);
  if (d < 1e-6f) return false;
  f3 wi = off / d;
  float cl = -dot(mk(L.n[0], L.n[1], L.n[2]), wi);
  if (cl <= 0.0f) return false;
  o.wi = wi;
  o.dist = d;
  o.li = mk(L.intensity[0], L.intensity[1], L.intensity[2]);
  o.pdf = d2 / (L.area * cl);
  o.cos_l = cl;
  return true;
}

// Solid-angle pdf of the quad light sampling point x from p at distance d.
__device__ __forceinline__ float light_pdf_li(const DLight& L, f3 p, f3 x, float d) {
  if (L.type == AR_LIGHT_POINT) return 0.0f;
  f3 wi = (x - p) / d;
  float cl = -dot(mk(L.n[0], L.n[1], L.n[2]), wi);
  if (cl <= 0.0f) return 0.0f;
  return (d * d) / (L.area * cl);
}

AR_HD float power_heuristic(float a, float b) {  // path_tracer.rs:201-205
  float a2 = a * a, b2 = b * b;
  return a2 / (a2 + b2);
}

// ---------------------------------------------------------------------------
// Camera (cameras.rs:146-149)
// ---------------------------------------------------------------------------
struct Cam {
  f3 origin, llc, horiz, vert, fwd;
  float area;
};
AR_HD Cam cam_from(const ar_camera& c) {
  Cam k;
  k.origin = mk(c.origin[0], c.origin[1], c.origin[2]);
  k.llc = mk(c.lower_left[0], c.lower_left[1], c.lower_left[2]);
  k.horiz = mk(c.horizontal[0], c.horizontal[1], c.horizontal[2]);
  k.vert = mk(c.vertical[0], c.vertical[1], c.vertical[2]);
  k.fwd = mk(c.forward[0], c.forward[1], c.forward[2]);
  k.area = c.area;
  return k;
}
AR_HD Ray camera_ray(const Cam& c, float u, float v) {
  f3 d = ((c.llc + c.horiz * u) + c.vert * v) - c.origin;
  Ray r;
  r.o = c.origin;
  r.d = normalize(d);
  r.tmin = 0.001f;
  r.tmax = kF32Max;
  return r;
}
AR_HD Ray make_ray(f3 o, f3 d, float tmin, float tmax) {
  Ray r;
  r.o = o;
  r.d = normalize(d);
  r.tmin = tmin;
  r.tmax = tmax;
  return r;
}

}  // namespace ar
