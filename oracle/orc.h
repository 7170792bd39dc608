// Oracle: a CPU restatement of the reference renderer's hot path
// (shiinamiyuki/agent-ray, Rust).  TEST INFRASTRUCTURE ONLY: it is the
// checker that the HIP library is compared against, and the CPU baseline that
// bench.py times.  Nothing in the product library links, loads or calls it.
//
// Arithmetic follows the reference's fp32 operation order (glam 0.32 SSE2
// kernels as restated below: dot = (x*x'+y*y')+z*z', cross = textbook
// two-product differences, normalize = v * (1/sqrt(dot(v,v))), lane min/max
// with SSE select semantics).  Build with -ffp-contract=off -fno-fast-math so
// the compiler never fuses a multiply and an add (Rust never contracts).
//
// Parity status of each part is written next to it; "parity unpinned" marks a
// function that no reference test, fixture or closed form pins bit-for-bit.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace orc {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;  // FRAC_1_PI
constexpr float kF32Eps = 1.1920929e-07f;          // f32::EPSILON
constexpr float kF32Max = 3.40282347e+38f;         // f32::MAX

// SSE MINPS/MAXPS select semantics: (a < b) ? a : b.  glam's Vec3A::min/max.
inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }
// Rust f32::min / f32::max (IEEE minNum/maxNum: a NaN operand is ignored).
inline float rs_min(float a, float b) { return std::fmin(a, b); }
inline float rs_max(float a, float b) { return std::fmax(a, b); }
// Rust f32::signum: +1 for +0 and positives, -1 for -0 and negatives.
inline float rs_signum(float a) {
  if (std::isnan(a)) return a;
  return std::signbit(a) ? -1.0f : 1.0f;
}

struct V3 {
  float x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  static V3 splat(float s) { return V3(s, s, s); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// glam Vec3A::dot (SSE2 dot3_in_x): (x*x' + y*y') + z*z'.
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glam Vec3A::cross: (zxy*rhs - self*rhs.zxy).zxy == textbook component order.
inline V3 cross(V3 a, V3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length_sq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glam normalize: self * (1.0 / length).
inline V3 normalize(V3 a) {
  float r = 1.0f / length(a);
  return a * r;
}
inline V3 vmin(V3 a, V3 b) { return V3(sse_min(a.x, b.x), sse_min(a.y, b.y), sse_min(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(sse_max(a.x, b.x), sse_max(a.y, b.y), sse_max(a.z, b.z)); }
// glam Vec3A::max_element / min_element (SSE2 shuffle reduction):
// max(max(x, z), max(y, z)).
inline float max_element(V3 a) { return sse_max(sse_max(a.x, a.z), sse_max(a.y, a.z)); }
inline float min_element(V3 a) { return sse_min(sse_min(a.x, a.z), sse_min(a.y, a.z)); }
inline V3 vabs(V3 a) { return V3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }

// Column-major 4x4 (glam Mat4: x_axis, y_axis, z_axis, w_axis columns).
struct M4 {
  float c[4][4];  // c[col][row]
  static M4 identity() {
    M4 m;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) m.c[i][j] = (i == j) ? 1.0f : 0.0f;
    return m;
  }
};
// transform_point3a: ((c0*x + c1*y) + c2*z) + c3, unfused (parity unpinned for
// non-identity matrices: glam's mul/add grouping is restated, not verified).
inline V3 transform_point(const M4& m, V3 p) {
  V3 r;
  float* o = &r.x;
  for (int k = 0; k < 3; ++k) {
    float a = m.c[0][k] * p.x;
    a = m.c[1][k] * p.y + a;
    a = m.c[2][k] * p.z + a;
    a = m.c[3][k] + a;
    o[k] = a;
  }
  return r;
}
inline V3 transform_vector(const M4& m, V3 v) {
  V3 r;
  float* o = &r.x;
  for (int k = 0; k < 3; ++k) {
    float a = m.c[0][k] * v.x;
    a = m.c[1][k] * v.y + a;
    a = m.c[2][k] * v.z + a;
    o[k] = a;
  }
  return r;
}
M4 inverse(const M4& m);  // cofactor expansion; parity unpinned vs glam
bool is_identity(const M4& m);

// 3x3 column-major (glam Mat3) used for the normal matrix.
struct M3 {
  float c[3][3];
};
M3 normal_matrix(const M4& m);  // Mat3::from_mat4(m).inverse().transpose()
inline V3 mul(const M3& m, V3 v) {
  // Mat3 * Vec3: (c0*x + c1*y) + c2*z.
  V3 r;
  float* o = &r.x;
  for (int k = 0; k < 3; ++k) {
    float a = m.c[0][k] * v.x;
    a = a + m.c[1][k] * v.y;
    a = a + m.c[2][k] * v.z;
    o[k] = a;
  }
  return r;
}

// ---------------------------------------------------------------------------
// geometry.rs
// ---------------------------------------------------------------------------
// Ray (geometry.rs:4-24); Ray::new normalizes the direction (geometry.rs:12-20).
struct Ray {
  V3 o, d;
  float tmin, tmax;
  static Ray make(V3 o, V3 d, float tmin, float tmax) {
    Ray r;
    r.o = o;
    r.d = normalize(d);
    r.tmin = tmin;
    r.tmax = tmax;
    return r;
  }
  V3 at(float t) const { return o + d * t; }  // geometry.rs:22-24
};

// AABB (geometry.rs:27-99).
struct AABB {
  V3 lo, hi;
  static AABB empty() {
    return AABB{V3::splat(INFINITY), V3::splat(-INFINITY)};
  }
  AABB merge(const AABB& b) const { return AABB{vmin(lo, b.lo), vmax(hi, b.hi)}; }
  // geometry.rs:54-57
  float surface_area() const {
    V3 d = vmax(hi - lo, V3(0, 0, 0));
    return 2.0f * ((d.x * d.y + d.y * d.z) + d.z * d.x);
  }
  V3 centroid() const { return (lo + hi) * 0.5f; }  // geometry.rs:60-62
  AABB transform(const M4& m) const;                 // geometry.rs:66-84
  // Slab test, geometry.rs:86-98 (inv_dir recomputed per call in the
  // reference; bitwise identical when precomputed).
  bool intersect(const Ray& r) const {
    V3 inv = V3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    V3 t0 = (lo - r.o) * inv;
    V3 t1 = (hi - r.o) * inv;
    V3 tmn = vmin(t0, t1);
    V3 tmx = vmax(t0, t1);
    float a = rs_max(max_element(tmn), r.tmin);
    float b = rs_min(min_element(tmx), r.tmax);
    return a <= b;
  }
};

struct HitInfo {
  float t;
  V3 n;
  float u, v;
};

// Moller-Trumbore, geometry.rs:171-206.  Pinned bit-for-bit by closed-form
// known-answer tests and brute force (tests/test_oracle.py).
inline bool triangle_hit(V3 v0, V3 e1, V3 e2, const Ray& r, HitInfo* out) {
  V3 h = cross(r.d, e2);
  float a = dot(e1, h);
  if (std::fabs(a) < 1e-8f) return false;
  float f = 1.0f / a;
  V3 s = r.o - v0;
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  V3 q = cross(s, e1);
  float v = f * dot(r.d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t < r.tmin || t > r.tmax) return false;
  if (out) {
    out->t = t;
    out->n = normalize(cross(e1, e2));
    out->u = u;
    out->v = v;
  }
  return true;
}

// Onb (geometry.rs:211-245).
struct Onb {
  V3 n, t, b;
  static Onb from_normal(V3 nn) {
    Onb o;
    o.n = normalize(nn);
    V3 a = std::fabs(o.n.x) > 0.9f ? V3(0, 1, 0) : V3(1, 0, 0);
    o.b = normalize(cross(o.n, a));
    o.t = cross(o.b, o.n);
    return o;
  }
  V3 to_local(V3 v) const { return V3(dot(v, t), dot(v, b), dot(v, n)); }
  // tangent*x + bitangent*y + normal*z, left to right.
  V3 to_world(V3 v) const { return (t * v.x + b * v.y) + n * v.z; }
};

// ---------------------------------------------------------------------------
// accel/bvh.rs
// ---------------------------------------------------------------------------
struct BVHNode {  // bvh.rs:32-51
  AABB box;
  uint32_t left_or_start, right_or_end, count;
  bool is_leaf() const { return count > 0; }
};

struct BVH {
  std::vector<BVHNode> nodes;
  std::vector<uint32_t> prims;  // re-ordered primitive indices
};

// Full-sweep SAH builder (bvh.rs:57-236).  Sort ties are broken by primitive
// index (the reference's sort_unstable_by leaves tie order unspecified:
// topology under exact centroid ties is parity unpinned).
BVH build_bvh(const std::vector<AABB>& prim_boxes);

struct TraceCounters {
  uint64_t nodes = 0;  // nodes popped and box-tested
  uint64_t tris = 0;   // Triangle::hit calls
};

struct Mesh {  // primitives/mesh.rs:7-16
  std::vector<V3> pos;
  std::vector<V3> nrm;  // empty when absent
  std::vector<float> uv;  // tex_coords, (u, v) per vertex; empty when absent
  std::vector<uint32_t> idx;
  std::vector<uint32_t> slots;
  uint32_t ntri() const { return (uint32_t)(idx.size() / 3); }
  AABB prim_aabb(uint32_t i) const;  // mesh.rs:81-94
};

struct RayHit {  // bvh.rs:10-20
  uint32_t inst, prim;
  V3 n;
  float u, v, t;
};

struct Blas {
  std::shared_ptr<Mesh> mesh;
  BVH bvh;
  AABB aabb;
  void build();
  bool intersect(const Ray& r, RayHit* hit, TraceCounters* c) const;  // bvh.rs:350-354
};

struct Instance {
  Blas* blas;
  M4 xf, inv;
  AABB world;
};

struct Tlas {
  std::vector<Instance> inst;
  BVH bvh;
  void build();
  bool intersect(const Ray& r, RayHit* hit, TraceCounters* c) const;  // bvh.rs:413-447
};

}  // namespace orc
