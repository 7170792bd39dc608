// Device-side data layouts and fp32 math for the MI355X (gfx950) hot path.
//
// All arithmetic that feeds a ray/triangle/box decision follows the
// reference's operation order (agent-ray src/geometry.rs with glam 0.32 SSE2
// semantics): dot = (x*x'+y*y')+z*z', cross = two-product differences,
// normalize = v * (1/sqrt(dot(v,v))).  The library is compiled with
// -ffp-contract=off so no multiply/add pair is fused, and HIP's default
// correctly rounded fp32 division and sqrt are kept.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AR_HD __host__ __device__ __forceinline__

// Bit casts usable in host and device code.
AR_HD float bits_f(int v) { return __builtin_bit_cast(float, v); }
AR_HD float bits_f(uint32_t v) { return __builtin_bit_cast(float, v); }
AR_HD int bits_i(float v) { return __builtin_bit_cast(int, v); }
AR_HD uint32_t bits_u(float v) { return __builtin_bit_cast(uint32_t, v); }

namespace ar {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kF32Eps = 1.1920929e-07f;
constexpr float kF32Max = 3.40282347e+38f;
constexpr uint32_t kNoHit = 0xffffffffu;

struct f3 {
  float x, y, z;
};
AR_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
AR_HD f3 splat(float s) { return f3{s, s, s}; }
AR_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AR_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AR_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AR_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
AR_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
AR_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
AR_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
AR_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
AR_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AR_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
AR_HD float length(f3 a) { return sqrtf(dot(a, a)); }
AR_HD f3 normalize(f3 a) {
  float r = 1.0f / sqrtf(dot(a, a));
  return a * r;
}
AR_HD float maxc(f3 a) { return fmaxf(fmaxf(a.x, a.z), fmaxf(a.y, a.z)); }
// Rust f32::signum (+1 for +0).
AR_HD float signum(float a) { return __builtin_copysignf(1.0f, a); }

// ---------------------------------------------------------------------------
// Acceleration structure layout in HBM
// ---------------------------------------------------------------------------
// Wide binary node: both children's boxes live in the parent so that one
// 64-byte fetch (four dwordx4 loads, one cache line half) tests two boxes.
//   a = (L.lo.x, L.lo.y, L.lo.z, L.hi.x)
//   b = (L.hi.y, L.hi.z, R.lo.x, R.lo.y)
//   c = (R.lo.z, R.hi.x, R.hi.y, R.hi.z)
//   d = (left ref, right ref, -, -)
// A child ref >= 0 is an internal node index; ref < 0 is a leaf whose first
// primitive record is at ~ref.  Leaf records carry a "last in leaf" bit.
struct alignas(16) WNode {
  float4 a, b, c;
  int4 d;
};
static_assert(sizeof(WNode) == 64, "node layout");

// 4-wide node of the near-first traversal.  Child c's box on axis k is
// [org[k] + qlo[k][c] * scl[k], org[k] + qhi[k][c] * scl[k]] (scl a power of
// two, so q * scl is exact and one rounding remains); the builder rounds the
// 8-bit planes outward until every dequantized box contains the child's
// exact box.  Empty slots have ref == kEmptyRef.  Load as four 16 B words:
//   a = (org.xyz, scl.x)  b = (scl.y, scl.z, qlo.x, qlo.y)
//   c = (qlo.z, qhi.x, qhi.y, qhi.z)  d = ref[0..3]
// where each q word packs children 0..3 in bytes 0..3.
struct alignas(64) QNode {
  float4 a, b, c;
  int4 d;
};
static_assert(sizeof(QNode) == 64, "qnode layout");
constexpr int kEmptyRef = (int)0x80000001;

// Triangle in BLAS leaf order: v0, e1 = v1 - v0, e2 = v2 - v0 (computed in
// fp32 exactly as the reference's Triangle::new).  v0.w = prim id | last bit.
struct alignas(16) Tri {
  float4 v0, e1, e2;
};
constexpr uint32_t kLastBit = 0x80000000u;
// A 4-wide tree leaf ref ~(k | kSingleTri) stands for Tri record k alone (its
// own quantized box in the parent); ~k without it for the records from k to
// the one tagged kLastBit (a binary leaf's range).  Records stay below 2^30.
// In the TLAS's 4-wide tree the same bit marks one instance record alone,
// in a slot holding its world box (SceneView::qinsts marks it last in its
// leaf).
constexpr uint32_t kSingleTri = 0x40000000u;

// One instance (TLAS leaf record), in TLAS leaf order.
struct alignas(16) Inst {
  float inv[12];      // world-to-local affine, column-major 3x4 (col c: inv[3c..3c+2])
  float box[6];       // BLAS root box (local space)
  int root;           // BLAS root ref (node index, or ~tri start for a leaf root)
  uint32_t id;        // object id | last-in-TLAS-leaf bit
  uint32_t identity;  // 1 when the transform is the identity
  uint32_t tri_base;  // first Tri record of this BLAS (global index)
  int qroot;          // BLAS root ref in the 4-wide node array
  uint32_t pad;
};
static_assert(sizeof(Inst) == 96, "instance layout");

struct SceneView {
  const WNode* nodes;   // binary trees (reference order)
  const QNode* qnodes;  // 4-wide trees (near-first order)
  const Tri* tris;       // binary trees' leaf order; hits name these records
  const Tri* qtris;      // 4-wide trees' leaf order (the same array as tris unless spatial splits)
  const Inst* insts;     // TLAS leaf order; hits name these records
  const Inst* qinsts;    // the same records for the 4-wide TLAS, kLastBit on every single-instance slot's
  int tlas_root;        // TLAS root ref (node index or ~first instance record)
  int qtlas_root;       // the same in the 4-wide node array
  float tlas_box[6];    // TLAS root box (world)
  uint32_t n_inst;
  uint32_t wide_ok;     // every 4-wide node's origin and scale are below 2^62 in magnitude
};

struct Ray {
  f3 o;
  float tmin;
  f3 d;
  float tmax;
};

// glam's SSE2 Vec3A min / max lanes (minps / maxps): the second operand
// when the comparison is false, so a NaN in either position follows the
// instruction, not IEEE minNum.
AR_HD float sel_min(float a, float b) { return a < b ? a : b; }
AR_HD float sel_max(float a, float b) { return a > b ? a : b; }

// Slab test with the reference's arithmetic and NaN behaviour
// (geometry.rs:86-98): t0 = (lo - o) * inv, t1 = (hi - o) * inv per axis,
// near = Vec3A::min(t0, t1), far = Vec3A::max(t0, t1), reduced as glam's
// max_element / min_element (max(max(x, z), max(y, z))), then clamped with
// f32::max(tmin) / f32::min(tmax) (a NaN operand ignored).  A ray lying in a
// box plane with a zero direction component makes 0 * inf = NaN, and the
// result then depends on exactly this order.  inv = 1/d precomputed
// (bitwise identical to recomputing per call).  Returns the entry distance
// in *tn.
AR_HD bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz, f3 o, f3 inv,
                float tmin, float tmax, float* tn) {
  float t0x = (lox - o.x) * inv.x, t1x = (hix - o.x) * inv.x;
  float t0y = (loy - o.y) * inv.y, t1y = (hiy - o.y) * inv.y;
  float t0z = (loz - o.z) * inv.z, t1z = (hiz - o.z) * inv.z;
  float nx = sel_min(t0x, t1x), fx = sel_max(t0x, t1x);
  float ny = sel_min(t0y, t1y), fy = sel_max(t0y, t1y);
  float nz = sel_min(t0z, t1z), fz = sel_max(t0z, t1z);
  float a = fmaxf(sel_max(sel_max(nx, nz), sel_max(ny, nz)), tmin);
  float b = fminf(sel_min(sel_min(fx, fz), sel_min(fy, fz)), tmax);
  *tn = a;
  return a <= b;
}

// The same decision with IEEE min / max, for an inv with no infinite
// component: no NaN can arise then, and min / max of non-NaN values agree
// with the select forms (a -0 / +0 difference compares equal).
AR_HD bool slab_fast(float lox, float loy, float loz, float hix, float hiy, float hiz, f3 o, f3 inv,
                     float tmin, float tmax, float* tn) {
  float t0x = (lox - o.x) * inv.x, t1x = (hix - o.x) * inv.x;
  float t0y = (loy - o.y) * inv.y, t1y = (hiy - o.y) * inv.y;
  float t0z = (loz - o.z) * inv.z, t1z = (hiz - o.z) * inv.z;
  float a = fmaxf(fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z)), tmin);
  float b = fminf(fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)), tmax);
  *tn = a;
  return a <= b;
}

// True when no component of inv is infinite (no zero or subnormal direction
// component): the slab test then never meets 0 * inf, and it is monotone in
// the box planes.
AR_HD bool finite_inv(f3 inv) {
  return fabsf(inv.x) < INFINITY && fabsf(inv.y) < INFINITY && fabsf(inv.z) < INFINITY;
}

// Moller-Trumbore (geometry.rs:171-206); returns t, u, v on hit.
AR_HD bool tri_hit(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float tmin, float tmax, float* t_out, float* u_out,
                   float* v_out) {
  f3 h = cross(d, e2);
  float a = dot(e1, h);
  if (fabsf(a) < 1e-8f) return false;
  float f = 1.0f / a;
  f3 s = o - v0;
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return false;
  f3 q = cross(s, e1);
  float v = f * dot(d, q);
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = f * dot(e2, q);
  if (t < tmin || t > tmax) return false;
  *t_out = t;
  *u_out = u;
  *v_out = v;
  return true;
}

// Affine transform with the reference's grouping ((c0*x + c1*y) + c2*z) + c3.
AR_HD f3 xform_point(const float* m, f3 p) {
  return f3{((m[0] * p.x + m[3] * p.y) + m[6] * p.z) + m[9], ((m[1] * p.x + m[4] * p.y) + m[7] * p.z) + m[10],
            ((m[2] * p.x + m[5] * p.y) + m[8] * p.z) + m[11]};
}
AR_HD f3 xform_vector(const float* m, f3 v) {
  return f3{(m[0] * v.x + m[3] * v.y) + m[6] * v.z, (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
            (m[2] * v.x + m[5] * v.y) + m[8] * v.z};
}

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 keyed by the job seed, counter =
// (dim / 4, sample, pixel, stream); word dim % 4; f32 = (w >> 8) * 2^-24.
// ---------------------------------------------------------------------------
AR_HD void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]) {
#ifdef __clang__
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Stateless-keyed sampler with a one-block cache (4 dimensions per Philox call).
struct Rng {
  uint32_t k0, k1, pixel, sample, stream, dim;
  uint32_t cache[4];
  uint32_t cached_block;
  AR_HD void init(uint64_t seed, uint32_t px, uint32_t smp, uint32_t strm, uint32_t d0 = 0) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    pixel = px;
    sample = smp;
    stream = strm;
    dim = d0;
    cached_block = 0xffffffffu;
  }
  AR_HD float next() {
    uint32_t blk = dim >> 2;
    if (blk != cached_block) {
      philox(blk, sample, pixel, stream, k0, k1, cache);
      cached_block = blk;
    }
    uint32_t w = cache[dim & 3];
    dim++;
    return (float)(w >> 8) * (1.0f / 16777216.0f);
  }
};

}  // namespace ar
