// Two-level BVH traversal for gfx950 (one ray per lane).
//
// Semantics follow the reference's TLAS::intersect / traverse_bvh
// (agent-ray src/accel/bvh.rs:247-296, 413-447):
//   * a node's children are box-tested against the current best t; a
//     deferred child is pushed with its entry distance and is skipped at pop
//     time when that distance exceeds the (shrunken) best t — exactly the
//     reference's "test the box when popped" decision;
//   * triangles in a leaf are tested in leaf order with Triangle::hit and
//     accepted with strict `<`;
//   * per instance: local ray = (M^-1 o, normalize(M^-1 d)), t range scaled
//     by s = |M^-1 d|, BLAS hit t divided by s, accepted against the world
//     best with strict `<`.
// REF order (left child first) reproduces the reference's results bit for
// bit, including which of two equal-t triangles wins.  NEAR order visits the
// nearer child first and breaks equal-t ties by leaf position, which is the
// reference's DFS order, so the same primitive wins.
//
// Stack: S entries per lane live in LDS (lane-strided so the 64 lanes of a
// wave hit distinct banks); deeper entries spill to a global array laid out
// [entry][thread] so spills coalesce.
#pragma once
#include "ar_device.h"

namespace ar {

struct Stack {
  int2* lds;          // &lds_base[threadIdx.x]; entry i at lds[i * stride]
  int2* spill;        // &spill_base[global thread]; entry i >= S at spill[(i - S) * spill_stride]
  uint32_t stride;
  uint32_t spill_stride;
  int S;
  int sp;
  __device__ __forceinline__ void push(int ref, float tnear) {
    int2 e = make_int2(ref, __float_as_int(tnear));
    if (sp < S)
      lds[sp * stride] = e;
    else
      spill[(size_t)(sp - S) * spill_stride] = e;
    ++sp;
  }
  __device__ __forceinline__ int2 pop() {
    --sp;
    return sp < S ? lds[sp * stride] : spill[(size_t)(sp - S) * spill_stride];
  }
};

struct LocalHit {
  float t;        // local t of the best hit (valid when pos != kNoHit)
  uint32_t pos;   // global Tri record index
  float u, v;
};

// BLAS traversal from `root` (box `rb`).  Closest mode: finds the best hit
// with t < tmax (strict), shrinking tmax.  Any mode: returns true at the
// first hit whose world t (t / scale) is < world_tmax.
template <bool ANY, bool NEAR>
__device__ __forceinline__ bool blas_traverse(const WNode* __restrict__ nodes, const Tri* __restrict__ tris,
                                              int root, const float* rb, f3 o, f3 d, float tmin, float tmax,
                                              LocalHit& best, Stack& stk, float scale, float world_tmax) {
  f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  float tn;
  if (!slab(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], o, inv, tmin, tmax, &tn)) return false;
  const int base = stk.sp;
  bool found = false;
  int cur = root;
  for (;;) {
    if (cur >= 0) {
      const WNode* n = nodes + cur;
      float4 a = n->a, b = n->b, c = n->c;
      int4 ch = n->d;
      float tl, tr;
      bool hl = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tmin, tmax, &tl);
      bool hr = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tmin, tmax, &tr);
      if (hl && hr) {
        bool swap = NEAR && (tr < tl);
        stk.push(swap ? ch.x : ch.y, swap ? tl : tr);
        cur = swap ? ch.y : ch.x;
        continue;
      }
      if (hl) {
        cur = ch.x;
        continue;
      }
      if (hr) {
        cur = ch.y;
        continue;
      }
    } else {
      int k = ~cur;
      for (;;) {
        const Tri* tp = tris + k;
        float4 v0 = tp->v0, e1 = tp->e1, e2 = tp->e2;
        uint32_t tag = __float_as_uint(v0.w);
        float th, uh, vh;
        if (tri_hit(mk(v0.x, v0.y, v0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), o, d, tmin, tmax, &th,
                    &uh, &vh)) {
          if (ANY) {
            if (th / scale < world_tmax) return true;
          } else if (th < tmax || (NEAR && found && th == tmax && (uint32_t)k < best.pos)) {
            tmax = th;
            best.t = th;
            best.pos = (uint32_t)k;
            best.u = uh;
            best.v = vh;
            found = true;
          }
        }
        if (tag & kLastBit) break;
        ++k;
      }
    }
    // pop the next deferred node still in range
    for (;;) {
      if (stk.sp == base) return found;
      int2 e = stk.pop();
      if (__int_as_float(e.y) <= tmax) {
        cur = e.x;
        break;
      }
    }
  }
}

struct WorldHit {
  float t;
  uint32_t pos;   // Tri record (global leaf order), kNoHit on miss
  uint32_t inst;  // instance record index (TLAS leaf order)
  float u, v;
};

// Instance leaf visit: transform, BLAS traverse, world acceptance.
template <bool ANY, bool NEAR>
__device__ __forceinline__ bool visit_instance(const SceneView& sv, uint32_t k, const Ray& r, WorldHit& best,
                                               Stack& stk) {
  const Inst* in = sv.insts + k;
  f3 lo, ld;
  if (in->identity) {
    lo = r.o;
    ld = r.d;
  } else {
    lo = xform_point(in->inv, r.o);
    ld = xform_vector(in->inv, r.d);
  }
  float s = length(ld);
  if (s < kF32Eps) return false;
  f3 ldn = normalize(ld);  // Ray::new renormalizes
  float wt = best.t;
  LocalHit lh;
  lh.pos = kNoHit;
  lh.t = 0.0f;
  lh.u = lh.v = 0.0f;
  bool hit = blas_traverse<ANY, NEAR>(sv.nodes, sv.tris, in->root, in->box, lo, ldn, r.tmin * s, wt * s, lh, stk,
                                      s, wt);
  if (ANY) return hit;
  if (!hit) return false;
  float tw = lh.t / s;
  if (tw < best.t || (NEAR && best.pos != kNoHit && tw == best.t && k < best.inst)) {
    best.t = tw;
    best.pos = lh.pos;
    best.inst = k;
    best.u = lh.u;
    best.v = lh.v;
    return true;
  }
  return false;
}

// Full query.  Closest: fills best (best.pos == kNoHit on miss).  Any:
// returns whether anything is hit in [tmin, tmax).
template <bool ANY, bool NEAR>
__device__ __forceinline__ bool trace_scene(const SceneView& sv, const Ray& r, WorldHit& best, Stack& stk) {
  best.t = r.tmax;
  best.pos = kNoHit;
  best.inst = kNoHit;
  best.u = best.v = 0.0f;
  f3 inv = mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
  float tn;
  const float* tb = sv.tlas_box;
  if (!slab(tb[0], tb[1], tb[2], tb[3], tb[4], tb[5], r.o, inv, r.tmin, r.tmax, &tn)) return false;
  const int base = stk.sp;
  bool found = false;
  int cur = sv.tlas_root;
  for (;;) {
    if (cur >= 0) {
      const WNode* n = sv.nodes + cur;
      float4 a = n->a, b = n->b, c = n->c;
      int4 ch = n->d;
      float tl, tr;
      bool hl = slab(a.x, a.y, a.z, a.w, b.x, b.y, r.o, inv, r.tmin, best.t, &tl);
      bool hr = slab(b.z, b.w, c.x, c.y, c.z, c.w, r.o, inv, r.tmin, best.t, &tr);
      if (hl && hr) {
        bool swap = NEAR && (tr < tl);
        stk.push(swap ? ch.x : ch.y, swap ? tl : tr);
        cur = swap ? ch.y : ch.x;
        continue;
      }
      if (hl) {
        cur = ch.x;
        continue;
      }
      if (hr) {
        cur = ch.y;
        continue;
      }
    } else {
      uint32_t k = (uint32_t)~cur;
      for (;;) {
        uint32_t id = sv.insts[k].id;
        if (visit_instance<ANY, NEAR>(sv, k, r, best, stk)) {
          if (ANY) {
            stk.sp = base;
            return true;
          }
          found = true;
        }
        if (id & kLastBit) break;
        ++k;
      }
    }
    for (;;) {
      if (stk.sp == base) return found;
      int2 e = stk.pop();
      if (__int_as_float(e.y) <= best.t) {
        cur = e.x;
        break;
      }
    }
  }
}

}  // namespace ar
