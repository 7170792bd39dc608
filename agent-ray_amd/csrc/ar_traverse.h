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
  AR_HD void push(int ref, float tnear) {
    int2 e = make_int2(ref, bits_i(tnear));
    if (sp < S)
      lds[sp * stride] = e;
    else
      spill[(size_t)(sp - S) * spill_stride] = e;
    ++sp;
  }
  AR_HD int2 pop() {
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
AR_HD bool blas_traverse(const WNode* __restrict__ nodes, const Tri* __restrict__ tris,
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
        uint32_t tag = bits_u(v0.w);
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
      if (bits_f(e.y) <= tmax) {
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
AR_HD bool visit_instance(const SceneView& sv, uint32_t k, const Ray& r, WorldHit& best,
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
AR_HD bool trace_scene(const SceneView& sv, const Ray& r, WorldHit& best, Stack& stk) {
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
      if (bits_f(e.y) <= best.t) {
        cur = e.x;
        break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Resumable form of trace_scene.  The whole state of one query lives in the
// Walker (registers), so a lane can finish its ray and start another while
// the other lanes of its wave are still walking (walk_items below) instead
// of idling until the wave's slowest ray ends.
//
// Work is split in two phases so that a wave's lanes run the same code at
// the same time: descend() visits internal nodes until the lane holds a
// leaf (or its level's stack is exhausted); leaf() then does the rarer work
// -- a leaf's triangles, entering or leaving an instance, ending the query.
//
// Arithmetic and decisions are those of trace_scene / visit_instance /
// blas_traverse, operation for operation, so results are bit-identical:
//   TLAS level: ray (wo, 1/wd, wtmin, bt), pops accepted when tnear <= bt;
//   BLAS level: local ray (M^-1 o, normalize(M^-1 d)), range (wtmin*s, bt*s)
//   at entry, local acceptance by strict < (NEAR ties: lower leaf position),
//   world acceptance of the local best as t/s < bt (NEAR ties: lower
//   instance record); any-hit ends at the first hit with t/s < bt.
// ---------------------------------------------------------------------------
constexpr int kLevelEnd = (int)0x80000000;  // cur: the current level's stack is exhausted

template <bool ANY, bool NEAR>
struct Walker {
  f3 wo, wd;              // world ray
  float wtmin;
  float bt;               // world best t (closest) or world tmax (any)
  uint32_t bpos, binst;   // world best: Tri record, instance record
  float bu, bv;
  f3 o, d, inv;           // ray of the current level
  float tmin, tmax;       // range of the current level (BLAS: tmax = local best once found)
  float s;                // BLAS: |M^-1 d|
  uint32_t lpos;          // BLAS-local best Tri record (kNoHit: none yet)
  float lu, lv;
  uint32_t ik;            // instance record being walked
  int base;               // stack depth at instance entry
  int cur;                // node ref, leaf ref, or kLevelEnd
  bool in_blas;

  // Starts a query.  Returns true when it is already finished (the TLAS
  // root box is missed).
  AR_HD bool start(const SceneView& sv, const Ray& r) {
    wo = r.o;
    wd = r.d;
    wtmin = r.tmin;
    bt = r.tmax;
    bpos = kNoHit;
    binst = kNoHit;
    bu = bv = 0.0f;
    o = wo;
    d = wd;
    inv = mk(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
    tmin = wtmin;
    tmax = bt;
    in_blas = false;
    cur = sv.tlas_root;
    float tn;
    const float* tb = sv.tlas_box;
    return !slab(tb[0], tb[1], tb[2], tb[3], tb[4], tb[5], o, inv, tmin, tmax, &tn);
  }

  // Pops the deferred nodes of the current level until one is in range;
  // kLevelEnd when the level has none left.
  AR_HD void pop_level(Stack& stk) {
    const int lo = in_blas ? base : 0;
    for (;;) {
      if (stk.sp == lo) {
        cur = kLevelEnd;
        return;
      }
      int2 e = stk.pop();
      if (bits_f(e.y) <= tmax) {
        cur = e.x;
        return;
      }
    }
  }

  // Visits internal nodes until the lane holds a leaf or kLevelEnd.
  AR_HD void descend(const SceneView& sv, Stack& stk) {
    while (cur >= 0) {
      const WNode* n = sv.nodes + cur;
      float4 a = n->a, b = n->b, c = n->c;
      int4 ch = n->d;
      float tl, tr;
      bool hl = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tmin, tmax, &tl);
      bool hr = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tmin, tmax, &tr);
      if (hl && hr) {
        bool swap = NEAR && (tr < tl);
        stk.push(swap ? ch.x : ch.y, swap ? tl : tr);
        cur = swap ? ch.y : ch.x;
      } else if (hl) {
        cur = ch.x;
      } else if (hr) {
        cur = ch.y;
      } else {
        pop_level(stk);
      }
    }
  }

  // Enters the first instance from record k on (same TLAS leaf) whose BLAS
  // root box the local ray hits.  False when the leaf has none left.
  AR_HD bool enter(const SceneView& sv, uint32_t k, const Stack& stk) {
    for (;;) {
      const Inst* in = sv.insts + k;
      f3 lo, ld;
      if (in->identity) {
        lo = wo;
        ld = wd;
      } else {
        lo = xform_point(in->inv, wo);
        ld = xform_vector(in->inv, wd);
      }
      float sc = length(ld);
      if (sc >= kF32Eps) {
        f3 ldn = normalize(ld);
        f3 li = mk(1.0f / ldn.x, 1.0f / ldn.y, 1.0f / ldn.z);
        float a = wtmin * sc, b = bt * sc;
        float tn;
        const float* rb = in->box;
        if (slab(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], lo, li, a, b, &tn)) {
          o = lo;
          d = ldn;
          inv = li;
          tmin = a;
          tmax = b;
          s = sc;
          ik = k;
          base = stk.sp;
          cur = in->root;
          in_blas = true;
          lpos = kNoHit;
          lu = lv = 0.0f;
          return true;
        }
      }
      if (in->id & kLastBit) return false;
      ++k;
    }
  }

  // Leaf phase: the lane holds a leaf ref or kLevelEnd.  Returns true when
  // the query is done; otherwise cur is a node, a leaf or kLevelEnd again.
  AR_HD bool leaf(const SceneView& sv, Stack& stk) {
    if (cur == kLevelEnd) {
      if (!in_blas) return true;  // TLAS exhausted: query done
      // instance finished: world acceptance of its local best
      if (!ANY && lpos != kNoHit) {
        float tw = tmax / s;
        if (tw < bt || (NEAR && bpos != kNoHit && tw == bt && ik < binst)) {
          bt = tw;
          bpos = lpos;
          binst = ik;
          bu = lu;
          bv = lv;
        }
      }
      in_blas = false;
      o = wo;
      d = wd;
      inv = mk(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
      tmin = wtmin;
      tmax = bt;
      if (!(sv.insts[ik].id & kLastBit) && enter(sv, ik + 1, stk)) return false;
    } else if (in_blas) {
      uint32_t k = (uint32_t)~cur;
      for (;;) {
        const Tri* tp = sv.tris + k;
        float4 v0 = tp->v0, e1 = tp->e1, e2 = tp->e2;
        uint32_t tag = bits_u(v0.w);
        float th, uh, vh;
        if (tri_hit(mk(v0.x, v0.y, v0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), o, d, tmin, tmax, &th,
                    &uh, &vh)) {
          if (ANY) {
            if (th / s < bt) {
              bpos = k;
              binst = ik;
              stk.sp = 0;
              return true;
            }
          } else if (th < tmax || (NEAR && lpos != kNoHit && th == tmax && k < lpos)) {
            tmax = th;
            lpos = k;
            lu = uh;
            lv = vh;
          }
        }
        if (tag & kLastBit) break;
        ++k;
      }
    } else if (enter(sv, (uint32_t)~cur, stk)) {
      return false;
    }
    pop_level(stk);
    return false;
  }
};

// Items [0, n) are split in chunks of 64; wave `wave` of `waves` owns chunks
// wave, wave + waves, ... and walks their items in order.  A lane that
// finishes takes the next unassigned item of its wave; idle lanes are
// refilled together once at least `refill_min` (1..64) of them wait, so
// items fetched together are neighbours (coherent rays stay together).
// Every item is started exactly once, and the loop ends when the wave's
// items are exhausted and every lane is idle.  n < 2^32.
//   src(i)      -> Ray for item i
//   sink(i, w)  consumes the finished walker of item i
template <bool ANY, bool NEAR, class Src, class Sink>
__device__ __forceinline__ void walk_items(const SceneView& sv, Stack& stk, uint64_t n, uint64_t wave,
                                           uint64_t waves, uint32_t refill_min, Src src, Sink sink) {
  const uint32_t lane = __lane_id();
  const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t next = 0;  // wave-uniform: position of the next unassigned item
  uint32_t item = 0;
  bool alive = false;
  Walker<ANY, NEAR> w;
  for (;;) {
    uint64_t idle = __ballot(!alive);
    uint32_t nidle = (uint32_t)__popcll(idle);
    uint64_t first = (wave + (next >> 6) * waves) * 64 + (next & 63);
    bool more = first < n;
    if (!more && nidle == 64) break;
    if (more && nidle >= refill_min) {
      if (!alive) {
        uint32_t p = next + (uint32_t)__popcll(idle & below);
        uint64_t it = (wave + (p >> 6) * waves) * 64 + (p & 63);
        if (it < n) {
          item = (uint32_t)it;
          alive = !w.start(sv, src(it));
          if (!alive) sink(it, w);
        }
      }
      next += nidle;
    }
    if (alive) {
      w.descend(sv, stk);
      if (w.leaf(sv, stk)) {
        sink(item, w);
        alive = false;
      }
    }
  }
}

}  // namespace ar
