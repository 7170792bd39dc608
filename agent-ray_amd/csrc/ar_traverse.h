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
// REF order (binary trees, left child first) reproduces the reference's
// results bit for bit, including which of two equal-t triangles wins.  NEAR
// order (4-wide trees, nearest child first) breaks equal-t ties by leaf
// position, which is the reference's DFS order, so the same primitive wins.
//
// Stack: S entries per lane live in LDS (lane-strided so the 64 lanes of a
// wave hit distinct banks); deeper entries spill to a global array laid out
// [entry][thread] so spills coalesce.
#pragma once
#include "ar_device.h"

// Instrumentation hooks, empty in the library's production kernels.  The
// counting build (trace_count.hip) defines them and compiles this header in
// a namespace of its own (AR_TRAVERSE_NS), so its instrumented walker never
// shares a symbol with the production one.
#ifndef AR_COUNT_NODE
#define AR_COUNT_NODE()  // one 64-byte node record read (4-wide or binary)
#endif
#ifndef AR_COUNT_TRI
#define AR_COUNT_TRI()  // one 48-byte triangle record read and tested
#endif
#ifndef AR_COUNT_INST
#define AR_COUNT_INST()  // one 96-byte instance record read at a TLAS leaf
#endif
#ifndef AR_COUNT_SPILL_W
#define AR_COUNT_SPILL_W()  // one 8-byte stack entry written to the global spill
#endif
#ifndef AR_COUNT_SPILL_R
#define AR_COUNT_SPILL_R()  // one 8-byte stack entry read back from the global spill
#endif
#ifndef AR_TRAVERSE_NS
#define AR_TRAVERSE_NS ar
#endif

namespace AR_TRAVERSE_NS {
using namespace ::ar;

// Stack pointers carry their address space in device code, so LDS entries
// are read and written with LDS instructions and spilled ones with global
// ones (a generic pointer would make the compiler merge both paths into
// flat accesses, which wait on both memory counters).
#if defined(__HIP_DEVICE_COMPILE__)
#define AR_LDS_AS __attribute__((address_space(3)))
#define AR_GLOBAL_AS __attribute__((address_space(1)))
#else
#define AR_LDS_AS
#define AR_GLOBAL_AS
#endif
using LdsEntryPtr = AR_LDS_AS int2*;
using SpillEntryPtr = AR_GLOBAL_AS unsigned long long*;  // (ref, tnear bits) packed low, high

struct Stack {
  LdsEntryPtr lds;      // &lds_base[threadIdx.x]; entry i at lds[i * stride]
  SpillEntryPtr spill;  // &spill_base[global thread]; entry i >= S at spill[(i - S) * spill_stride]
  uint32_t stride;
  uint32_t spill_stride;
  int S;
  int sp;
  AR_HD void push(int ref, float tnear) {
    int2 e = make_int2(ref, bits_i(tnear));
    if (sp < S) {
      lds[sp * stride] = e;
    } else {
      AR_COUNT_SPILL_W();
      spill[(size_t)(sp - S) * spill_stride] = (unsigned long long)(uint32_t)e.x | ((unsigned long long)(uint32_t)e.y << 32);
    }
    ++sp;
  }
  AR_HD int2 pop() {
    --sp;
    int2 e;
    if (sp < S) {
      e = lds[sp * stride];
    } else {
      AR_COUNT_SPILL_R();
      const unsigned long long v = spill[(size_t)(sp - S) * spill_stride];
      e = make_int2((int)(uint32_t)v, (int)(uint32_t)(v >> 32));
    }
    return e;
  }
};

struct WorldHit {
  float t;
  uint32_t pos;   // Tri record (global leaf order), kNoHit on miss
  uint32_t inst;  // instance record index (TLAS leaf order)
  float u, v;
};

// ---------------------------------------------------------------------------
// Two-level traversal as a resumable walker.  The whole state of one query
// lives in the Walker (registers), so a lane can finish its ray and start
// another while the other lanes of its wave are still walking (walk_items
// below) instead of idling until the wave's slowest ray ends; trace_scene
// runs one walker to completion.
//
// REF order walks the binary trees left child first (bvh.rs:247-296) and
// reproduces the reference bit for bit, ties included.  NEAR order walks the
// 4-wide trees (QNode), nearest child first.  Their quantized child boxes
// contain the exact ones, and with a finite 1/d the slab test is monotone in
// the box planes, so every leaf the exact tree reaches is reached and the
// closest hit is the same; equal-t hits are resolved by lower leaf position
// / instance record, which is the reference's DFS winner.  A level whose
// 1/d has an infinite component (the 0 * inf = NaN cases of the slab test)
// is walked on the binary tree instead, so those rays see the reference's
// exact box decisions.  What remains different is a hit that the exact
// slab test loses to rounding at a box face (a ray grazing the box edge at
// the hit point): the 4-wide boxes are larger and keep it.
//
// Work is split in two phases so that a wave's lanes run the same code at
// the same time: descend() visits internal nodes until the lane holds a
// leaf (or its level's stack is exhausted); leaf() then does the rarer work
// -- a leaf's triangles, entering or leaving an instance, ending the query.
//
// Per level (the reference's TLAS::intersect / Instance::intersect /
// traverse_bvh arithmetic, operation for operation):
//   TLAS level: ray (wo, 1/wd, wtmin, bt), pops accepted when tnear <= bt;
//   BLAS level: local ray (M^-1 o, normalize(M^-1 d)), range (wtmin*s, bt*s)
//   at entry, local acceptance by strict < (NEAR ties: lower leaf position),
//   world acceptance of the local best as t/s < bt (NEAR ties: lower
//   instance record); any-hit ends at the first hit with t/s < bt.
// ---------------------------------------------------------------------------
constexpr int kLevelEnd = (int)0x80000000;  // cur: the current level's stack is exhausted

// Near-first closest hit culls boxes, pops and triangle ranges against
// best * kCullSlack rather than best: a hit tied with the final best is then
// reached whatever the visiting order (the slack covers the few-ulp gap
// between a box's entry distance, a triangle's t and t * s / s), and ties
// are broken by (instance record, leaf position).  The answer is therefore
// independent of the order, which depends on the wave a ray shares.
constexpr float kCullSlack = 1.0f + 0x1p-18f;

// Whether a level's ray walks the 4-wide trees (near-first order): finite
// 1/d bounded by 2^32 and an origin bounded by 2^64, with node origins and
// scales below 2^62 (SceneView::wide_ok), keep every term of the 4-wide box
// test finite and far from underflow-only ranges; other rays walk the
// binary trees with the exact slab test.
AR_HD bool wide_ray(const SceneView& sv, f3 o, f3 inv) {
  const float mi = fmaxf(fmaxf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z));
  const float mo = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z));
  return sv.wide_ok != 0u && mi <= 0x1p32f && mo <= 0x1p64f;
}

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
  int park0, park1, park2, park3;  // parked triangle leaves in the order reached (kLevelEnd: free)
  bool in_blas;
  bool fast;              // 1/d of this level is finite: IEEE min / max slab
  bool wide;              // this level walks the 4-wide tree (near-first, fast)

  // Bound for culling and for the triangle range (see kCullSlack).
  AR_HD float tcull() const { return (NEAR && !ANY) ? tmax * kCullSlack : tmax; }

  // Starts a query on an empty stack.  Returns true when it is already
  // finished (the TLAS root box is missed, or the TLAS is one leaf whose
  // instances are all missed).
  AR_HD bool start(const SceneView& sv, const Ray& r, const Stack& stk) {
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
    park0 = park1 = park2 = park3 = kLevelEnd;
    fast = finite_inv(inv);
    wide = NEAR && fast && wide_ray(sv, o, inv);
    cur = wide ? sv.qtlas_root : sv.tlas_root;
    float tn;
    const float* tb = sv.tlas_box;
    if (!slab(tb[0], tb[1], tb[2], tb[3], tb[4], tb[5], o, inv, tmin, tmax, &tn)) return true;
    // a TLAS that is a single leaf: enter its instances right away
    if (cur < 0 && !enter(sv, (uint32_t)~cur, stk)) return true;  // (a range: single slots sit in nodes)
    return false;
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
      if (bits_f(e.y) <= tcull()) {
        cur = e.x;
        return;
      }
    }
  }

  // One 4-wide node: box-test the four children, continue with the nearest
  // hit, push the others farthest first.
  AR_HD void visit_wide(const SceneView& sv, Stack& stk) {
    AR_COUNT_NODE();
    const QNode* n = sv.qnodes + cur;
    const float4 a = n->a, b = n->b, c = n->c;
    const int4 rf = n->d;
    // Plane t per child and axis as one fma, q * A + B -/+ E, with A = scl / d
    // (exact: scl is a power of two) and B = (org - o) / d per node and axis.
    // With u = 2^-24, fmaf(q, A, B) differs from the (fmaf(q, scl, org) - o)
    // * (1 / d) of the exact-box test by at most about
    // u (770 |A| + 4.03 |B| + 1.01 |org / d|) (rounding of org + q scl and of
    // both formulas), and rounding B -/+ E and the fma adds at most
    // u (255 |A| + 2 |B| + 2 E).  So with, per axis,
    //   E = 2^-12 |A| + 2^-20 |B| + 2^-21 |org / d| + 2^-126
    // (2^-126 covers underflow; wide_ray keeps every term finite) the near
    // planes' q * A + (B - E) are not above, and the far planes'
    // q * A + (B + E) not below, the exact-box test's values.  Every child it
    // accepts is accepted here with an entry distance that is not larger:
    // the walk is conservative and the triangle tests decide the hit.  E is
    // per axis because an axis the ray is nearly parallel to has a huge A,
    // and its bound must not loosen the other axes.
    const float Ax = a.w * inv.x, Ay = b.x * inv.y, Az = b.y * inv.z;
    const float Bx = (a.x - o.x) * inv.x, By = (a.y - o.y) * inv.y, Bz = (a.z - o.z) * inv.z;
    const float Ex = fmaf(fabsf(Ax), 0x1p-12f, fmaf(fabsf(Bx), 0x1p-20f, fmaf(fabsf(a.x * inv.x), 0x1p-21f, 0x1p-126f)));
    const float Ey = fmaf(fabsf(Ay), 0x1p-12f, fmaf(fabsf(By), 0x1p-20f, fmaf(fabsf(a.y * inv.y), 0x1p-21f, 0x1p-126f)));
    const float Ez = fmaf(fabsf(Az), 0x1p-12f, fmaf(fabsf(Bz), 0x1p-20f, fmaf(fabsf(a.z * inv.z), 0x1p-21f, 0x1p-126f)));
    const float Bnx = Bx - Ex, Bny = By - Ey, Bnz = Bz - Ez;
    const float Bfx = Bx + Ex, Bfy = By + Ey, Bfz = Bz + Ez;
    // With finite 1/d the entry plane of every child on axis k is its low
    // plane when 1/d_k >= 0 and its high plane otherwise.
    const bool px = inv.x >= 0.0f, py = inv.y >= 0.0f, pz = inv.z >= 0.0f;
    const uint32_t lx = bits_u(b.z), ly = bits_u(b.w), lz = bits_u(c.x);
    const uint32_t hx = bits_u(c.y), hy = bits_u(c.z), hz = bits_u(c.w);
    const uint32_t nx = px ? lx : hx, ny = py ? ly : hy, nz = pz ? lz : hz;
    const uint32_t fx = px ? hx : lx, fy = py ? hy : ly, fz = pz ? hz : lz;
    const float tc = tcull();
    float t[4];
    int r[4] = {rf.x, rf.y, rf.z, rf.w};
#ifdef __clang__
#pragma unroll
#endif
    for (int k = 0; k < 4; ++k) {
      const int sh = 8 * k;
      const float tnx = fmaf((float)((nx >> sh) & 255u), Ax, Bnx);
      const float tny = fmaf((float)((ny >> sh) & 255u), Ay, Bny);
      const float tnz = fmaf((float)((nz >> sh) & 255u), Az, Bnz);
      const float tfx = fmaf((float)((fx >> sh) & 255u), Ax, Bfx);
      const float tfy = fmaf((float)((fy >> sh) & 255u), Ay, Bfy);
      const float tfz = fmaf((float)((fz >> sh) & 255u), Az, Bfz);
      const float tn = fmaxf(fmaxf(fmaxf(tnx, tny), tnz), tmin);
      const float tf = fminf(fminf(fminf(tfx, tfy), tfz), tc);
      const bool h = tn <= tf && r[k] != kEmptyRef;
      t[k] = h ? tn : INFINITY;
      r[k] = h ? r[k] : kEmptyRef;
    }
    // nearest first: a sorting network of selects over the (t, ref) pairs
    auto cx = [&](int i, int j) {
      const bool sw = t[j] < t[i];
      const float ti = sw ? t[j] : t[i], tj = sw ? t[i] : t[j];
      const int ri = sw ? r[j] : r[i], rj = sw ? r[i] : r[j];
      t[i] = ti;
      t[j] = tj;
      r[i] = ri;
      r[j] = rj;
    };
    cx(0, 1);
    cx(2, 3);
    cx(0, 2);
    cx(1, 3);
    cx(1, 2);
    int hold = kEmptyRef;
    float th = 0.0f;
#ifdef __clang__
#pragma unroll
#endif
    for (int k = 3; k >= 0; --k) {
      if (r[k] != kEmptyRef) {
        if (hold != kEmptyRef) stk.push(hold, th);
        hold = r[k];
        th = t[k];
      }
    }
    if (hold != kEmptyRef)
      cur = hold;
    else
      pop_level(stk);
  }

  // One binary node (both child boxes in one WNode).  FAST: IEEE min / max
  // slab (finite 1/d); otherwise the reference's select order.
  template <bool FAST>
  AR_HD void visit_binary(const SceneView& sv, Stack& stk) {
    AR_COUNT_NODE();
    const WNode* n = sv.nodes + cur;
    float4 a = n->a, b = n->b, c = n->c;
    int4 ch = n->d;
    float tl, tr;
    bool hl, hr;
    const float tc = tcull();
    if (FAST) {
      hl = slab_fast(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tmin, tc, &tl);
      hr = slab_fast(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tmin, tc, &tr);
    } else {
      hl = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, tmin, tc, &tl);
      hr = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, tmin, tc, &tr);
    }
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

  AR_HD void visit(const SceneView& sv, Stack& stk) {
    if (NEAR && wide)
      visit_wide(sv, stk);
    else if (fast)
      visit_binary<true>(sv, stk);
    else
      visit_binary<false>(sv, stk);
  }

  // Node phase.  A lane visits internal nodes and parks the triangle leaves
  // it reaches (up to SLOTS, 3 or 4, in the order reached) instead of
  // testing them at once, so the lanes of a wave keep doing node work
  // together; the leaf phase then tests the parked leaves.  A lane can act
  // while it holds a node, or a triangle leaf and a free park slot; a lane
  // that has parked nothing yet is "searching".  A step that parks a leaf
  // also visits the node popped after it.  Every lane stays in the
  // loop until go(can, searching) -- wave-uniform on the device, see
  // walk_items -- ends it for all of them.  Alone, a lane stops at its first
  // parked leaf or when it cannot act.  The visiting order then depends on
  // the wave, and the answer does not (kCullSlack).
  //
  // The loop has one back-edge and one (wave-uniform) exit, and a lane that
  // cannot act stays in it, masked: a `continue` for such lanes makes the
  // compiler split the loop into an inner loop that lanes leave one by one,
  // after which lane masks computed in the loop (e.g. park0 != kLevelEnd)
  // only keep the lanes of its last iteration.
  template <uint32_t SLOTS = 3, class Go>
  AR_HD void descend(const SceneView& sv, Stack& stk, Go go) {
    static_assert(SLOTS == 3 || SLOTS == 4, "three or four park slots");
    for (;;) {
      const bool node = cur >= 0;
      const int last_slot = SLOTS == 4 ? park3 : park2;
      const bool can = last_slot == kLevelEnd && (node || (cur != kLevelEnd && in_blas));
      const bool searching = can && park0 == kLevelEnd;
      if (!go(can, searching)) return;
      if (can) {
        if (!node) {
          if (park0 == kLevelEnd)
            park0 = cur;
          else if (park1 == kLevelEnd)
            park1 = cur;
          else if (SLOTS == 3 || park2 == kLevelEnd)
            park2 = cur;
          else
            park3 = cur;
          pop_level(stk);
        }
        // A lane that has just parked a leaf visits the node it popped in
        // the same iteration (while a slot is still free), so parking does
        // not cost it an iteration of node work.
        const int free_slot = SLOTS == 4 ? park3 : park2;
        if (cur >= 0 && free_slot == kLevelEnd) visit(sv, stk);
      }
    }
  }
  AR_HD void descend(const SceneView& sv, Stack& stk) {
    descend(sv, stk, [](bool can, bool searching) { return searching; });
  }

  // The instance records of the current TLAS form: the 4-wide TLAS's copy
  // ends every single-instance slot's record with kLastBit.
  AR_HD const Inst* tlas_insts(const SceneView& sv) const { return (NEAR && wide) ? sv.qinsts : sv.insts; }

  // Enters the first instance from record k on (same TLAS leaf) whose BLAS
  // root box the local ray hits (a single-instance slot's ref carries
  // kSingleTri; its record is the last of its leaf).  False when the leaf
  // has none left.  Called at the TLAS level.
  AR_HD bool enter(const SceneView& sv, uint32_t k, const Stack& stk) {
    const Inst* ib = tlas_insts(sv);
    k &= ~kSingleTri;
    for (;;) {
      AR_COUNT_INST();
      const Inst* in = ib + k;
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
          fast = finite_inv(li);
          wide = NEAR && fast && wide_ray(sv, lo, li);
          cur = wide ? in->qroot : in->root;
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

  // One triangle: record k of the level's tree form (4-wide trees: qtris,
  // binary: tris).  Hits and near-first ties use the record's key (e1.w),
  // its index in tris.  Returns true when an any-hit query is answered.
  AR_HD bool tri(const SceneView& sv, Stack& stk, uint32_t k, uint32_t& tag) {
    AR_COUNT_TRI();
    const Tri* tp = (NEAR && wide ? sv.qtris : sv.tris) + k;
    float4 v0 = tp->v0, e1 = tp->e1, e2 = tp->e2;
    tag = bits_u(v0.w);
    if (NEAR) k = bits_u(e1.w);
    float th, uh, vh;
    if (tri_hit(mk(v0.x, v0.y, v0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), o, d, tmin, tcull(), &th, &uh,
                &vh)) {
      if (ANY) {
        if (th / s < bt) {
          bpos = k;
          binst = ik;
          stk.sp = 0;
          return true;
        }
      } else if (NEAR) {
        // world-space comparison right away, ties by (instance, leaf position)
        float tw = th / s;
        if (tw < bt || (bpos != kNoHit && tw == bt && (ik < binst || (ik == binst && k < bpos)))) {
          bt = tw;
          bpos = k;
          binst = ik;
          bu = uh;
          bv = vh;
          tmax = bt * s;
        }
      } else if (th < tmax) {
        tmax = th;
        lpos = k;
        lu = uh;
        lv = vh;
      }
    }
    return false;
  }

  // All triangles of one leaf, in leaf order (one for a single-triangle
  // ref of the 4-wide trees).
  AR_HD bool tri_leaf(const SceneView& sv, Stack& stk, int ref) {
    const uint32_t r = (uint32_t)~ref;
    const uint32_t last = (r & kSingleTri) ? 0u : kLastBit;  // tag bit that ends the leaf
    for (uint32_t k = r & ~kSingleTri;; ++k) {
      uint32_t tag;
      if (tri(sv, stk, k, tag)) return true;
      if ((tag & last) || !last) return false;
    }
  }

  // The parked leaves' triangles as one loop (park0's, then park1's, ...),
  // so a lane moves on to its next leaf without waiting for the wave.
  AR_HD bool tri_parked(const SceneView& sv, Stack& stk) {
    uint32_t k = (uint32_t)~park0;
    park0 = kLevelEnd;
    for (;;) {
      const bool single = (k & kSingleTri) != 0u;
      k &= ~kSingleTri;
      uint32_t tag;
      if (tri(sv, stk, k, tag)) {
        park1 = park2 = park3 = kLevelEnd;
        return true;
      }
      if (!single && !(tag & kLastBit)) {
        ++k;
      } else if (park1 != kLevelEnd) {
        k = (uint32_t)~park1;
        park1 = park2;
        park2 = park3;
        park3 = kLevelEnd;
      } else {
        return false;
      }
    }
  }

  // Leaf phase: the parked leaves, then what the lane holds -- a node (the
  // wave ended the node phase first: nothing to do), a triangle leaf (tested
  // here only if nothing was parked; otherwise the next node phase parks
  // it), a TLAS leaf (enter its instances) or kLevelEnd (instance finished,
  // or the query is done).  Returns true when the query is done.
  AR_HD bool leaf(const SceneView& sv, Stack& stk) {
    const bool had = park0 != kLevelEnd;
    if (had && tri_parked(sv, stk)) return true;
    if (cur >= 0) return false;
    if (cur != kLevelEnd) {
      if (in_blas) {
        if (had) return false;
        if (tri_leaf(sv, stk, cur)) return true;
      } else if (enter(sv, (uint32_t)~cur, stk)) {
        return false;
      }
      pop_level(stk);
    }
    while (cur == kLevelEnd) {
      if (!in_blas) return true;  // TLAS exhausted: query done
      // instance finished: world acceptance of its local best (reference
      // order; near-first order has compared every hit in world space)
      if (!ANY && !NEAR && lpos != kNoHit) {
        float tw = tmax / s;
        if (tw < bt) {
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
      fast = finite_inv(inv);
      wide = NEAR && fast && wide_ray(sv, o, inv);
      tmin = wtmin;
      tmax = bt;
      if (!(tlas_insts(sv)[ik].id & kLastBit) && enter(sv, ik + 1, stk)) return false;
      pop_level(stk);
    }
    return false;
  }
};

// One query to completion.  Closest: fills best (best.pos == kNoHit on a
// miss) and returns whether anything was hit.  Any: returns whether
// anything is hit in [tmin, tmax).
template <bool ANY, bool NEAR>
AR_HD bool trace_scene(const SceneView& sv, const Ray& r, WorldHit& best, Stack& stk) {
  Walker<ANY, NEAR> w;
  bool done = w.start(sv, r, stk);
  while (!done) {
    w.descend(sv, stk);
    done = w.leaf(sv, stk);
  }
  best.t = w.bt;
  best.pos = w.bpos;
  best.inst = w.binst;
  best.u = w.bu;
  best.v = w.bv;
  return w.bpos != kNoHit;
}

// The sequence of chunks of 64 items one wave walks in walk_items (below),
// kept apart from the wave intrinsics so that the host emulation of the
// walker runs this same bookkeeping.  Wave `wv` of `nw` first walks its
// static share, chunks wv, wv + nw, ..., wv + (ns - 1) nw; chunk k >= ns of
// its sequence is ns * nw + take(), take() handing out 0, 1, 2, ... across
// the launch (a counter).  ns is half of the chunks per wave when the
// launch has counters and at least 4 chunks per wave; otherwise every chunk
// is static (take() is never called).  `off` is the next unassigned item of
// the current chunk `ca`, sequence index `ka`.  All values are wave-uniform.
struct ChunkCursor {
  uint32_t wv, nw, ns;
  uint32_t ka, ca, off;
  bool dyn;

  AR_HD static bool uses_counters(uint64_t n, uint32_t nw, bool counters) {
    return counters && (uint32_t)((n + 63) >> 6) >= 4u * nw;
  }
  template <class Take>
  AR_HD uint32_t chunk_at(uint32_t k, Take& take) const {
    return k < ns ? wv + k * nw : ns * nw + take();
  }
  template <class Take>
  AR_HD void init(uint32_t wave, uint32_t waves, uint64_t n, bool counters, Take& take) {
    wv = wave;
    nw = waves;
    dyn = uses_counters(n, waves, counters);
    ns = dyn ? ((uint32_t)((n + 63) >> 6) / nw) >> 1 : 0xffffffffu;
    ka = 0;
    off = 0;
    ca = chunk_at(0, take);
  }
  // Whether items remain for this wave.
  AR_HD bool more(uint64_t n) const { return (uint64_t)ca * 64 + off < n; }
  // Start of a refill of `nidle` lanes: a refill that reaches past the
  // current chunk takes the next chunk now (its lanes use it at once), so no
  // chunk is taken without being used.  Returns that chunk (0 if none).
  template <class Take>
  AR_HD uint32_t refill_begin(uint32_t nidle, Take& take) const {
    return off + nidle > 64 ? chunk_at(ka + 1, take) : 0u;
  }
  // The item of refill position p = off + (rank of the lane among the idle
  // ones); cb is refill_begin's chunk.
  AR_HD uint64_t item(uint32_t p, uint32_t cb) const { return (uint64_t)(p < 64 ? ca : cb) * 64 + (p & 63); }
  template <class Take>
  AR_HD void refill_end(uint32_t nidle, uint32_t cb, Take& take) {
    off += nidle;
    if (off >= 64) {
      off -= 64;
      ++ka;
      ca = off > 0 ? cb : chunk_at(ka, take);
    }
  }
};

#ifdef __HIP__  // wave intrinsics: HIP translation units only
// Node-phase rule of a wave whose lanes walk together: the wave keeps
// walking nodes while more than search_min of its active lanes have parked
// nothing, or while no active lane has leaf-phase work.  Wave-uniform.
__device__ __forceinline__ bool wave_quorum(uint32_t search_min, bool searching) {
  const uint64_t sm = __ballot(searching);
  return (uint32_t)__popcll(sm) > search_min || (sm != 0ull && __ballot(!searching) == 0ull);
}

// trace_scene for kernels whose lanes each carry one query to completion
// (no refill): the active lanes run their node phases together under
// wave_quorum, as walk_items does, instead of each stopping at its first
// parked leaf.  Same answer as trace_scene.
template <bool ANY, bool NEAR, uint32_t SLOTS = 3>
__device__ __forceinline__ bool trace_scene_wave(const SceneView& sv, const Ray& r, WorldHit& best, Stack& stk,
                                                 uint32_t search_min = 8) {
  Walker<ANY, NEAR> w;
  bool done = w.start(sv, r, stk);
  while (!done) {
    w.template descend<SLOTS>(sv, stk, [search_min](bool, bool searching) {
      return wave_quorum(search_min, searching);
    });
    done = w.leaf(sv, stk);
  }
  best.t = w.bt;
  best.pos = w.bpos;
  best.inst = w.binst;
  best.u = w.bu;
  best.v = w.bv;
  return w.bpos != kNoHit;
}

// Items [0, n) are split in chunks of 64, and each wave walks a sequence of
// chunks.  Wave `wave` of `waves` starts with a static share, chunks wave,
// wave + waves, ..., wave + (ns - 1) waves, where ns is half of the chunks
// per wave.  The chunks from ns * waves on are then handed out one at a
// time, in order, from a counter that all waves of the launch share: a wave
// whose rays were cheap takes more of them, so the waves end together.  Ray
// costs vary widely (misses next to deep rays), and a fixed stride can alias
// with the image layout: with static shares alone, a closest-hit launch over
// 2^24 rays had its waves end anywhere between 41% and 100% of its
// duration.  Same-address atomics serialise at about 10 ns each, so the
// counter is used only from 4 chunks per wave on; a smaller launch (few or
// cheap rays: a 36-triangle scene at 2.7 chunks per wave ran 40% slower
// with it) keeps every chunk static and touches no counter.
//
// `work` holds two counters that are zero when the launch starts: [0] the
// chunks taken beyond the static shares, [1] the waves that have finished.
// The last wave to finish sets both back to zero, so launches that follow
// on the same stream, and replays of a captured graph, find them zero
// again.  Launches that share a `work` must not overlap.  Every value taken
// from [0] is used before the wave counts itself in [1], so each taking
// atomic has been performed when the last wave clears the counters.  Every
// wave of the grid must call walk_items exactly once with the same n,
// waves and work.  With work == nullptr every chunk is static.
//
// A lane that finishes takes the next unassigned item of its wave; idle
// lanes are refilled together once at least `refill_min` (1..64) of them
// wait, so items fetched together are neighbours (coherent rays stay
// together).  Every item is started exactly once, and the loop ends when the
// wave's chunks are exhausted and every lane is idle.  In its node phase a
// lane parks up to SLOTS (3 or 4) triangle leaves.  n < 2^32.
//   src(i)      -> Ray for item i
//   sink(i, w)  consumes the finished walker of item i
template <bool ANY, bool NEAR, uint32_t SLOTS, class Src, class Sink>
__device__ __forceinline__ void walk_items(const SceneView& sv, Stack& stk, uint64_t n, uint64_t wave,
                                           uint64_t waves, uint32_t* work, uint32_t refill_min,
                                           uint32_t search_min, Src src, Sink sink) {
  const uint32_t lane = __lane_id();
  const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
  // Chunk ids and positions are wave-uniform and below 2^32 (n < 2^32):
  // kept in scalar registers.
  auto take = [&]() -> uint32_t {
    const uint64_t act = __ballot(1);
    const int l0 = __ffsll((unsigned long long)act) - 1;
    uint32_t v = 0;
    if ((int)lane == l0) v = atomicAdd(&work[0], 1u);
    return __builtin_amdgcn_readfirstlane(__shfl(v, l0));
  };
  ChunkCursor cc;
  cc.init(__builtin_amdgcn_readfirstlane((uint32_t)wave), __builtin_amdgcn_readfirstlane((uint32_t)waves), n,
          work != nullptr, take);
  uint32_t item = 0;
  bool alive = false;
  Walker<ANY, NEAR> w;
  for (;;) {
    uint64_t idle = __ballot(!alive);
    uint32_t nidle = (uint32_t)__popcll(idle);
    bool more = cc.more(n);
    if (!more && nidle == 64) break;
    if (more && nidle >= refill_min) {
      const uint32_t cb = cc.refill_begin(nidle, take);
      if (!alive) {
        const uint64_t it = cc.item(cc.off + (uint32_t)__popcll(idle & below), cb);
        if (it < n) {
          item = (uint32_t)it;
          alive = !w.start(sv, src(it), stk);
          if (!alive) sink(it, w);
        }
      }
      cc.refill_end(nidle, cb, take);
    }
    if (alive) {
      w.template descend<SLOTS>(sv, stk, [search_min](bool, bool searching) {
        return wave_quorum(search_min, searching);
      });
      if (w.leaf(sv, stk)) {
        sink(item, w);
        alive = false;
      }
    }
  }
  if (cc.dyn) {
    // The wave is done; the last wave of the launch clears the counters.
    const uint64_t act = __ballot(1);
    const int l0 = __ffsll((unsigned long long)act) - 1;
    if ((int)lane == l0 && atomicAdd(&work[1], 1u) == cc.nw - 1u) {
      atomicExch(&work[0], 0u);
      atomicExch(&work[1], 0u);
    }
  }
}

#endif  // __HIP__

}  // namespace AR_TRAVERSE_NS
