// Test helper (built by tests/test_host_traverse.py with g++, never shipped):
// runs the library's traversal code -- ar_traverse.h, compiled for the host
// -- over host copies of a scene's trees, so the CPU suite can compare it
// with the oracle ray by ray.  Two drivers per query type:
//   * trace_scene, one ray after another (no leaf parking across lanes);
//   * an emulation of walk_items for 64-lane waves with wave-local refill,
//     the library's chunk bookkeeping (ChunkCursor, with its shared counter)
//     and maximal leaf parking (3 or 4 park slots), which checks that every
//     ray is started once and that the visiting order does not change the
//     answer;
//   * host_chunk_schedule: the chunk bookkeeping alone, for many waves
//     advancing in a random interleaving.
#include <cstdio>
#include <vector>

// Built a second time with -DHOST_COUNTS: the walker's counting hooks then
// count node, triangle and instance records read (host_counts).
#ifdef HOST_COUNTS
static unsigned long long g_counts[3];
#define AR_COUNT_NODE() (++g_counts[0])
#define AR_COUNT_TRI() (++g_counts[1])
#define AR_COUNT_INST() (++g_counts[2])
#endif

#include "../../agent-ray_amd/csrc/ar_internal.h"
#include "../../agent-ray_amd/csrc/ar_traverse.h"

using namespace ar;

namespace {

Ray ray_at(const float* rays, uint64_t i) {
  const float* q = rays + 8 * i;
  Ray r;
  r.o = mk(q[0], q[1], q[2]);
  r.tmin = q[3];
  r.d = mk(q[4], q[5], q[6]);
  r.tmax = q[7];
  return r;
}

template <bool ANY, bool NEAR>
void record(const Walker<ANY, NEAR>& w, const SceneView& sv, uint64_t i, float* t, uint32_t* prim, uint32_t* inst) {
  if (ANY) {
    t[i] = w.bpos != kNoHit ? 1.0f : 0.0f;
    prim[i] = inst[i] = 0;
    return;
  }
  t[i] = w.bpos != kNoHit ? w.bt : INFINITY;
  prim[i] = w.bpos != kNoHit ? (bits_u(sv.tris[w.bpos].v0.w) & ~kLastBit) : kNoHit;
  inst[i] = w.bpos != kNoHit ? (sv.insts[w.binst].id & ~kLastBit) : kNoHit;
}

template <bool ANY, bool NEAR>
int run(const SceneView& sv, const float* rays, uint64_t n, int S, int waves, uint32_t slots, float* t0, uint32_t* p0, uint32_t* i0,
        float* t1, uint32_t* p1, uint32_t* i1) {
  std::vector<int2> lds(64 * 64);
  std::vector<unsigned long long> spill(64 * 512);
  for (uint64_t i = 0; i < n; ++i) {
    Stack stk{lds.data(), spill.data(), 1, 1, S, 0};
    Walker<ANY, NEAR> w;
    bool done = w.start(sv, ray_at(rays, i), stk);
    while (!done) {
      w.descend(sv, stk);
      done = w.leaf(sv, stk);
    }
    if (stk.sp != 0) return -1;
    record(w, sv, i, t0, p0, i0);
  }
  std::vector<uint8_t> seen(n, 0);
  uint32_t counter = 0;  // the launch's shared chunk counter (waves run one after another here)
  auto take = [&]() { return counter++; };
  for (int wave = 0; wave < waves; ++wave) {
    Walker<ANY, NEAR> w[64];
    Stack stk[64];
    bool alive[64] = {};
    uint64_t item[64] = {};
    for (int l = 0; l < 64; ++l) stk[l] = Stack{lds.data() + 64 * l, spill.data() + 512 * l, 1, 1, S, 0};
    ChunkCursor cc;
    cc.init((uint32_t)wave, (uint32_t)waves, n, true, take);
    const uint32_t refill_min = 16;
    for (;;) {
      uint64_t idle = 0;
      for (int l = 0; l < 64; ++l)
        if (!alive[l]) idle |= 1ull << l;
      uint32_t nidle = (uint32_t)__builtin_popcountll(idle);
      bool more = cc.more(n);
      if (!more && nidle == 64) break;
      if (more && nidle >= refill_min) {
        const uint32_t cb = cc.refill_begin(nidle, take);
        for (int l = 0; l < 64; ++l) {
          if (alive[l]) continue;
          uint64_t below = l == 0 ? 0ull : ((~0ull) >> (64 - l));
          uint64_t it = cc.item(cc.off + (uint32_t)__builtin_popcountll(idle & below), cb);
          if (it >= n) continue;
          if (seen[it]++) return -2;
          item[l] = it;
          alive[l] = !w[l].start(sv, ray_at(rays, it), stk[l]);
          if (!alive[l]) record(w[l], sv, it, t1, p1, i1);
        }
        cc.refill_end(nidle, cb, take);
      }
      for (int l = 0; l < 64; ++l) {
        if (!alive[l]) continue;
        auto any_time = [](bool can, bool) { return can; };  // park whenever possible
        if (slots == 4)
          w[l].template descend<4>(sv, stk[l], any_time);
        else
          w[l].template descend<3>(sv, stk[l], any_time);
        if (w[l].leaf(sv, stk[l])) {
          if (stk[l].sp != 0) return -3;
          record(w[l], sv, item[l], t1, p1, i1);
          alive[l] = false;
        }
      }
    }
  }
  for (uint64_t i = 0; i < n; ++i)
    if (seen[i] != 1) return -4;
  return 0;
}

}  // namespace

// The chunk bookkeeping of walk_items (ChunkCursor) for `waves` waves of 64
// lanes that advance in a random interleaving (xorshift, `seed`): at each
// step a random unfinished wave refills its idle lanes if it has items left
// and at least refill_min idle lanes, as walk_items does, and otherwise
// retires a random subset of its busy lanes.  A finished wave counts itself,
// and the last one clears the shared counter, as on the device.  Returns 0
// when every item in [0, n) was started exactly once and the counter ends at
// zero; *takes = chunks taken from the counter (0 when the launch keeps
// static shares).
extern "C" int host_chunk_schedule(uint64_t n, uint32_t waves, int counters, uint32_t refill_min, uint64_t seed,
                                   uint64_t* takes) {
  struct Wave {
    ChunkCursor cc;
    uint64_t busy;
    bool done;
  };
  uint32_t counter = 0, finished = 0;
  uint64_t ntake = 0;
  auto take = [&]() {
    ++ntake;
    return counter++;
  };
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
  auto rnd = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  std::vector<Wave> ws(waves);
  for (uint32_t w = 0; w < waves; ++w) {
    ws[w].cc.init(w, waves, n, counters != 0, take);
    ws[w].busy = 0;
    ws[w].done = false;
  }
  std::vector<uint8_t> seen(n, 0);
  uint32_t left = waves;
  while (left) {
    Wave& v = ws[rnd() % waves];
    if (v.done) continue;
    const uint64_t idle = ~v.busy;
    const uint32_t nidle = (uint32_t)__builtin_popcountll(idle);
    const bool more = v.cc.more(n);
    if (!more && nidle == 64) {
      v.done = true;
      --left;
      if (v.cc.dyn && ++finished == waves) counter = finished = 0;
      continue;
    }
    if (more && nidle >= refill_min) {
      const uint32_t cb = v.cc.refill_begin(nidle, take);
      for (int l = 0; l < 64; ++l) {
        if (v.busy >> l & 1) continue;
        const uint64_t below = l == 0 ? 0ull : ((~0ull) >> (64 - l));
        const uint64_t it = v.cc.item(v.cc.off + (uint32_t)__builtin_popcountll(idle & below), cb);
        if (it >= n) continue;
        if (seen[it]++) return -1;
        v.busy |= 1ull << l;
      }
      v.cc.refill_end(nidle, cb, take);
    } else {
      v.busy &= rnd();  // some lanes finish their rays
    }
  }
  for (uint64_t i = 0; i < n; ++i)
    if (seen[i] != 1) return -2;
  if (counter != 0) return -3;
  *takes = ntake;
  return 0;
}

extern "C" int host_traverse(ar_scene* s, const float* rays, uint64_t n, int any, int near, int lds_stack, int slots,
                             float* t0, uint32_t* p0, uint32_t* i0, float* t1, uint32_t* p1, uint32_t* i1) {
  HostAccel h;
  if (build_host_accel(s, h) != AR_OK) return -10;
  SceneView sv{};
  sv.nodes = h.nodes.data();
  sv.qnodes = h.qnodes.data();
  sv.tris = h.tris.data();
  sv.qtris = h.qtris.empty() ? sv.tris : h.qtris.data();
  sv.insts = h.insts.data();
  sv.qinsts = h.qinsts.data();
  sv.tlas_root = h.tlas_root;
  sv.qtlas_root = h.qtlas_root;
  for (int k = 0; k < 6; ++k) sv.tlas_box[k] = h.tlas_box[k];
  sv.n_inst = (uint32_t)h.insts.size();
  sv.wide_ok = h.wide_ok;
  if (slots != 3 && slots != 4) return -11;
  if (any && near) return run<true, true>(sv, rays, n, lds_stack, 3, (uint32_t)slots, t0, p0, i0, t1, p1, i1);
  if (any) return run<true, false>(sv, rays, n, lds_stack, 3, (uint32_t)slots, t0, p0, i0, t1, p1, i1);
  if (near) return run<false, true>(sv, rays, n, lds_stack, 3, (uint32_t)slots, t0, p0, i0, t1, p1, i1);
  return run<false, false>(sv, rays, n, lds_stack, 3, (uint32_t)slots, t0, p0, i0, t1, p1, i1);
}

// sizes[o] = the largest bottom-level leaf of object o (Tri records between
// two leaf-end tags, in the binary trees' leaf order).
extern "C" int host_max_leaf(ar_scene* s, uint32_t* sizes) {
  HostAccel h;
  if (build_host_accel(s, h) != AR_OK) return -10;
  for (size_t o = 0; o < h.objs.size(); ++o) {
    uint32_t run = 0, best = 0;
    for (uint32_t r = 0; r < h.objs[o].n_tris; ++r) {
      ++run;
      if (bits_u(h.tris[h.objs[o].tri_base + r].v0.w) & kLastBit) {
        best = run > best ? run : best;
        run = 0;
      }
    }
    sizes[o] = best;
  }
  return 0;
}

// counts[0] = Tri records, counts[1] = the 4-wide trees' own records (0 when
// they share the binary trees' records).
extern "C" int host_record_counts(ar_scene* s, uint64_t* counts) {
  HostAccel h;
  if (build_host_accel(s, h) != AR_OK) return -10;
  counts[0] = h.tris.size();
  counts[1] = h.qtris.size();
  return 0;
}

#ifdef HOST_COUNTS
// counts[0..2] = node, triangle and instance records read since the last call
// (both drivers of host_traverse count), then reset.
extern "C" void host_counts(unsigned long long* counts) {
  for (int i = 0; i < 3; ++i) {
    counts[i] = g_counts[i];
    g_counts[i] = 0;
  }
}
#endif
