// Host SAH builder.  Produces the same binary tree as the reference's
// full-sweep builder (agent-ray src/accel/bvh.rs:71-236): for every node the
// candidate splits, their SAH costs (same fp32 expression), the strict `<`
// winner (axis order 0,1,2, split position ascending), the leaf rule
// (n <= 4, parent area < f32::EPSILON, or no split cheaper than n) and the
// leaf primitive order (the parent's winning-axis order) are identical.
// Ties in the centroid sort are broken by primitive index.
//
// Instead of re-sorting each node on three axes (O(n log^2 n)), each axis
// list is sorted once and stably partitioned at every split, which keeps each
// list sorted: O(n) work per tree level.
#include <algorithm>
#include <array>
#include <cmath>
#include <future>
#include <numeric>

#include "ar_internal.h"

namespace ar {
namespace {

inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }

inline void merge_box(HBox& a, const HBox& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = sse_min(a.lo[k], b.lo[k]);
    a.hi[k] = sse_max(a.hi[k], b.hi[k]);
  }
}
inline HBox empty_box() {
  HBox b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
  return b;
}
// geometry.rs:54-57
inline float surface_area(const HBox& b) {
  float dx = sse_max(b.hi[0] - b.lo[0], 0.0f), dy = sse_max(b.hi[1] - b.lo[1], 0.0f),
        dz = sse_max(b.hi[2] - b.lo[2], 0.0f);
  return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

struct Build {
  const std::vector<HBox>& boxes;
  std::vector<float> cen[3];
  std::vector<uint32_t> ord[3];   // per-axis sorted primitive ids, partitioned in place
  std::vector<uint32_t> tmp;      // partition scratch, indexed like ord
  std::vector<float> suffix_sa;   // scratch, indexed like ord
  std::vector<uint8_t> go_left;   // per primitive
  std::vector<uint32_t> natural;  // identity order (the root's arrival order)

  explicit Build(const std::vector<HBox>& b) : boxes(b) {}

  struct Sub {
    std::vector<BinNode> nodes;
    std::vector<uint32_t> prims;
    uint32_t depth = 1;
  };

  // Node over ord[*][b, e); `arrive` holds the arrival order of this node's
  // primitives (root: natural order; otherwise the parent's winning axis).
  Sub build(uint32_t b, uint32_t e, const uint32_t* arrive, int level) {
    uint32_t n = e - b;
    HBox box = empty_box();
    for (uint32_t i = b; i < e; ++i) merge_box(box, boxes[ord[0][i]]);
    int best_axis = -1;
    uint32_t best_k = 0;
    if (n > 4) {
      float parent_sa = surface_area(box);
      if (!(parent_sa < 1.1920929e-07f)) {
        const float C_TRAV = 1.0f;
        float best_cost = (float)n;
        for (int axis = 0; axis < 3; ++axis) {
          const uint32_t* o = ord[axis].data();
          HBox acc = empty_box();
          for (uint32_t i = e; i-- > b;) {
            merge_box(acc, boxes[o[i]]);
            suffix_sa[i] = surface_area(acc);
          }
          HBox pre = empty_box();
          for (uint32_t k = 1; k < n; ++k) {
            merge_box(pre, boxes[o[b + k - 1]]);
            float cost = C_TRAV + (surface_area(pre) * (float)k + suffix_sa[b + k] * (float)(n - k)) / parent_sa;
            if (cost < best_cost) {
              best_cost = cost;
              best_axis = axis;
              best_k = k;
            }
          }
        }
      }
    }
    if (best_axis < 0) {
      Sub leaf;
      BinNode nd;
      nd.box = box;
      nd.left_or_start = 0;
      nd.right_or_end = 0;
      nd.count = n;
      leaf.nodes.push_back(nd);
      leaf.prims.assign(arrive, arrive + n);
      return leaf;
    }
    // Partition the two other axis lists stably by side.
    uint32_t m = b + best_k;
    for (uint32_t i = b; i < e; ++i) go_left[ord[best_axis][i]] = i < m ? 1 : 0;
    for (int axis = 0; axis < 3; ++axis) {
      if (axis == best_axis) continue;
      uint32_t* o = ord[axis].data();
      uint32_t l = b, r = m;
      for (uint32_t i = b; i < e; ++i) {
        uint32_t p = o[i];
        if (go_left[p]) tmp[l++] = p; else tmp[r++] = p;
      }
      std::copy(tmp.begin() + b, tmp.begin() + e, o + b);
    }
    const uint32_t* ab = ord[best_axis].data();
    Sub L, R;
    if (level < 3 && n > 8192) {
      auto fut = std::async(std::launch::async, [&]() { return build(m, e, ab + m, level + 1); });
      L = build(b, m, ab + b, level + 1);
      R = fut.get();
    } else {
      L = build(b, m, ab + b, level + 1);
      R = build(m, e, ab + m, level + 1);
    }
    // merge (bvh.rs:157-193): root, left subtree, right subtree.
    Sub out;
    uint32_t ln = (uint32_t)L.nodes.size(), lp = (uint32_t)L.prims.size();
    out.nodes.reserve(1 + L.nodes.size() + R.nodes.size());
    BinNode root;
    root.box = box;
    root.left_or_start = 1;
    root.right_or_end = 1 + ln;
    root.count = 0;
    out.nodes.push_back(root);
    for (auto nd : L.nodes) {
      if (nd.count == 0) {
        nd.left_or_start += 1;
        nd.right_or_end += 1;
      }
      out.nodes.push_back(nd);
    }
    for (auto nd : R.nodes) {
      if (nd.count == 0) {
        nd.left_or_start += 1 + ln;
        nd.right_or_end += 1 + ln;
      } else {
        nd.left_or_start += lp;
      }
      out.nodes.push_back(nd);
    }
    // Left-subtree leaf starts index L.prims, which keep their positions.
    out.prims = std::move(L.prims);
    out.prims.insert(out.prims.end(), R.prims.begin(), R.prims.end());
    out.depth = 1 + std::max(L.depth, R.depth);
    return out;
  }
};

}  // namespace

BinTree build_sah(const std::vector<HBox>& boxes, int nthreads) {
  (void)nthreads;
  BinTree t;
  uint32_t n = (uint32_t)boxes.size();
  if (n == 0) return t;
  Build bd(boxes);
  for (int a = 0; a < 3; ++a) {
    bd.cen[a].resize(n);
    for (uint32_t i = 0; i < n; ++i) bd.cen[a][i] = (boxes[i].lo[a] + boxes[i].hi[a]) * 0.5f;
    bd.ord[a].resize(n);
    std::iota(bd.ord[a].begin(), bd.ord[a].end(), 0u);
    const std::vector<float>& c = bd.cen[a];
    std::sort(bd.ord[a].begin(), bd.ord[a].end(), [&](uint32_t x, uint32_t y) {
      if (c[x] < c[y]) return true;
      if (c[y] < c[x]) return false;
      return x < y;
    });
  }
  bd.tmp.resize(n);
  bd.suffix_sa.resize(n);
  bd.go_left.resize(n);
  bd.natural.resize(n);
  std::iota(bd.natural.begin(), bd.natural.end(), 0u);
  Build::Sub s = bd.build(0, n, bd.natural.data(), 0);
  t.nodes = std::move(s.nodes);
  t.prims = std::move(s.prims);
  t.depth = s.depth;
  return t;
}

// ---- spatial-split SAH (near-first trees only) ------------------------------
// The object-split SAH puts every triangle in exactly one leaf, so a large or
// slanted triangle inflates the box of every node above it.  Here a node may
// instead be split by a plane: the triangles crossing it are referenced from
// both sides, each with the box of its part on that side (Stich, Friedrich
// and Dietrich, "Spatial Splits in Bounding Volume Hierarchies", HPG 2009).
// Per node the cheaper of
//   * the best object split (full sweep over the references' box centroids
//     up to 4096 references, 256 centroid bins above), and
//   * the best of 32 planes per axis, tried when the object split's children
//     overlap by more than 1e-5 of the root's area,
// is taken when it is cheaper than a leaf, with the SAH constants and leaf
// rule of the reference's builder (C_trav = C_isect = 1; n <= 4, a node area
// below f32::EPSILON, or no cheaper split makes a leaf).  A reference that
// straddles the chosen plane goes to one side whole when that is cheaper
// ("reference unsplitting").
//
// Boxes stay conservative: clip points are computed in double and rounded
// outward, and every reference box is widened by 2^-16 of its triangle's
// extent.  The device tests the triangle (v0, v0 + fl(v1 - v0), v0 +
// fl(v2 - v0)), within 2^-24 of its extent of the exact one, so a point it
// can report as a hit lies inside the box of a reference whose region holds
// it, however small that region is against the triangle.
namespace {

struct SRef {
  uint32_t tri;
  HBox box;
};

inline bool box_ok(const HBox& b) {
  return b.lo[0] <= b.hi[0] && b.lo[1] <= b.hi[1] && b.lo[2] <= b.hi[2];
}
inline float down_f(double d) {
  float f = (float)d;
  return (double)f > d ? std::nextafter(f, -INFINITY) : f;
}
inline float up_f(double d) {
  float f = (float)d;
  return (double)f < d ? std::nextafter(f, INFINITY) : f;
}
inline float overlap_area(const HBox& a, const HBox& b) {
  HBox c;
  for (int k = 0; k < 3; ++k) {
    c.lo[k] = sse_max(a.lo[k], b.lo[k]);
    c.hi[k] = sse_min(a.hi[k], b.hi[k]);
  }
  return box_ok(c) ? surface_area(c) : 0.0f;
}

struct SplitBuild {
  // per triangle: its three vertices, contiguous (the clip loop reads them
  // once per reference and axis instead of through the index array)
  std::vector<std::array<float, 9>> tv;
  std::vector<float> margin;  // per triangle: 2^-16 of its extent
  float min_overlap = 0.0f;   // area of overlap that makes spatial splits worth trying

  static constexpr uint32_t kSweepMax = 4096, kObjBins = 256, kSpatialBins = 32;

  struct Sub {
    std::vector<BinNode> nodes;
    std::vector<SRef> refs;
    uint32_t depth = 1;
  };

  void load(uint32_t tri, double v[3][3]) const {
    const std::array<float, 9>& t = tv[tri];
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) v[c][k] = t[3 * c + k];
  }
  // Parts of reference r's triangle (vertices v) inside r.box on each side
  // of the plane x_axis = p; a side the part does not reach comes back empty
  // (!box_ok).
  void clip(const SRef& r, int axis, float p, HBox& L, HBox& R) const {
    double v[3][3];
    load(r.tri, v);
    clip(v, r, axis, p, L, R);
  }
  void clip(const double v[3][3], const SRef& r, int axis, float p, HBox& L, HBox& R) const {
    double llo[3], lhi[3], rlo[3], rhi[3];
    for (int k = 0; k < 3; ++k) {
      llo[k] = rlo[k] = INFINITY;
      lhi[k] = rhi[k] = -INFINITY;
    }
    auto add = [](double* lo, double* hi, const double* q) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], q[k]);
        hi[k] = std::max(hi[k], q[k]);
      }
    };
    const double pd = p;
    for (int e = 0; e < 3; ++e) {
      const double* a = v[e];
      const double* b = v[(e + 1) % 3];
      if (a[axis] <= pd) add(llo, lhi, a);
      if (a[axis] >= pd) add(rlo, rhi, a);
      if ((a[axis] < pd && b[axis] > pd) || (a[axis] > pd && b[axis] < pd)) {
        const double t = (pd - a[axis]) / (b[axis] - a[axis]);
        double q[3];
        for (int k = 0; k < 3; ++k) q[k] = a[k] + t * (b[k] - a[k]);
        q[axis] = pd;
        add(llo, lhi, q);
        add(rlo, rhi, q);
      }
    }
    const float m = margin[r.tri];
    for (int k = 0; k < 3; ++k) {
      L.lo[k] = sse_max(down_f(llo[k]) - m, r.box.lo[k]);
      L.hi[k] = sse_min(up_f(lhi[k]) + m, r.box.hi[k]);
      R.lo[k] = sse_max(down_f(rlo[k]) - m, r.box.lo[k]);
      R.hi[k] = sse_min(up_f(rhi[k]) + m, r.box.hi[k]);
    }
  }

  static float centroid(const SRef& r, int a) { return (r.box.lo[a] + r.box.hi[a]) * 0.5f; }

  struct ObjSplit {
    float cost = INFINITY;
    int axis = -1;
    uint32_t k = 0;       // sweep: first k of the axis order go left
    float cmin = 0, cinv = 0;  // binned: bin of a centroid
    uint32_t bin = 0;     // binned: bins < bin go left
    bool binned = false;
    HBox lb, rb;
  };

  uint32_t obj_bin(const ObjSplit& s, float c) const {
    float f = (c - s.cmin) * s.cinv;
    int b = (int)f;
    return (uint32_t)std::min(std::max(b, 0), (int)kObjBins - 1);
  }

  ObjSplit object_split(const std::vector<SRef>& refs, float parent_sa) const {
    ObjSplit best;
    const uint32_t n = (uint32_t)refs.size();
    if (n <= kSweepMax) {
      std::vector<uint32_t> ord(n);
      std::vector<float> suf(n);
      for (int axis = 0; axis < 3; ++axis) {
        std::iota(ord.begin(), ord.end(), 0u);
        std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
          const float cx = centroid(refs[x], axis), cy = centroid(refs[y], axis);
          return cx < cy || (cx == cy && x < y);
        });
        HBox acc = empty_box();
        for (uint32_t i = n; i-- > 0;) {
          merge_box(acc, refs[ord[i]].box);
          suf[i] = surface_area(acc);
        }
        HBox pre = empty_box();
        for (uint32_t k = 1; k < n; ++k) {
          merge_box(pre, refs[ord[k - 1]].box);
          const float cost = 1.0f + (surface_area(pre) * (float)k + suf[k] * (float)(n - k)) / parent_sa;
          if (cost < best.cost) {
            best.cost = cost;
            best.axis = axis;
            best.k = k;
          }
        }
      }
      if (best.axis >= 0) {  // the two children's boxes
        const int axis = best.axis;
        std::iota(ord.begin(), ord.end(), 0u);
        std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
          const float cx = centroid(refs[x], axis), cy = centroid(refs[y], axis);
          return cx < cy || (cx == cy && x < y);
        });
        best.lb = best.rb = empty_box();
        for (uint32_t i = 0; i < n; ++i) merge_box(i < best.k ? best.lb : best.rb, refs[ord[i]].box);
      }
      return best;
    }
    HBox cb = empty_box();
    for (const SRef& r : refs)
      for (int a = 0; a < 3; ++a) {
        const float c = centroid(r, a);
        cb.lo[a] = sse_min(cb.lo[a], c);
        cb.hi[a] = sse_max(cb.hi[a], c);
      }
    for (int axis = 0; axis < 3; ++axis) {
      const float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.0f)) continue;
      ObjSplit s;
      s.binned = true;
      s.axis = axis;
      s.cmin = cb.lo[axis];
      s.cinv = (float)kObjBins / ext;
      HBox bb[kObjBins];
      uint32_t cnt[kObjBins] = {};
      for (auto& b : bb) b = empty_box();
      for (const SRef& r : refs) {
        const uint32_t b = obj_bin(s, centroid(r, axis));
        merge_box(bb[b], r.box);
        ++cnt[b];
      }
      float suf[kObjBins];
      HBox acc = empty_box();
      for (uint32_t i = kObjBins; i-- > 0;) {
        merge_box(acc, bb[i]);
        suf[i] = surface_area(acc);
      }
      HBox pre = empty_box();
      uint32_t nl = 0;
      for (uint32_t b = 1; b < kObjBins; ++b) {
        merge_box(pre, bb[b - 1]);
        nl += cnt[b - 1];
        if (nl == 0 || nl == n) continue;
        const float cost = 1.0f + (surface_area(pre) * (float)nl + suf[b] * (float)(n - nl)) / parent_sa;
        if (cost < best.cost) {
          best = s;
          best.cost = cost;
          best.bin = b;
          best.lb = pre;
          best.rb = empty_box();
          for (uint32_t i = b; i < kObjBins; ++i) merge_box(best.rb, bb[i]);
        }
      }
    }
    return best;
  }

  struct SpatialSplit {
    float cost = INFINITY;
    int axis = -1;
    float plane = 0;
    HBox lb, rb;
    uint32_t nl = 0, nr = 0;
  };

  SpatialSplit spatial_split(const std::vector<SRef>& refs, const HBox& box, float parent_sa) const {
    SpatialSplit best;
    constexpr uint32_t NB = kSpatialBins;
    for (int axis = 0; axis < 3; ++axis) {
      const float lo = box.lo[axis], ext = box.hi[axis] - lo;
      if (!(ext > 0.0f)) continue;
      float plane[NB + 1];
      for (uint32_t i = 0; i <= NB; ++i) plane[i] = lo + ext * ((float)i / (float)NB);
      auto bin_of = [&](float x) {
        int b = (int)((x - lo) * ((float)NB / ext));
        b = std::min(std::max(b, 0), (int)NB - 1);
        return (uint32_t)b;
      };
      HBox bb[NB];
      uint32_t enter[NB] = {}, leave[NB] = {};
      for (auto& b : bb) b = empty_box();
      for (const SRef& r : refs) {
        const uint32_t b0 = bin_of(r.box.lo[axis]), b1 = bin_of(r.box.hi[axis]);
        ++enter[b0];
        ++leave[b1];
        SRef cur = r;
        uint32_t b = b0;
        double v[3][3];
        if (b0 < b1) load(r.tri, v);
        for (; b < b1; ++b) {
          HBox L, R;
          clip(v, cur, axis, plane[b + 1], L, R);
          if (box_ok(L)) merge_box(bb[b], L);
          if (!box_ok(R)) break;
          cur.box = R;
        }
        if (b == b1) merge_box(bb[b1], cur.box);
      }
      HBox suf[NB];
      uint32_t nsuf[NB];
      HBox acc = empty_box();
      uint32_t nacc = 0;
      for (uint32_t i = NB; i-- > 0;) {
        merge_box(acc, bb[i]);
        nacc += leave[i];
        suf[i] = acc;
        nsuf[i] = nacc;
      }
      HBox pre = empty_box();
      uint32_t nl = 0;
      for (uint32_t i = 1; i < NB; ++i) {
        merge_box(pre, bb[i - 1]);
        nl += enter[i - 1];
        const uint32_t nr = nsuf[i];
        if (nl == 0 || nr == 0 || !box_ok(pre) || !box_ok(suf[i])) continue;
        const float cost = 1.0f + (surface_area(pre) * (float)nl + surface_area(suf[i]) * (float)nr) / parent_sa;
        if (cost < best.cost) {
          best.cost = cost;
          best.axis = axis;
          best.plane = plane[i];
          best.lb = pre;
          best.rb = suf[i];
          best.nl = nl;
          best.nr = nr;
        }
      }
    }
    return best;
  }

  Sub leaf(const HBox& box, std::vector<SRef>&& refs) const {
    Sub s;
    BinNode nd;
    nd.box = box;
    nd.left_or_start = 0;
    nd.right_or_end = 0;
    nd.count = (uint32_t)refs.size();
    s.nodes.push_back(nd);
    s.refs = std::move(refs);
    return s;
  }

  // `extra`: references this subtree may add by spatial splits.  What a node
  // leaves unused is divided between its children in proportion to their
  // references, so the tree does not depend on the order of the parallel
  // subtree builds.
  Sub build(std::vector<SRef>&& refs, uint32_t depth, size_t extra) {
    const uint32_t n = (uint32_t)refs.size();
    HBox box = empty_box();
    for (const SRef& r : refs) merge_box(box, r.box);
    const float parent_sa = surface_area(box);
    if (n <= 4 || depth >= 64 || parent_sa < 1.1920929e-07f) return leaf(box, std::move(refs));
    ObjSplit os = object_split(refs, parent_sa);
    SpatialSplit ss;
    if (extra > 0 && (os.axis < 0 || overlap_area(os.lb, os.rb) > min_overlap))
      ss = spatial_split(refs, box, parent_sa);
    if (!(std::min(os.cost, ss.cost) < (float)n)) return leaf(box, std::move(refs));
    std::vector<SRef> L, R;
    L.reserve(n / 2 + 1);
    R.reserve(n / 2 + 1);
    if (ss.cost < os.cost) {
      const int a = ss.axis;
      const float p = ss.plane;
      const float al = surface_area(ss.lb), ar = surface_area(ss.rb);
      const float nl = (float)ss.nl, nr = (float)ss.nr;
      for (const SRef& r : refs) {
        if (r.box.hi[a] <= p) {
          L.push_back(r);
        } else if (r.box.lo[a] >= p) {
          R.push_back(r);
        } else {
          HBox lu = ss.lb, ru = ss.rb;
          merge_box(lu, r.box);
          merge_box(ru, r.box);
          const float c_split = al * nl + ar * nr;
          const float c_left = surface_area(lu) * nl + ar * (nr - 1.0f);
          const float c_right = al * (nl - 1.0f) + surface_area(ru) * nr;
          if (extra > 0 && c_split < c_left && c_split < c_right) {
            HBox lb, rb;
            clip(r, a, p, lb, rb);
            const bool lk = box_ok(lb), rk = box_ok(rb);
            if (lk) L.push_back({r.tri, lb});
            if (rk) R.push_back({r.tri, rb});
            if (!lk && !rk) L.push_back(r);  // not reached: r.box meets the triangle
            if (lk && rk) --extra;
          } else if (c_left <= c_right) {
            L.push_back(r);
          } else {
            R.push_back(r);
          }
        }
      }
    } else if (os.binned) {
      for (const SRef& r : refs) (obj_bin(os, centroid(r, os.axis)) < os.bin ? L : R).push_back(r);
    } else {
      const int axis = os.axis;
      std::vector<uint32_t> ord(n);
      std::iota(ord.begin(), ord.end(), 0u);
      std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
        const float cx = centroid(refs[x], axis), cy = centroid(refs[y], axis);
        return cx < cy || (cx == cy && x < y);
      });
      for (uint32_t i = 0; i < n; ++i) (i < os.k ? L : R).push_back(refs[ord[i]]);
    }
    if (L.empty() || R.empty()) return leaf(box, std::move(refs));
    std::vector<SRef>().swap(refs);
    const size_t extra_l = (size_t)((double)extra * (double)L.size() / (double)(L.size() + R.size()));
    const size_t extra_r = extra - extra_l;
    Sub A, B;
    if (depth < 4 && n > 8192) {
      auto fut = std::async(std::launch::async, [&]() { return build(std::move(R), depth + 1, extra_r); });
      A = build(std::move(L), depth + 1, extra_l);
      B = fut.get();
    } else {
      A = build(std::move(L), depth + 1, extra_l);
      B = build(std::move(R), depth + 1, extra_r);
    }
    Sub out;
    const uint32_t an = (uint32_t)A.nodes.size(), ap = (uint32_t)A.refs.size();
    out.nodes.reserve(1 + A.nodes.size() + B.nodes.size());
    BinNode root;
    root.box = box;
    root.left_or_start = 1;
    root.right_or_end = 1 + an;
    root.count = 0;
    out.nodes.push_back(root);
    for (auto nd : A.nodes) {
      if (nd.count == 0) {
        nd.left_or_start += 1;
        nd.right_or_end += 1;
      }
      out.nodes.push_back(nd);
    }
    for (auto nd : B.nodes) {
      if (nd.count == 0) {
        nd.left_or_start += 1 + an;
        nd.right_or_end += 1 + an;
      } else {
        nd.left_or_start += ap;
      }
      out.nodes.push_back(nd);
    }
    out.refs = std::move(A.refs);
    out.refs.insert(out.refs.end(), B.refs.begin(), B.refs.end());
    out.depth = 1 + std::max(A.depth, B.depth);
    return out;
  }
};

}  // namespace

BinTree build_split_sah(const std::vector<HBox>& boxes, const float* pos, const uint32_t* idx, size_t max_refs,
                        std::vector<uint32_t>& ref_tri, std::vector<HBox>& ref_box) {
  BinTree t;
  ref_tri.clear();
  ref_box.clear();
  const uint32_t n = (uint32_t)boxes.size();
  if (n == 0) return t;
  SplitBuild sb;
  sb.tv.resize(n);
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) sb.tv[i][3 * c + k] = pos[3 * (size_t)idx[3 * (size_t)i + c] + k];
  sb.margin.resize(n);
  std::vector<SRef> refs(n);
  HBox root = empty_box();
  for (uint32_t i = 0; i < n; ++i) {
    const HBox& b = boxes[i];
    const float ext = sse_max(sse_max(b.hi[0] - b.lo[0], b.hi[1] - b.lo[1]), b.hi[2] - b.lo[2]);
    const float m = ext * 0x1p-16f;
    sb.margin[i] = m;
    refs[i].tri = i;
    for (int k = 0; k < 3; ++k) {
      refs[i].box.lo[k] = b.lo[k] - m;
      refs[i].box.hi[k] = b.hi[k] + m;
    }
    merge_box(root, refs[i].box);
  }
  sb.min_overlap = 1e-5f * surface_area(root);
  SplitBuild::Sub s = sb.build(std::move(refs), 0, max_refs > n ? max_refs - n : 0);
  t.nodes = std::move(s.nodes);
  t.depth = s.depth;
  const size_t r = s.refs.size();
  ref_tri.resize(r);
  ref_box.resize(r);
  t.prims.resize(r);
  for (size_t i = 0; i < r; ++i) {
    ref_tri[i] = s.refs[i].tri;
    ref_box[i] = s.refs[i].box;
    t.prims[i] = (uint32_t)i;
  }
  return t;
}

}  // namespace ar

extern "C" ar_status ar_bvh_build_host(const float* boxes, uint32_t n, float* out_boxes, uint32_t* out_meta,
                                       uint32_t* out_prims, uint32_t* out_nodes, uint32_t* out_depth) {
  if ((!boxes && n) || !out_boxes || !out_meta || !out_prims || !out_nodes)
    return ar::fail(AR_ERR_INVALID_ARG, "ar_bvh_build_host: NULL argument");
  std::vector<ar::HBox> b(n);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      b[i].lo[k] = boxes[6 * i + k];
      b[i].hi[k] = boxes[6 * i + 3 + k];
    }
  ar::BinTree t = ar::build_sah(b, 0);
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    for (int k = 0; k < 3; ++k) {
      out_boxes[6 * i + k] = t.nodes[i].box.lo[k];
      out_boxes[6 * i + 3 + k] = t.nodes[i].box.hi[k];
    }
    out_meta[3 * i] = t.nodes[i].left_or_start;
    out_meta[3 * i + 1] = t.nodes[i].right_or_end;
    out_meta[3 * i + 2] = t.nodes[i].count;
  }
  for (size_t i = 0; i < t.prims.size(); ++i) out_prims[i] = t.prims[i];
  *out_nodes = (uint32_t)t.nodes.size();
  if (out_depth) *out_depth = t.depth;
  return AR_OK;
}
