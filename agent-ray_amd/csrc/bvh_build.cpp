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
