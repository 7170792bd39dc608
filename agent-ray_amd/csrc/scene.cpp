// Scene handle: collects Scene::new's inputs (agent-ray src/scene.rs:125-139),
// builds each object's BLAS and the TLAS on the host with the reference's
// SAH (accel/bvh.rs:317-410), flattens them into the wide-node layout and
// uploads everything once.  The committed scene stays resident in HBM.
#include <algorithm>
#include <array>
#include <atomic>
#include <future>
#include <thread>
#include <cmath>

#include "ar_internal.h"

using namespace ar;

namespace {

bool is_identity16(const float* m) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      if (m[4 * c + r] != (c == r ? 1.0f : 0.0f)) return false;
  return true;
}

// 4x4 inverse by cofactors (column-major in and out).  Identity maps to
// itself exactly.
void inverse16(const float* m, float* out) {
  if (is_identity16(m)) {
    for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return;
  }
  float inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  float rd = 1.0f / det;
  for (int i = 0; i < 16; ++i) out[i] = inv[i] * rd;
}

// Mat3::from_mat4(m).inverse().transpose(), column-major 3x3 out.
void normal_matrix(const float* m, float* out) {
  float a[3][3];  // a[col][row]
  bool ident = true;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      a[c][r] = m[4 * c + r];
      if (a[c][r] != (c == r ? 1.0f : 0.0f)) ident = false;
    }
  if (ident) {
    for (int i = 0; i < 9; ++i) out[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    return;
  }
  auto e = [&](int row, int col) { return a[col][row]; };
  float cof[3][3];
  cof[0][0] = e(1, 1) * e(2, 2) - e(1, 2) * e(2, 1);
  cof[0][1] = e(1, 2) * e(2, 0) - e(1, 0) * e(2, 2);
  cof[0][2] = e(1, 0) * e(2, 1) - e(1, 1) * e(2, 0);
  cof[1][0] = e(0, 2) * e(2, 1) - e(0, 1) * e(2, 2);
  cof[1][1] = e(0, 0) * e(2, 2) - e(0, 2) * e(2, 0);
  cof[1][2] = e(0, 1) * e(2, 0) - e(0, 0) * e(2, 1);
  cof[2][0] = e(0, 1) * e(1, 2) - e(0, 2) * e(1, 1);
  cof[2][1] = e(0, 2) * e(1, 0) - e(0, 0) * e(1, 2);
  cof[2][2] = e(0, 0) * e(1, 1) - e(0, 1) * e(1, 0);
  float det = e(0, 0) * cof[0][0] + e(0, 1) * cof[0][1] + e(0, 2) * cof[0][2];
  float id = 1.0f / det;
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) out[3 * col + row] = cof[row][col] * id;
}

// transform_point3a with the reference grouping.
void xform_pt16(const float* m, const float* p, float* o) {
  for (int k = 0; k < 3; ++k) {
    float a = m[k] * p[0];
    a = m[4 + k] * p[1] + a;
    a = m[8 + k] * p[2] + a;
    o[k] = m[12 + k] + a;
  }
}

inline float smin(float a, float b) { return a < b ? a : b; }
inline float smax(float a, float b) { return a > b ? a : b; }

// AABB::transform (geometry.rs:66-84)
HBox transform_box(const HBox& b, const float* m) {
  HBox r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = INFINITY;
    r.hi[k] = -INFINITY;
  }
  for (int c = 0; c < 8; ++c) {
    float p[3] = {(c & 1) ? b.hi[0] : b.lo[0], (c & 2) ? b.hi[1] : b.lo[1], (c & 4) ? b.hi[2] : b.lo[2]};
    float o[3];
    xform_pt16(m, p, o);
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = smin(r.lo[k], o[k]);
      r.hi[k] = smax(r.hi[k], o[k]);
    }
  }
  return r;
}

// Flatten a binary tree: every internal node becomes one wide node holding
// both child boxes.  Returns the root ref; leaf refs are ~(prim_base + start).
struct Flat {
  int root_ref;
  HBox root_box;
};
Flat flatten(const BinTree& t, std::vector<WNode>& out, uint32_t prim_base) {
  Flat f;
  f.root_box = t.nodes[0].box;
  std::vector<int> wide(t.nodes.size(), -1);
  uint32_t base = (uint32_t)out.size();
  uint32_t k = 0;
  for (size_t i = 0; i < t.nodes.size(); ++i)
    if (t.nodes[i].count == 0) wide[i] = (int)(base + k++);
  out.resize(base + k);
  auto ref = [&](uint32_t i) -> int {
    const BinNode& n = t.nodes[i];
    return n.count == 0 ? wide[i] : ~(int)(prim_base + n.left_or_start);
  };
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const BinNode& n = t.nodes[i];
    if (n.count != 0) continue;
    const HBox& L = t.nodes[n.left_or_start].box;
    const HBox& R = t.nodes[n.right_or_end].box;
    WNode w;
    w.a = make_float4(L.lo[0], L.lo[1], L.lo[2], L.hi[0]);
    w.b = make_float4(L.hi[1], L.hi[2], R.lo[0], R.lo[1]);
    w.c = make_float4(R.lo[2], R.hi[0], R.hi[1], R.hi[2]);
    w.d = make_int4(ref(n.left_or_start), ref(n.right_or_end), 0, 0);
    out[wide[i]] = w;
  }
  f.root_ref = ref(0);
  return f;
}

// ---- 4-wide trees for the near-first traversal ------------------------------
static float box_area(const HBox& b) {
  float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Smallest power of two >= x (x > 0, finite).
static float pow2_at_least(float x) {
  int e = 0;
  float m = std::frexp(x, &e);  // x = m * 2^e, m in [0.5, 1)
  return m == 0.5f ? std::ldexp(1.0f, e - 1) : std::ldexp(1.0f, e);
}

// The kernel's dequantization: q * s is exact (s a power of two, q < 256),
// so this single rounding is what the device computes too.
static float deq(float org, uint32_t q, float s) { return org + (float)q * s; }

// Quantizes the child boxes of one 4-wide node relative to the node box,
// rounding each plane outward until the dequantized box contains the exact
// one; an axis whose planes cannot all be covered doubles its scale.
static QNode quantize_node(const HBox& nb, const std::vector<const HBox*>& cb, const int* refs) {
  uint32_t w[16];
  float org[3], scl[3];
  uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    org[k] = nb.lo[k];
    float ext = nb.hi[k] - nb.lo[k];
    float s = ext > 0.0f ? pow2_at_least(ext / 255.0f) : 1.0f;
    for (;;) {
      bool ok = true;
      uint32_t wl = 0, wh = 0;
      for (size_t c = 0; c < cb.size() && ok; ++c) {
        float lo = cb[c]->lo[k], hi = cb[c]->hi[k];
        double fl = std::floor(((double)lo - (double)org[k]) / (double)s);
        double fh = std::ceil(((double)hi - (double)org[k]) / (double)s);
        uint32_t a = (uint32_t)std::min(std::max(fl, 0.0), 255.0);
        uint32_t b = (uint32_t)std::min(std::max(fh, 0.0), 255.0);
        while (a > 0 && deq(org[k], a, s) > lo) --a;
        while (b < 255 && deq(org[k], b, s) < hi) ++b;
        if (deq(org[k], a, s) > lo || deq(org[k], b, s) < hi) ok = false;
        wl |= a << (8 * c);
        wh |= b << (8 * c);
      }
      if (ok) {
        scl[k] = s;
        qlo[k] = wl;
        qhi[k] = wh;
        break;
      }
      s *= 2.0f;
    }
  }
  memcpy(&w[0], org, 12);
  memcpy(&w[3], &scl[0], 4);
  memcpy(&w[4], &scl[1], 8);
  w[6] = qlo[0];
  w[7] = qlo[1];
  w[8] = qlo[2];
  w[9] = qhi[0];
  w[10] = qhi[1];
  w[11] = qhi[2];
  memcpy(&w[12], refs, 16);
  QNode q;
  memcpy(&q, w, sizeof(q));
  return q;
}

// Collapses a binary tree into 4-wide nodes, choosing each node's slots by
// a surface-area cost model solved exactly over the binary tree (dynamic
// programming, bottom up).  A ray is taken to visit a 4-wide node with
// probability proportional to its box's area.  Every triangle sits in a
// slot of its own box, wherever the collapse puts it, so its expected tests
// are the same for every collapse, and the cost to minimise is the total
// area of the 4-wide nodes.  F(m, k) is the least such area in the subtree
// of binary node m placed in at most k slots of its parent:
//   * an internal node in one slot is a 4-wide node: its area plus its
//     children's best split of four slots, min over j of
//     F(left, j) + F(right, 4 - j);
//   * with k >= 2 slots it may instead be opened: min over j of
//     F(left, j) + F(right, k - j);
//   * a leaf of one triangle is a triangle slot (0);
//   * a bottom-level leaf of more than four triangles is a leaf slot (0);
//   * a bottom-level leaf of two to four triangles in one slot is a 4-wide
//     node whose children are its triangles (ref ~(record | kSingleTri)),
//     and with as many slots as triangles its triangles sit directly in the
//     parent's slots instead (0);
//   * a top-level leaf (instances) is a leaf slot (0).
// Against opening the slot of largest area until four slots are filled, on
// a 2^20-triangle soup: 30.7 instead of 32.6 node visits per ray, closest
// hit 5.73 instead of 6.08 ms per 2^24 rays.
// Slots keep the binary tree's left-to-right order.  Leaves keep their
// ranges (ref ~(prim_base + start)), so leaf positions and the Tri records
// are shared with the binary form, and the answers do not depend on the
// collapse (near-first ties are broken by leaf position).  Nodes are stored
// in preorder.
struct QFlat {
  int root_ref;
  uint32_t depth;
};

// Expected SAH cost of a binary tree (C_trav = C_isect = 1) relative to its
// root's area: the node and triangle tests of a random ray through the root.
static double sah_cost(const BinTree& t) {
  const double root = box_area(t.nodes[0].box);
  double c = 0.0;
  for (const BinNode& n : t.nodes) c += (double)box_area(n.box) * (n.count ? (double)n.count : 1.0);
  return root > 0.0 ? c / root : 0.0;
}
struct QPlan {
  const BinTree& t;
  const std::vector<HBox>* prim_boxes;
  // per binary node: F(m, k) for k = 1..4, and the choice that reaches it:
  // 0 keep as one slot, j in 1..3 open with j slots for the left child,
  // 4 the leaf's triangles in slots of their own
  std::vector<std::array<float, 5>> F;
  std::vector<std::array<uint8_t, 5>> choice;
  std::vector<uint8_t> node_split;  // j of the 4-wide node made of m (internal m)

  QPlan(const BinTree& tree, const std::vector<HBox>* pb) : t(tree), prim_boxes(pb) {
    const size_t n = t.nodes.size();
    F.resize(n);
    choice.resize(n);
    node_split.assign(n, 1);
    // preorder: children come after their parent
    for (size_t ii = n; ii-- > 0;) solve((uint32_t)ii);
  }
  void solve(uint32_t m) {
    const BinNode& nd = t.nodes[m];
    const float A = box_area(nd.box);
    auto& f = F[m];
    auto& c = choice[m];
    if (nd.count != 0) {
      // one slot: a node of triangle slots (two to four triangles); a leaf of
      // more (the SAH leaf rule allows them) stays a range of Tri records
      const bool node = prim_boxes && nd.count >= 2 && nd.count <= 4;
      for (int k = 1; k <= 4; ++k) {
        const bool inline_tris = node && (uint32_t)k >= nd.count;
        f[k] = node && !inline_tris ? A : 0.0f;
        c[k] = inline_tris ? 4 : 0;
      }
      return;
    }
    const auto& L = F[nd.left_or_start];
    const auto& R = F[nd.right_or_end];
    float best = INFINITY;
    for (int j = 1; j <= 3; ++j)
      if (L[j] + R[4 - j] < best) {
        best = L[j] + R[4 - j];
        node_split[m] = (uint8_t)j;
      }
    f[1] = A + best;
    c[1] = 0;
    for (int k = 2; k <= 4; ++k) {
      f[k] = f[1];
      c[k] = 0;
      for (int j = 1; j < k; ++j)
        if (L[j] + R[k - j] < f[k]) {
          f[k] = L[j] + R[k - j];
          c[k] = (uint8_t)j;
        }
    }
  }
  // The slots binary node m occupies when given k of them.
  struct Slot {
    uint32_t node;
    int tri;  // >= 0: the leaf's triangle `tri` in a slot of its own
  };
  void slots(uint32_t m, int k, std::vector<Slot>& out) const {
    const BinNode& nd = t.nodes[m];
    const uint8_t ch = choice[m][k];
    if (ch == 0) {
      out.push_back({m, -1});
    } else if (ch == 4) {
      for (uint32_t j = 0; j < nd.count; ++j) out.push_back({m, (int)j});
    } else {
      slots(nd.left_or_start, ch, out);
      slots(nd.right_or_end, k - ch, out);
    }
  }
};

static int qbvh_rec(const QPlan& plan, uint32_t i, uint32_t depth, std::vector<QNode>& out, uint32_t prim_base,
                    QFlat& f) {
  const BinTree& t = plan.t;
  const std::vector<HBox>* prim_boxes = plan.prim_boxes;
  const BinNode& n = t.nodes[i];
  if (n.count == 1 || (n.count != 0 && (!prim_boxes || n.count > 4))) return ~(int)(prim_base + n.left_or_start);
  f.depth = std::max(f.depth, depth + 1);
  const uint32_t me = (uint32_t)out.size();
  out.emplace_back();
  int refs[4] = {kEmptyRef, kEmptyRef, kEmptyRef, kEmptyRef};
  std::vector<const HBox*> boxes;
  if (n.count != 0) {  // a leaf of 2..4 triangles: one node with a slot per triangle
    for (uint32_t j = 0; j < n.count; ++j) {
      refs[j] = ~(int)((prim_base + n.left_or_start + j) | kSingleTri);
      boxes.push_back(&(*prim_boxes)[t.prims[n.left_or_start + j]]);
    }
    out[me] = quantize_node(n.box, boxes, refs);
    return (int)me;
  }
  std::vector<QPlan::Slot> sl;
  const int j = plan.node_split[i];
  plan.slots(n.left_or_start, j, sl);
  plan.slots(n.right_or_end, 4 - j, sl);
  for (size_t k = 0; k < sl.size(); ++k) {
    const BinNode& c = t.nodes[sl[k].node];
    if (sl[k].tri >= 0) {
      refs[k] = ~(int)((prim_base + c.left_or_start + (uint32_t)sl[k].tri) | kSingleTri);
      boxes.push_back(&(*prim_boxes)[t.prims[c.left_or_start + (uint32_t)sl[k].tri]]);
    } else {
      refs[k] = qbvh_rec(plan, sl[k].node, depth + 1, out, prim_base, f);
      boxes.push_back(&c.box);
    }
  }
  out[me] = quantize_node(n.box, boxes, refs);
  return (int)me;
}
static QFlat build_qbvh(const BinTree& t, std::vector<QNode>& out, uint32_t prim_base,
                        const std::vector<HBox>* prim_boxes) {
  QFlat f{0, 0};
  const QPlan plan(t, prim_boxes);
  f.root_ref = qbvh_rec(plan, 0, 0, out, prim_base, f);
  return f;
}

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  size_t n = std::max<size_t>(v.size() * sizeof(T), 16);
  hipError_t e = b.alloc(n);
  if (e != hipSuccess) return e;
  if (!v.empty()) e = hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

}  // namespace

extern "C" {

ar_status ar_scene_create(ar_scene** out) {
  if (!out) return fail(AR_ERR_INVALID_ARG, "ar_scene_create: out is NULL");
  *out = new ar_scene();
  return AR_OK;
}

ar_status ar_scene_destroy(ar_scene* s) {
  if (!s) return AR_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->scratch_event) (void)hipEventDestroy(s->scratch_event);
  destroy_workspace(s->ws);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return AR_OK;
}

ar_status ar_scene_add_material(ar_scene* s, const ar_material* m, uint32_t* out_id) {
  if (!s || !m) return fail(AR_ERR_INVALID_ARG, "ar_scene_add_material: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (m->type > AR_MAT_DIELECTRIC) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown material type");
  DMaterial d{};
  d.type = m->type;
  d.light_id = -1;
  // ConductorBsdf::new / DielectricBsdf::new clamp alpha (microfacet.rs:94-99,171-176)
  d.roughness = std::min(std::max(m->roughness, 1e-3f), 1.0f);
  d.eta = m->eta;
  for (int k = 0; k < 3; ++k) d.albedo[k] = m->albedo[k];
  d.tex = -1;
  if (m->albedo_texture != 0) {
    if (m->type != AR_MAT_LAMBERTIAN) return fail(AR_ERR_UNSUPPORTED_TYPE, "only a Lambertian albedo can be textured");
    if (m->albedo_texture > s->tex_info.size()) return fail(AR_ERR_INVALID_ARG, "unknown texture id");
    d.tex = (int32_t)m->albedo_texture - 1;
  }
  s->materials.push_back(d);
  if (out_id) *out_id = (uint32_t)s->materials.size() - 1;
  return AR_OK;
}

ar_status ar_scene_add_mesh(ar_scene* s, const float* positions, uint32_t nv, const uint32_t* indices,
                            uint32_t nt, const float* normals, const uint32_t* slots, uint32_t n_slots,
                            const uint32_t* mats, uint32_t n_mats, const float* transform,
                            uint32_t* out_object_id) try {
  if (!s || !positions || !indices || !slots || !mats)
    return fail(AR_ERR_INVALID_ARG, "ar_scene_add_mesh: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (nt == 0 || nv == 0) return fail(AR_ERR_INVALID_ARG, "ar_scene_add_mesh: empty mesh");
  if (n_slots != 1 && n_slots != nt)
    return fail(AR_ERR_INVALID_ARG, "material_slots must have length 1 or the triangle count");
  if (n_mats == 0) return fail(AR_ERR_INVALID_ARG, "an object needs at least one material");
  for (uint64_t i = 0; i < 3ull * nt; ++i)
    if (indices[i] >= nv) return fail(AR_ERR_INVALID_ARG, "triangle index out of range");
  for (uint32_t i = 0; i < n_mats; ++i)
    if (mats[i] >= s->materials.size()) return fail(AR_ERR_INVALID_ARG, "material id out of range");
  for (uint64_t i = 0; i < 3ull * nv; ++i)
    if (!std::isfinite(positions[i])) return fail(AR_ERR_INVALID_ARG, "non-finite vertex position");
  HostMesh m;
  m.pos.assign(positions, positions + 3ull * nv);
  if (normals) m.nrm.assign(normals, normals + 3ull * nv);
  m.idx.assign(indices, indices + 3ull * nt);
  m.slots.assign(slots, slots + n_slots);
  m.mats.assign(mats, mats + n_mats);
  if (transform) {
    for (int i = 0; i < 16; ++i) m.xf[i] = transform[i];
  } else {
    for (int i = 0; i < 16; ++i) m.xf[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  }
  s->meshes.push_back(std::move(m));
  if (out_object_id) *out_object_id = (uint32_t)s->meshes.size() - 1;
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_scene_add_mesh");
}

ar_status ar_scene_set_builder(ar_scene* s, uint32_t builder) {
  if (!s) return fail(AR_ERR_INVALID_ARG, "ar_scene_set_builder: NULL scene");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (builder > AR_BUILDER_SAH_SPLIT) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown tree builder");
  s->builder = builder;
  return AR_OK;
}

ar_status ar_scene_add_texture(ar_scene* s, const uint8_t* rgb, uint32_t width, uint32_t height, uint32_t* out_id) try {
  if (!s || !rgb) return fail(AR_ERR_INVALID_ARG, "ar_scene_add_texture: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 28))
    return fail(AR_ERR_INVALID_ARG, "ar_scene_add_texture: empty or oversized image");
  // srgb_to_linear of every 8-bit level (texture.rs:144-151), as
  // ImageTexture::load applies it to p / 255 (texture.rs:80-89)
  static float lut[256];
  static bool lut_ready = false;
  if (!lut_ready) {
    for (int k = 0; k < 256; ++k) {
      const float x = (float)k / 255.0f;
      lut[k] = x <= 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f);
    }
    lut_ready = true;
  }
  const uint64_t n = (uint64_t)width * height;
  const uint32_t first = (uint32_t)s->texels.size();
  if ((uint64_t)first + n > (1ull << 31)) return fail(AR_ERR_INVALID_ARG, "ar_scene_add_texture: too many texels");
  s->texels.reserve(first + n);
  for (uint64_t i = 0; i < n; ++i)
    s->texels.push_back(make_float4(lut[rgb[3 * i]], lut[rgb[3 * i + 1]], lut[rgb[3 * i + 2]], 0.0f));
  s->tex_info.push_back(make_uint4(first, width, height, 0u));
  if (out_id) *out_id = (uint32_t)s->tex_info.size();  // ids start at 1; 0 means no texture
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_scene_add_texture");
}

ar_status ar_scene_set_mesh_texcoords(ar_scene* s, uint32_t object_id, const float* uv, uint32_t nv) try {
  if (!s || !uv) return fail(AR_ERR_INVALID_ARG, "ar_scene_set_mesh_texcoords: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (object_id >= s->meshes.size()) return fail(AR_ERR_INVALID_ARG, "unknown object id");
  HostMesh& m = s->meshes[object_id];
  if ((uint64_t)nv * 3 != m.pos.size()) return fail(AR_ERR_INVALID_ARG, "tex_coords need one (u, v) per vertex");
  for (uint64_t i = 0; i < 2ull * nv; ++i)
    if (!std::isfinite(uv[i])) return fail(AR_ERR_INVALID_ARG, "non-finite texture coordinate");
  m.uv.assign(uv, uv + 2ull * nv);
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_scene_set_mesh_texcoords");
}

ar_status ar_scene_add_light(ar_scene* s, const ar_light* l, uint32_t* out_id) {
  if (!s || !l) return fail(AR_ERR_INVALID_ARG, "ar_scene_add_light: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (l->type > AR_LIGHT_QUAD) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown light type");
  uint32_t id = (uint32_t)s->lights.size();
  if (l->type == AR_LIGHT_QUAD) {
    f3 eu = mk(l->edge_u[0], l->edge_u[1], l->edge_u[2]), ev = mk(l->edge_v[0], l->edge_v[1], l->edge_v[2]);
    if (length(cross(eu, ev)) <= 0.0f) return fail(AR_ERR_INVALID_ARG, "degenerate quad light");
    DMaterial d{};
    d.tex = -1;  // emitters are never textured
    d.type = AR_MAT_LAMBERTIAN;
    d.light_id = (int32_t)id;
    d.roughness = 1.0f;
    d.eta = 1.0f;
    for (int k = 0; k < 3; ++k) d.emission[k] = l->intensity[k];
    s->materials.push_back(d);
    uint32_t mid = (uint32_t)s->materials.size() - 1;
    f3 a = mk(l->position[0], l->position[1], l->position[2]);
    f3 b = a + eu, c = (a + eu) + ev, dd = a + ev;
    float pos[12] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, dd.x, dd.y, dd.z};
    uint32_t idx[6] = {0, 1, 2, 0, 2, 3};
    uint32_t slot = 0;
    ar_status st = ar_scene_add_mesh(s, pos, 4, idx, 2, nullptr, &slot, 1, &mid, 1, nullptr, nullptr);
    if (st != AR_OK) return st;
  }
  s->lights.push_back(*l);
  if (out_id) *out_id = id;
  return AR_OK;
}

ar_status ar_scene_set_light_distribution(ar_scene* s, uint32_t mode) {
  if (!s) return fail(AR_ERR_INVALID_ARG, "NULL scene");
  if (mode > AR_LIGHTDIST_POWER) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown light distribution");
  s->dist_mode = mode;
  return AR_OK;
}

}  // extern "C"

namespace ar {

// Host half of ar_scene_commit: BLAS per object, Tri records in leaf order,
// instance records in TLAS leaf order, flattened node array, per-object
// shading records.  Touches no device.
ar_status build_host_accel(const ar_scene* s, HostAccel& h) {
  std::vector<WNode>& nodes = h.nodes;
  std::vector<Tri>& tris = h.tris;
  std::vector<Inst> inst_by_obj(s->meshes.size());
  std::vector<HBox> world(s->meshes.size());
  std::vector<DObject>& objs = h.objs;
  objs.assign(s->meshes.size(), DObject{});
  std::vector<uint32_t>&gidx = h.gidx, &gslots = h.gslots, &gmats = h.gmats, &tri_obj = h.tri_obj;
  std::vector<float>& gnrm = h.gnrm;
  uint32_t max_depth = 0;
  bool any_uv = false;
  for (const HostMesh& m : s->meshes) any_uv = any_uv || !m.uv.empty();
  auto mesh_boxes = [](const HostMesh& m) {
    const uint32_t nt = (uint32_t)(m.idx.size() / 3);
    std::vector<HBox> boxes(nt);
    for (uint32_t i = 0; i < nt; ++i) {  // mesh.rs:81-94
      const float* v0 = &m.pos[3 * m.idx[3 * i]];
      const float* v1 = &m.pos[3 * m.idx[3 * i + 1]];
      const float* v2 = &m.pos[3 * m.idx[3 * i + 2]];
      for (int k = 0; k < 3; ++k) {
        boxes[i].lo[k] = smin(smin(v0[k], v1[k]), v2[k]);
        boxes[i].hi[k] = smax(smax(v0[k], v1[k]), v2[k]);
      }
    }
    return boxes;
  };
  // Spatial-split trees (AR_BUILDER_SAH_SPLIT), built per object first and
  // kept where they lower the SAH cost by at least 5%: their references
  // cost a second copy of the triangle records, which the 4-wide walk then
  // reads while shading reads the first (a room of boxes, whose split trees
  // split nothing, rendered 3-4% slower with them).  When some object keeps
  // one, every object's 4-wide tree reads records of its own (qtris), in the
  // leaf order of the tree it is collapsed from.
  struct SplitTree {
    BinTree bin, st;
    std::vector<uint32_t> ref_tri;
    std::vector<HBox> ref_box;
    bool keep = false;
  };
  std::vector<SplitTree> split(s->builder == AR_BUILDER_SAH_SPLIT ? s->meshes.size() : 0);
  {
    // objects are built by a few threads in parallel (a room of 117 objects
    // committed in 2.7 s with one); each big build also splits its own work
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t oi; (oi = next.fetch_add(1)) < split.size();) {
        const HostMesh& m = s->meshes[oi];
        const std::vector<HBox> boxes = mesh_boxes(m);
        const uint32_t nt = (uint32_t)boxes.size();
        SplitTree& sp = split[oi];
        sp.bin = build_sah(boxes, 0);
        sp.st = build_split_sah(boxes, m.pos.data(), m.idx.data(), 2 * (size_t)nt, sp.ref_tri, sp.ref_box);
        sp.keep = nt > 0 && sah_cost(sp.st) < 0.95 * sah_cost(sp.bin);
        if (!sp.keep) {
          sp.st = BinTree();
          std::vector<uint32_t>().swap(sp.ref_tri);
          std::vector<HBox>().swap(sp.ref_box);
        }
      }
    };
    const size_t nthr = std::min<size_t>(split.size(), std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
    std::vector<std::future<void>> fs;
    for (size_t i = 1; i < nthr; ++i) fs.push_back(std::async(std::launch::async, work));
    work();
    for (auto& f : fs) f.get();
  }
  bool any_split = false;
  for (const SplitTree& sp : split) any_split = any_split || sp.keep;
  for (size_t oi = 0; oi < s->meshes.size(); ++oi) {
    const HostMesh& m = s->meshes[oi];
    uint32_t nt = (uint32_t)(m.idx.size() / 3);
    std::vector<HBox> boxes = mesh_boxes(m);
    BinTree t;
    if (s->builder == AR_BUILDER_LBVH) {
      ar_status be = build_lbvh(boxes, t);
      if (be != AR_OK) return be;
    } else if (!split.empty()) {
      t = std::move(split[oi].bin);
    } else {
      t = build_sah(boxes, 0);
    }
    max_depth = std::max(max_depth, t.depth);
    uint32_t tri_base = (uint32_t)tris.size();
    for (size_t i = 0; i < t.nodes.size(); ++i) {
      const BinNode& n = t.nodes[i];
      if (n.count == 0) continue;
      for (uint32_t k = 0; k < n.count; ++k) {
        uint32_t prim = t.prims[n.left_or_start + k];
        const float* v0 = &m.pos[3 * m.idx[3 * prim]];
        const float* v1 = &m.pos[3 * m.idx[3 * prim + 1]];
        const float* v2 = &m.pos[3 * m.idx[3 * prim + 2]];
        Tri tr;
        uint32_t tag = prim | (k + 1 == n.count ? kLastBit : 0u);
        float tagf;
        memcpy(&tagf, &tag, 4);
        // e1.w: the record's own index, the key near-first ties and hits use
        const uint32_t key = tri_base + n.left_or_start + k;
        float keyf;
        memcpy(&keyf, &key, 4);
        tr.v0 = make_float4(v0[0], v0[1], v0[2], tagf);
        tr.e1 = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], keyf);
        tr.e2 = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.0f);
        if (tris.size() != tri_base + n.left_or_start + k) return fail(AR_ERR_STATE, "leaf order mismatch");
        tris.push_back(tr);
      }
    }
    for (uint32_t i = 0; i < nt; ++i) tri_obj.push_back((uint32_t)oi);
    Flat f = flatten(t, nodes, tri_base);
    QFlat qf;
    if (any_split) {
      // the 4-wide tree's own records: copies of the binary tree's, keyed by
      // their index there, in the leaf order of the split tree (kept) or of
      // the binary tree itself
      const SplitTree& sp = split[oi];
      const BinTree& qt = sp.keep ? sp.st : t;
      std::vector<uint32_t> at(nt);
      for (uint32_t p = 0; p < nt; ++p) at[t.prims[p]] = p;
      const uint32_t qbase = (uint32_t)h.qtris.size();
      for (const BinNode& n : qt.nodes) {
        for (uint32_t k = 0; k < n.count; ++k) {
          const uint32_t p = n.left_or_start + k;
          if (h.qtris.size() != qbase + p) return fail(AR_ERR_STATE, "leaf order mismatch");
          const uint32_t tri = sp.keep ? sp.ref_tri[qt.prims[p]] : qt.prims[p];
          Tri tr = tris[tri_base + at[tri]];
          uint32_t tag;
          memcpy(&tag, &tr.v0.w, 4);
          tag = (tag & ~kLastBit) | (k + 1 == n.count ? kLastBit : 0u);
          memcpy(&tr.v0.w, &tag, 4);
          h.qtris.push_back(tr);
        }
      }
      qf = build_qbvh(qt, h.qnodes, qbase, sp.keep ? &sp.ref_box : &boxes);
    } else {
      qf = build_qbvh(t, h.qnodes, tri_base, &boxes);
    }
    h.qmax_depth = std::max(h.qmax_depth, qf.depth);
    Inst& in = inst_by_obj[oi];
    memset(&in, 0, sizeof(in));
    float inv[16];
    inverse16(m.xf, inv);
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 3; ++r) in.inv[3 * c + r] = inv[4 * c + r];
    for (int k = 0; k < 3; ++k) {
      in.box[k] = f.root_box.lo[k];
      in.box[3 + k] = f.root_box.hi[k];
    }
    in.root = f.root_ref;
    in.qroot = qf.root_ref;
    in.id = (uint32_t)oi;
    in.identity = is_identity16(m.xf) ? 1u : 0u;
    in.tri_base = tri_base;
    world[oi] = transform_box(f.root_box, m.xf);  // Instance::new (bvh.rs:373-377)
    DObject& d = objs[oi];
    memset(&d, 0, sizeof(d));
    normal_matrix(m.xf, d.nmat);
    d.idx_base = (uint32_t)gidx.size();
    gidx.insert(gidx.end(), m.idx.begin(), m.idx.end());
    if (!m.nrm.empty()) {
      d.nrm_base = (int32_t)(gnrm.size() / 3);
      gnrm.insert(gnrm.end(), m.nrm.begin(), m.nrm.end());
    } else {
      d.nrm_base = -1;
    }
    d.slot_base = (uint32_t)gslots.size();
    d.n_slots = (uint32_t)m.slots.size();
    gslots.insert(gslots.end(), m.slots.begin(), m.slots.end());
    d.mat_base = (uint32_t)gmats.size();
    d.n_mats = (uint32_t)m.mats.size();
    gmats.insert(gmats.end(), m.mats.begin(), m.mats.end());
    d.tri_base = tri_base;
    d.n_tris = nt;
    // shading records of this object's Tri records (scene.rs:41-48, 75-84)
    for (uint32_t r = 0; r < nt; ++r) {
      uint32_t tag;
      memcpy(&tag, &tris[tri_base + r].v0.w, 4);
      const uint32_t prim = tag & ~kLastBit;
      const uint32_t slot = m.slots.size() == 1 ? m.slots[0] : m.slots[prim];
      const uint32_t mat = m.mats[slot < m.mats.size() - 1 ? slot : (uint32_t)m.mats.size() - 1];
      float n[3][3] = {};
      const uint32_t has = m.nrm.empty() ? 0u : 1u;
      if (has)
        for (int c = 0; c < 3; ++c)
          for (int k = 0; k < 3; ++k) n[c][k] = m.nrm[3 * (size_t)m.idx[3 * prim + c] + k];
      float mf, hf, of;
      const uint32_t ou = (uint32_t)oi;
      memcpy(&mf, &mat, 4);
      memcpy(&hf, &has, 4);
      memcpy(&of, &ou, 4);
      h.gshade.push_back(make_float4(n[0][0], n[0][1], n[0][2], mf));
      h.gshade.push_back(make_float4(n[1][0], n[1][1], n[1][2], hf));
      h.gshade.push_back(make_float4(n[2][0], n[2][1], n[2][2], of));
      if (any_uv) {  // scene.rs:55-64
        float t[3][2] = {};
        const float has_uv = m.uv.empty() ? 0.0f : 1.0f;
        if (!m.uv.empty())
          for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 2; ++k) t[c][k] = m.uv[2 * (size_t)m.idx[3 * prim + c] + k];
        h.guv.push_back(make_float4(t[0][0], t[0][1], t[1][0], t[1][1]));
        h.guv.push_back(make_float4(t[2][0], t[2][1], has_uv, 0.0f));
      }
    }
  }
  // TLAS over instance world boxes (bvh.rs:395-410); instance records are
  // stored in TLAS leaf order.
  BinTree tl = build_sah(world, 0);
  h.tlas_depth = tl.depth;
  h.max_depth = max_depth;
  std::vector<Inst>& insts = h.insts;
  for (size_t i = 0; i < tl.nodes.size(); ++i) {
    const BinNode& n = tl.nodes[i];
    if (n.count == 0) continue;
    for (uint32_t k = 0; k < n.count; ++k) {
      Inst in = inst_by_obj[tl.prims[n.left_or_start + k]];
      if (k + 1 == n.count) in.id |= kLastBit;
      insts.push_back(in);
    }
  }
  Flat tf = flatten(tl, nodes, 0);
  // TLAS leaves of two to four instances become nodes with a world-box slot
  // per instance record (kSingleTri refs), as bottom-level leaves do with
  // triangles: the box test culls instances before their records are read
  QFlat tq = build_qbvh(tl, h.qnodes, 0, &world);
  h.qinsts = insts;
  for (const BinNode& n : tl.nodes)
    if (n.count >= 2 && n.count <= 4)
      for (uint32_t k = 0; k < n.count; ++k) h.qinsts[n.left_or_start + k].id |= kLastBit;
  h.tlas_root = tf.root_ref;
  if (tris.size() >= kSingleTri || h.qtris.size() >= kSingleTri)
    return fail(AR_ERR_INVALID_ARG, "scene has 2^30 or more triangles (or triangle references)");
  h.qtlas_root = tq.root_ref;
  h.qtlas_depth = tq.depth;
  // the 4-wide box test evaluates q * (scl / d) + (org - o) / d per plane;
  // bounded origins and scales keep every term finite for the rays it accepts
  h.wide_ok = 1;
  for (const QNode& q : h.qnodes) {
    const float v[6] = {q.a.x, q.a.y, q.a.z, q.a.w, q.b.x, q.b.y};
    for (float x : v)
      if (!(std::fabs(x) < 0x1p62f)) h.wide_ok = 0;
  }
  for (int k = 0; k < 3; ++k) {
    h.tlas_box[k] = tf.root_box.lo[k];
    h.tlas_box[3 + k] = tf.root_box.hi[k];
  }
  return AR_OK;
}

}  // namespace ar

extern "C" {

ar_status ar_scene_commit(ar_scene* s) try {
  if (!s) return fail(AR_ERR_INVALID_ARG, "NULL scene");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  if (s->meshes.empty()) return fail(AR_ERR_INVALID_ARG, "scene has no objects");
  HostAccel h;
  ar_status be = build_host_accel(s, h);
  if (be != AR_OK) return be;
  s->tlas_depth = h.tlas_depth;
  const uint32_t max_depth = h.max_depth;
  // Stack bound: the binary walk pushes at most one entry per level, the
  // 4-wide walk at most three.
  s->stack_depth = std::max(h.max_depth + h.tlas_depth + 2, 3 * (h.qmax_depth + h.qtlas_depth) + 2);
  // Materials / lights / distribution (lights.rs:115-235).
  std::vector<DLight> dl(s->lights.size());
  bool emit = false;
  for (size_t i = 0; i < s->lights.size(); ++i) {
    const ar_light& l = s->lights[i];
    DLight& d = dl[i];
    memset(&d, 0, sizeof(d));
    d.type = l.type;
    for (int k = 0; k < 3; ++k) {
      d.pos[k] = l.position[k];
      d.intensity[k] = l.intensity[k];
      d.eu[k] = l.edge_u[k];
      d.ev[k] = l.edge_v[k];
    }
    if (l.type == AR_LIGHT_QUAD) {
      emit = true;
      f3 c = cross(mk(d.eu[0], d.eu[1], d.eu[2]), mk(d.ev[0], d.ev[1], d.ev[2]));
      d.area = length(c);
      f3 n = normalize(c);
      d.n[0] = n.x;
      d.n[1] = n.y;
      d.n[2] = n.z;
    }
  }
  uint32_t nl = (uint32_t)dl.size();
  s->has_dist = s->dist_mode != AR_LIGHTDIST_NONE && nl > 0;
  if (s->has_dist) {
    if (s->dist_mode == AR_LIGHTDIST_UNIFORM) {
      for (uint32_t i = 0; i < nl; ++i) dl[i].pmf = 1.0f / (float)nl;
    } else {
      std::vector<float> w(nl);
      float total = 0.0f;
      for (uint32_t i = 0; i < nl; ++i) {
        float p[3];
        for (int k = 0; k < 3; ++k)
          p[k] = dl[i].type == AR_LIGHT_POINT ? dl[i].intensity[k] * (4.0f * kPi)
                                               : dl[i].intensity[k] * (kPi * dl[i].area);
        w[i] = std::fmax(0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2], 0.0f);
      }
      for (uint32_t i = 0; i < nl; ++i) total = total + w[i];
      if (total == 0.0f) {
        float u = 1.0f / (float)nl;
        for (uint32_t i = 0; i < nl; ++i) {
          dl[i].pmf = u;
          dl[i].cdf = (float)(i + 1) * u;
        }
      } else {
        float acc = 0.0f;
        for (uint32_t i = 0; i < nl; ++i) {
          dl[i].pmf = w[i] / total;
          acc += dl[i].pmf;
          dl[i].cdf = acc;
        }
        dl[nl - 1].cdf = 1.0f;
      }
    }
  }
  AR_HIP(upload(s->d_nodes, h.nodes));
  AR_HIP(upload(s->d_qnodes, h.qnodes));
  AR_HIP(upload(s->d_tris, h.tris));
  AR_HIP(upload(s->d_qtris, h.qtris));
  AR_HIP(upload(s->d_insts, h.insts));
  AR_HIP(upload(s->d_qinsts, h.qinsts));
  AR_HIP(upload(s->d_objects, h.objs));
  AR_HIP(upload(s->d_materials, s->materials));
  AR_HIP(upload(s->d_lights, dl));
  AR_HIP(upload(s->d_idx, h.gidx));
  AR_HIP(upload(s->d_nrm, h.gnrm));
  AR_HIP(upload(s->d_slots, h.gslots));
  AR_HIP(upload(s->d_objmats, h.gmats));
  AR_HIP(upload(s->d_tri_obj, h.tri_obj));
  AR_HIP(upload(s->d_stri, h.gshade));
  AR_HIP(upload(s->d_suv, h.guv));
  s->has_uv = !h.guv.empty();
  AR_HIP(upload(s->d_tex, s->texels));
  AR_HIP(upload(s->d_texinfo, s->tex_info));
  s->view.nodes = s->d_nodes.as<WNode>();
  s->view.qnodes = s->d_qnodes.as<QNode>();
  s->view.qtlas_root = h.qtlas_root;
  s->view.tris = s->d_tris.as<Tri>();
  s->view.qtris = h.qtris.empty() ? s->view.tris : s->d_qtris.as<Tri>();
  s->view.insts = s->d_insts.as<Inst>();
  s->view.qinsts = s->d_qinsts.as<Inst>();
  s->view.tlas_root = h.tlas_root;
  for (int k = 0; k < 6; ++k) s->view.tlas_box[k] = h.tlas_box[k];
  s->view.n_inst = (uint32_t)h.insts.size();
  s->view.wide_ok = h.wide_ok;
  s->n_lights = nl;
  s->has_emitters = emit;
  s->h_lights = dl;
  s->h_objects = h.objs;
  s->info.n_objects = (uint32_t)s->meshes.size();
  s->info.n_triangles = (uint32_t)h.tris.size();
  s->info.n_nodes = (uint32_t)h.nodes.size();
  s->info.n_lights = nl;
  s->info.max_depth = max_depth;
  if (!s->tcfg_set) {
    // Dynamic chunk hand-out pays a serialised atomic per chunk of 64 rays:
    // worth it where rays cost enough (a 36-triangle box ran 33% slower with
    // it at 11 chunks per wave, a 280k-triangle room 11-13% faster).
    const uint32_t bal = h.tris.size() >= 4096 ? 1u : 0u;
    s->tcfg.balance = bal;
    s->tcfg_any.balance = bal;
  }
  s->info.device_bytes = s->d_nodes.bytes + s->d_qnodes.bytes + s->d_tris.bytes + s->d_qtris.bytes + s->d_insts.bytes + s->d_qinsts.bytes + s->d_objects.bytes +
                         s->d_materials.bytes + s->d_lights.bytes + s->d_idx.bytes + s->d_nrm.bytes +
                         s->d_slots.bytes + s->d_objmats.bytes + s->d_tri_obj.bytes + s->d_stri.bytes;
  AR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  s->committed = true;
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_scene_commit");
}

ar_status ar_scene_get_info(const ar_scene* s, ar_scene_info* info) {
  if (!s || !info) return fail(AR_ERR_INVALID_ARG, "NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "scene not committed");
  *info = s->info;
  return AR_OK;
}

}  // extern "C"
