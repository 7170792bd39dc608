// Oracle (test infrastructure only): two-level SAH BVH restated from the
// reference's src/accel/bvh.rs and src/primitives/mesh.rs.
#include <algorithm>
#include <future>
#include <numeric>

#include "orc.h"

namespace orc {

bool is_identity(const M4& m) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      if (m.c[i][j] != (i == j ? 1.0f : 0.0f)) return false;
  return true;
}

// General 4x4 inverse by cofactors.  glam's Mat4::inverse is not available
// here; for the identity and for translations/power-of-two scales the result
// is exact either way.  Other matrices: parity unpinned.
M4 inverse(const M4& mm) {
  const float* m = &mm.c[0][0];  // column-major flat: m[col*4+row]
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
  M4 r;
  float* o = &r.c[0][0];
  float rd = 1.0f / det;
  for (int i = 0; i < 16; ++i) o[i] = inv[i] * rd;
  if (is_identity(mm)) return M4::identity();
  return r;
}

// scene.rs:33 — Mat3::from_mat4(transform).inverse().transpose().
M3 normal_matrix(const M4& m) {
  float a[3][3];  // a[col][row]
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = m.c[i][j];
  M3 r;
  bool ident = true;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (a[i][j] != (i == j ? 1.0f : 0.0f)) ident = false;
  if (ident) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.c[i][j] = (i == j) ? 1.0f : 0.0f;
    return r;
  }
  // element (row, col) = a[col][row]
  auto e = [&](int row, int col) { return a[col][row]; };
  float c00 = e(1, 1) * e(2, 2) - e(1, 2) * e(2, 1);
  float c01 = e(1, 2) * e(2, 0) - e(1, 0) * e(2, 2);
  float c02 = e(1, 0) * e(2, 1) - e(1, 1) * e(2, 0);
  float det = e(0, 0) * c00 + e(0, 1) * c01 + e(0, 2) * c02;
  float id = 1.0f / det;
  // inverse(row,col) = cofactor(col,row)/det; transpose -> cofactor(row,col)/det
  float cof[3][3];
  cof[0][0] = c00;
  cof[0][1] = c01;
  cof[0][2] = c02;
  cof[1][0] = e(0, 2) * e(2, 1) - e(0, 1) * e(2, 2);
  cof[1][1] = e(0, 0) * e(2, 2) - e(0, 2) * e(2, 0);
  cof[1][2] = e(0, 1) * e(2, 0) - e(0, 0) * e(2, 1);
  cof[2][0] = e(0, 1) * e(1, 2) - e(0, 2) * e(1, 1);
  cof[2][1] = e(0, 2) * e(1, 0) - e(0, 0) * e(1, 2);
  cof[2][2] = e(0, 0) * e(1, 1) - e(0, 1) * e(1, 0);
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) r.c[col][row] = cof[row][col] * id;
  return r;
}

AABB AABB::transform(const M4& m) const {  // geometry.rs:66-84
  V3 corners[8] = {V3(lo.x, lo.y, lo.z), V3(hi.x, lo.y, lo.z), V3(lo.x, hi.y, lo.z),
                   V3(hi.x, hi.y, lo.z), V3(lo.x, lo.y, hi.z), V3(hi.x, lo.y, hi.z),
                   V3(lo.x, hi.y, hi.z), V3(hi.x, hi.y, hi.z)};
  AABB r = AABB::empty();
  for (auto c : corners) {
    V3 tc = transform_point(m, c);
    r.lo = vmin(r.lo, tc);
    r.hi = vmax(r.hi, tc);
  }
  return r;
}

AABB Mesh::prim_aabb(uint32_t i) const {  // mesh.rs:81-94
  V3 v0 = pos[idx[3 * i]], v1 = pos[idx[3 * i + 1]], v2 = pos[idx[3 * i + 2]];
  return AABB{vmin(vmin(v0, v1), v2), vmax(vmax(v0, v1), v2)};
}

// ---------------------------------------------------------------------------
// SAH build (bvh.rs:57-236)
// ---------------------------------------------------------------------------
namespace {

struct Builder {
  const std::vector<AABB>& boxes;
  explicit Builder(const std::vector<AABB>& b) : boxes(b) {}

  // bvh.rs:71-150.  Returns split position or -1 for "make a leaf".
  long sah_split(std::vector<uint32_t>& indices) {
    size_t n = indices.size();
    std::vector<AABB> pb(n);
    for (size_t i = 0; i < n; ++i) pb[i] = boxes[indices[i]];
    AABB parent = AABB::empty();
    for (auto& b : pb) parent = parent.merge(b);
    float parent_sa = parent.surface_area();
    if (parent_sa < kF32Eps) return -1;
    const float C_TRAV = 1.0f;
    float leaf_cost = (float)n;
    float best_cost = leaf_cost;
    int best_axis = -1;
    size_t best_k = 0;
    std::vector<uint32_t> sorted(n);
    std::iota(sorted.begin(), sorted.end(), 0u);
    std::vector<AABB> prefix(n), suffix(n);
    auto sort_axis = [&](int axis) {
      std::sort(sorted.begin(), sorted.end(), [&](uint32_t a, uint32_t b) {
        float ca = pb[a].centroid()[axis], cb = pb[b].centroid()[axis];
        if (ca < cb) return true;
        if (cb < ca) return false;
        return indices[a] < indices[b];  // tie-break: primitive index
      });
    };
    for (int axis = 0; axis < 3; ++axis) {
      sort_axis(axis);
      prefix[0] = pb[sorted[0]];
      for (size_t i = 1; i < n; ++i) prefix[i] = prefix[i - 1].merge(pb[sorted[i]]);
      suffix[n - 1] = pb[sorted[n - 1]];
      for (size_t i = n - 1; i-- > 0;) suffix[i] = suffix[i + 1].merge(pb[sorted[i]]);
      for (size_t k = 1; k < n; ++k) {
        float cost = C_TRAV + (prefix[k - 1].surface_area() * (float)k +
                               suffix[k].surface_area() * (float)(n - k)) /
                                  parent_sa;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_k = k;
        }
      }
    }
    if (best_axis < 0) return -1;
    if (best_axis != 2) sort_axis(best_axis);
    std::vector<uint32_t> old = indices;
    for (size_t i = 0; i < n; ++i) indices[i] = old[sorted[i]];
    return (long)best_k;
  }

  // bvh.rs:199-236: returns (nodes, prims) of the subtree with node 0 the
  // root; subtrees below depth 3 run on their own threads (the reference uses
  // rayon::join below depth 8; the result does not depend on it).
  BVH build(std::vector<uint32_t> indices, int depth) {
    AABB box = AABB::empty();
    for (auto i : indices) box = box.merge(boxes[i]);
    long split = indices.size() <= 4 ? -1 : sah_split(indices);
    if (split < 0) {
      BVH leaf;
      leaf.nodes.push_back(BVHNode{box, 0, 0, (uint32_t)indices.size()});
      leaf.prims = std::move(indices);
      return leaf;
    }
    std::vector<uint32_t> right(indices.begin() + split, indices.end());
    indices.resize(split);
    BVH l, r;
    if (depth < 3 && indices.size() + right.size() > 4096) {
      auto fut = std::async(std::launch::async, [&]() { return build(std::move(right), depth + 1); });
      l = build(std::move(indices), depth + 1);
      r = fut.get();
    } else {
      l = build(std::move(indices), depth + 1);
      r = build(std::move(right), depth + 1);
    }
    return merge(box, std::move(l), std::move(r));
  }

  // bvh.rs:157-193: rebase child/leaf references and concatenate.
  static BVH merge(const AABB& box, BVH l, BVH r) {
    uint32_t ln = (uint32_t)l.nodes.size(), lp = (uint32_t)l.prims.size();
    for (auto& n : l.nodes) {
      if (n.is_leaf()) continue;
      n.left_or_start += 1;
      n.right_or_end += 1;
    }
    for (auto& n : r.nodes) {
      if (n.is_leaf()) {
        n.left_or_start += lp;
      } else {
        n.left_or_start += 1 + ln;
        n.right_or_end += 1 + ln;
      }
    }
    BVH out;
    out.nodes.reserve(1 + l.nodes.size() + r.nodes.size());
    out.nodes.push_back(BVHNode{box, 1, 1 + ln, 0});
    out.nodes.insert(out.nodes.end(), l.nodes.begin(), l.nodes.end());
    out.nodes.insert(out.nodes.end(), r.nodes.begin(), r.nodes.end());
    out.prims = std::move(l.prims);
    out.prims.insert(out.prims.end(), r.prims.begin(), r.prims.end());
    return out;
  }
};

}  // namespace

BVH build_bvh(const std::vector<AABB>& prim_boxes) {
  if (prim_boxes.empty()) return BVH();
  std::vector<uint32_t> idx(prim_boxes.size());
  std::iota(idx.begin(), idx.end(), 0u);
  Builder b(prim_boxes);
  return b.build(std::move(idx), 0);
}

// ---------------------------------------------------------------------------
// Traversal (bvh.rs:247-296): left-first DFS, 64-entry stack, current best t
// used for box culling, strict `<` acceptance (first found wins ties).
// ---------------------------------------------------------------------------
template <class LeafFn>
static bool traverse(const BVH& bvh, const Ray& ray, LeafFn&& leaf, TraceCounters* c) {
  if (bvh.nodes.empty()) return false;
  uint32_t stack[64];
  int sp = 0;
  bool found = false;
  float tmax = ray.tmax;
  stack[sp++] = 0;
  while (sp > 0) {
    const BVHNode& node = bvh.nodes[stack[--sp]];
    Ray probe = ray;
    probe.tmax = tmax;
    if (c) c->nodes++;
    if (!node.box.intersect(probe)) continue;
    if (node.is_leaf()) {
      for (uint32_t k = node.left_or_start; k < node.left_or_start + node.count; ++k) {
        float t;
        if (leaf(bvh.prims[k], tmax, &t) && t < tmax) {
          tmax = t;
          found = true;
        }
      }
    } else {
      if (sp + 2 > 64) abort();  // bvh.rs:287 debug_assert
      stack[sp++] = node.right_or_end;
      stack[sp++] = node.left_or_start;
    }
  }
  return found;
}

void Blas::build() {
  std::vector<AABB> pb(mesh->ntri());
  for (uint32_t i = 0; i < mesh->ntri(); ++i) pb[i] = mesh->prim_aabb(i);
  bvh = build_bvh(pb);
  aabb = bvh.nodes.empty() ? AABB::empty() : bvh.nodes[0].box;
}

bool Blas::intersect(const Ray& ray, RayHit* hit, TraceCounters* c) const {
  RayHit best;
  bool any = traverse(
      bvh, ray,
      [&](uint32_t prim, float tmax, float* tout) {  // mesh.rs:96-116
        const Mesh& m = *mesh;
        V3 v0 = m.pos[m.idx[3 * prim]], v1 = m.pos[m.idx[3 * prim + 1]], v2 = m.pos[m.idx[3 * prim + 2]];
        Ray lr = ray;
        lr.tmax = tmax;
        HitInfo hi;
        if (c) c->tris++;
        if (!triangle_hit(v0, v1 - v0, v2 - v0, lr, &hi)) return false;
        *tout = hi.t;
        if (hi.t < tmax) {
          best.inst = 0;
          best.prim = prim;
          best.n = hi.n;
          best.u = hi.u;
          best.v = hi.v;
          best.t = hi.t;
        }
        return true;
      },
      c);
  if (any && hit) *hit = best;
  return any;
}

void Tlas::build() {
  for (auto& in : inst) {
    in.inv = inverse(in.xf);
    in.world = in.blas->aabb.transform(in.xf);
  }
  std::vector<AABB> pb(inst.size());
  for (size_t i = 0; i < inst.size(); ++i) pb[i] = inst[i].world;
  bvh = build_bvh(pb);
}

bool Tlas::intersect(const Ray& ray, RayHit* hit, TraceCounters* c) const {
  RayHit best;
  bool any = traverse(
      bvh, ray,
      [&](uint32_t id, float tmax, float* tout) {  // bvh.rs:427-445
        const Instance& in = inst[id];
        V3 lo = transform_point(in.inv, ray.o);
        V3 ld = transform_vector(in.inv, ray.d);
        float s = length(ld);
        if (s < kF32Eps) return false;
        Ray lr = Ray::make(lo, ld, ray.tmin * s, tmax * s);
        RayHit h;
        if (!in.blas->intersect(lr, &h, c)) return false;
        h.t = h.t / s;
        h.inst = id;
        *tout = h.t;
        if (h.t < tmax) best = h;
        return true;
      },
      c);
  if (any && hit) *hit = best;
  return any;
}

}  // namespace orc
