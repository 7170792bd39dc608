// Device BVH builder (an alternative to the host full-sweep SAH of
// agent-ray src/accel/bvh.rs:71-236 for large meshes, e.g. scene swaps in a
// server): a linear BVH after Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees" (HPG 2012).
//   1. k_lbvh_morton: 30-bit Morton code of each primitive's centroid in the
//      centroid bounds (10 bits per axis), value = primitive index;
//   2. rocPRIM radix sort of (code, index) pairs (stable: equal codes keep
//      index order);
//   3. k_lbvh_karras: one thread per internal node i finds its key range
//      [min(i, j), max(i, j)] and its split from the common-prefix lengths
//      delta (ties between equal codes broken by index), giving the two
//      children (leaf = sorted position, internal = node index).
// The host then walks the hierarchy from the root into the preorder tree the
// rest of the commit consumes (BinTree: leaves of up to four primitives,
// contiguous in leaf order; boxes are unions of the primitives' boxes, exact
// min/max).  The answer of a query does not depend on the tree, only the
// work does; SAH trees trace faster, this one builds in milliseconds.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "ar_internal.h"
#include "ar_kernels.h"

namespace ar {

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {  // 10 bits -> every third bit
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_lbvh_morton(const float* __restrict__ boxes, uint32_t n, float lx, float ly, float lz, float sx,
                              float sy, float sz, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* b = boxes + 6 * (size_t)i;
    const float c[3] = {(b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f};
    const float q[3] = {(c[0] - lx) * sx, (c[1] - ly) * sy, (c[2] - lz) * sz};
    uint32_t m[3];
    for (int k = 0; k < 3; ++k) m[k] = (uint32_t)fminf(fmaxf(q[k] * 1024.0f, 0.0f), 1023.0f);
    keys[i] = (expand_bits10(m[0]) << 2) | (expand_bits10(m[1]) << 1) | expand_bits10(m[2]);
    vals[i] = i;
  }
}

// Length of the common prefix of sorted keys i and j (index bits appended
// for equal keys); -1 outside [0, n).
__device__ __forceinline__ int lbvh_delta(const uint32_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

// child[i] = (left, right): >= 0 internal node, < 0 leaf ~position;
// range[i] = (first, last) sorted positions covered by internal node i.
__global__ void k_lbvh_karras(const uint32_t* __restrict__ keys, int n, int2* __restrict__ child,
                              int2* __restrict__ range) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
    const int d = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lbvh_delta(keys, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (lbvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = lbvh_delta(keys, n, i, j);
    int s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (lbvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    child[i] = make_int2(lo == gamma ? ~gamma : gamma, hi == gamma + 1 ? ~(gamma + 1) : gamma + 1);
    range[i] = make_int2(lo, hi);
  }
}

static inline void box_union(HBox& a, const HBox& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::min(a.lo[k], b.lo[k]);
    a.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
}

ar_status build_lbvh(const std::vector<HBox>& boxes, BinTree& out) {
  out = BinTree();
  const uint32_t n = (uint32_t)boxes.size();
  if (n == 0) return AR_OK;
  if (n >= (1u << 30)) return fail(AR_ERR_INVALID_ARG, "build_lbvh: too many primitives");
  // centroid bounds on the host (one pass)
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const HBox& b : boxes)
    for (int k = 0; k < 3; ++k) {
      const float c = (b.lo[k] + b.hi[k]) * 0.5f;
      lo[k] = std::min(lo[k], c);
      hi[k] = std::max(hi[k], c);
    }
  float sc[3];
  for (int k = 0; k < 3; ++k) sc[k] = hi[k] > lo[k] ? 1.0f / (hi[k] - lo[k]) : 0.0f;
  std::vector<uint32_t> order(n);
  std::vector<int2> child(n > 1 ? n - 1 : 0), range(n > 1 ? n - 1 : 0);
  if (n > 1) {
    DevBuf d_boxes, d_kv, d_child, d_tmp;
    const size_t kv = (size_t)n * sizeof(uint32_t);
    AR_HIP(d_boxes.alloc((size_t)n * sizeof(HBox)));
    AR_HIP(d_kv.alloc(4 * kv));
    AR_HIP(d_child.alloc(2 * (size_t)(n - 1) * sizeof(int2)));
    uint32_t* keys = d_kv.as<uint32_t>();
    uint32_t* vals = keys + n;
    uint32_t* keys_s = vals + n;
    uint32_t* vals_s = keys_s + n;
    AR_HIP(hipMemcpy(d_boxes.p, boxes.data(), (size_t)n * sizeof(HBox), hipMemcpyHostToDevice));
    const uint32_t g = std::min<uint32_t>((n + 255) / 256, 65535u);
    hipLaunchKernelGGL(k_lbvh_morton, dim3(g), dim3(256), 0, nullptr, d_boxes.as<float>(), n, lo[0], lo[1], lo[2],
                       sc[0], sc[1], sc[2], keys, vals);
    AR_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    AR_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys_s, vals, vals_s, n, 0, 30, nullptr));
    AR_HIP(d_tmp.alloc(std::max<size_t>(tmp_bytes, 16)));
    AR_HIP(rocprim::radix_sort_pairs(d_tmp.p, tmp_bytes, keys, keys_s, vals, vals_s, n, 0, 30, nullptr));
    int2* d_c = d_child.as<int2>();
    int2* d_r = d_c + (n - 1);
    const uint32_t g2 = std::min<uint32_t>((n - 1 + 255) / 256, 65535u);
    hipLaunchKernelGGL(k_lbvh_karras, dim3(g2), dim3(256), 0, nullptr, keys_s, (int)n, d_c, d_r);
    AR_HIP(hipGetLastError());
    AR_HIP(hipMemcpy(order.data(), vals_s, kv, hipMemcpyDeviceToHost));
    AR_HIP(hipMemcpy(child.data(), d_c, (size_t)(n - 1) * sizeof(int2), hipMemcpyDeviceToHost));
    AR_HIP(hipMemcpy(range.data(), d_r, (size_t)(n - 1) * sizeof(int2), hipMemcpyDeviceToHost));
  } else {
    order[0] = 0;
  }
  // preorder walk from the root (internal node 0); ranges of <= 4 become leaves
  constexpr uint32_t kLeafMax = 4;
  out.nodes.reserve(2 * (size_t)n);
  out.prims.reserve(n);
  struct Item {
    int ref;           // >= 0 internal, < 0 leaf ~position
    uint32_t parent;   // node index of the parent, UINT32_MAX for the root
    uint32_t depth;
    bool right;        // the parent's right child
  };
  std::vector<Item> stack;
  stack.push_back({n > 1 ? 0 : ~0, UINT32_MAX, 1, false});
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const uint32_t idx = (uint32_t)out.nodes.size();
    BinNode nd{};
    int first, last;
    if (it.ref < 0) {
      first = last = ~it.ref;
    } else {
      first = range[it.ref].x;
      last = range[it.ref].y;
    }
    if (it.ref < 0 || (uint32_t)(last - first + 1) <= kLeafMax) {
      nd.left_or_start = (uint32_t)out.prims.size();
      nd.count = (uint32_t)(last - first + 1);
      nd.right_or_end = 0;  // as build_sah's leaves
      nd.box = boxes[order[first]];
      for (int k = first; k <= last; ++k) {
        out.prims.push_back(order[k]);
        box_union(nd.box, boxes[order[k]]);
      }
      out.nodes.push_back(nd);
    } else {
      out.nodes.push_back(nd);  // children and box filled below
      stack.push_back({child[it.ref].y, idx, it.depth + 1, true});
      stack.push_back({child[it.ref].x, idx, it.depth + 1, false});
    }
    out.depth = std::max(out.depth, it.depth);
    if (it.parent != UINT32_MAX) {
      BinNode& p = out.nodes[it.parent];
      if (it.right)
        p.right_or_end = idx;
      else
        p.left_or_start = idx;
    }
  }
  // internal boxes: children follow their parent in preorder, so a reverse
  // sweep sees both children before the parent
  for (size_t i = out.nodes.size(); i-- > 0;) {
    BinNode& nd = out.nodes[i];
    if (nd.count != 0) continue;
    nd.box = out.nodes[nd.left_or_start].box;
    box_union(nd.box, out.nodes[nd.right_or_end].box);
  }
  return AR_OK;
}

}  // namespace ar

// The device builder on an explicit list of boxes, in ar_bvh_build_host's
// output format (for the tests).
extern "C" ar_status ar_bvh_build_device(const float* boxes, uint32_t n, float* out_boxes, uint32_t* out_meta,
                                         uint32_t* out_prims, uint32_t* out_nodes, uint32_t* out_depth) try {
  if ((!boxes && n) || !out_boxes || !out_meta || !out_prims || !out_nodes)
    return ar::fail(AR_ERR_INVALID_ARG, "ar_bvh_build_device: NULL argument");
  std::vector<ar::HBox> b(n);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      b[i].lo[k] = boxes[6 * i + k];
      b[i].hi[k] = boxes[6 * i + 3 + k];
    }
  ar::BinTree t;
  ar_status e = ar::build_lbvh(b, t);
  if (e != AR_OK) return e;
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
} catch (...) {
  return ar::fail_exception("ar_bvh_build_device");
}
