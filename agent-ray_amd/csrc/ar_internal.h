// Host-side internals of libagentray: scene handle, device buffers, errors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/agentray.h"
#include "ar_device.h"

namespace ar {

// ---- errors -----------------------------------------------------------------
void set_error(const std::string& msg);
ar_status fail(ar_status code, const std::string& msg);
// The status for the exception in flight (host memory -> AR_ERR_OUT_OF_BUDGET,
// anything else -> AR_ERR_INVALID_ARG).  Entry points that allocate on the
// host are function-try-blocks ending in it, so nothing unwinds through the
// C ABI.
ar_status fail_exception(const char* fn);
#define AR_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::ar::fail(AR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- device buffer (owned) ---------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// ---- SAH BVH over a set of boxes --------------------------------------------
struct HBox {
  float lo[3], hi[3];
};

// Binary SAH tree in the reference's form (bvh.rs:32-51): preorder node
// array, leaf = (start, count) into `prims`.
struct BinNode {
  HBox box;
  uint32_t left_or_start, right_or_end, count;
};
struct BinTree {
  std::vector<BinNode> nodes;
  std::vector<uint32_t> prims;
  uint32_t depth = 0;
};
// Exact full-sweep SAH (same split decisions as bvh.rs:71-150) computed with
// per-axis presorted lists and stable partitioning: O(n log n) per level.
BinTree build_sah(const std::vector<HBox>& boxes, int nthreads);
// Linear BVH built on the device (Morton order, Karras hierarchy), the same
// form as build_sah's output with leaves of up to four primitives.
ar_status build_lbvh(const std::vector<HBox>& boxes, BinTree& out);
// SAH tree with spatial splits over triangle references, for near-first
// traversal only: a triangle may be referenced from several leaves, each
// time with the box of its part inside that leaf's region.  `pos` / `idx`
// are the mesh (3 floats per vertex, 3 indices per triangle), `boxes` the
// triangles' boxes.  Leaf position p of the result holds triangle
// ref_tri[p] with box ref_box[p]; `prims` is the identity.  Spatial splits
// stop once the references reach max_refs.
BinTree build_split_sah(const std::vector<HBox>& boxes, const float* pos, const uint32_t* idx, size_t max_refs,
                        std::vector<uint32_t>& ref_tri, std::vector<HBox>& ref_box);

// ---- materials / lights as stored on the device -----------------------------
struct alignas(16) DMaterial {
  uint32_t type;
  int32_t light_id;  // >= 0 for an emitter (quad light geometry)
  float roughness, eta;
  float albedo[3];
  int32_t tex = -1;  // albedo texture (index into the texture table), -1: constant albedo
  float emission[3];
  float pad1;
};
static_assert(sizeof(DMaterial) == 48, "material layout");

struct alignas(16) DLight {
  uint32_t type;
  float area;
  float pmf;  // selection probability (light distribution)
  float cdf;  // inclusive CDF value (power distribution)
  float pos[3];
  float pad0;
  float intensity[3];
  float pad1;
  float eu[3];
  float pad2;
  float ev[3];
  float pad3;
  float n[3];
  float pad4;
};
static_assert(sizeof(DLight) == 96, "light layout");

// Per-object shading record (scene.rs:16-88).
struct alignas(16) DObject {
  float nmat[9];        // normal matrix, column-major 3x3
  uint32_t idx_base;    // first index (3 per triangle) in the global index array
  int32_t nrm_base;     // first vertex normal, or -1 when the mesh has none
  uint32_t slot_base;   // first material slot entry
  uint32_t n_slots;     // 1 or triangle count
  uint32_t mat_base;    // first entry of this object's material-id list
  uint32_t n_mats;
  uint32_t tri_base;    // first Tri record (leaf order)
  uint32_t n_tris;
  uint32_t pad[3];
};
static_assert(sizeof(DObject) == 80, "object layout");

struct HostMesh {
  std::vector<float> pos, nrm, uv;  // uv: tex_coords, empty when absent
  std::vector<uint32_t> idx, slots, mats;
  float xf[16];
};

}  // namespace ar

namespace ar {
struct Workspace;
}

struct ar_scene {
  ar::Workspace* ws = nullptr;  // render workspace used by ar_render
  // paths (BDPT: samples) per chunk of ar_render when sample_chunk is 0, by
  // integrator; 0 = the library default; set by ar_render_autotune_chunk
  uint64_t chunk_target[3] = {0, 0, 0};
  std::vector<ar::DMaterial> materials;
  std::vector<ar::HostMesh> meshes;
  std::vector<float4> texels;    // linear texels of every texture, back to back
  std::vector<uint4> tex_info;   // per texture: (first texel, width, height, -)
  std::vector<ar_light> lights;
  uint32_t dist_mode = AR_LIGHTDIST_POWER;
  uint32_t builder = AR_BUILDER_SAH_SPLIT;  // bottom-level tree builder
  bool committed = false;
  // host copies of what was uploaded (used by tests through the info call)
  ar_scene_info info{};
  // device-resident scene
  ar::DevBuf d_scratch;
  hipEvent_t scratch_event = nullptr;
  uint32_t tlas_depth = 0;
  uint32_t stack_depth = 0;  // traversal stack entries a query can need (either tree form)
  // Launch configurations until the scene is tuned: the 12-entry LDS stack
  // with refills of 8 lanes renders the room's PT, BDPT and GDPT jobs 12-13%
  // faster than the 8-entry stack with refills of 24.  `balance` is set at
  // commit from the triangle count unless a configuration was set before.
  ar_trace_config tcfg{256, 12, 8, 8, 8, 3, 1, 0.0f};      // closest-hit launch configuration (autotuned)
  ar_trace_config tcfg_any{256, 12, 8, 8, 8, 3, 1, 0.0f};  // any-hit (occlusion, shadow rays), tuned separately
  bool tcfg_set = false;  // a configuration was set or tuned
  ar::DevBuf d_nodes, d_qnodes, d_tris, d_qtris, d_insts, d_qinsts, d_objects, d_materials, d_lights, d_idx, d_nrm, d_slots, d_objmats,
      d_tri_obj, d_stri, d_suv, d_tex, d_texinfo;
  ar::SceneView view{};
  uint32_t n_lights = 0;
  bool has_dist = false;
  bool has_emitters = false;
  bool has_uv = false;  // some mesh has tex_coords (d_suv holds the per-Tri-record uvs)
  std::vector<ar::DLight> h_lights;
  std::vector<ar::DObject> h_objects;
  hipStream_t stream = nullptr;
  std::mutex mu;
};

namespace ar {
hipStream_t pick_stream(const ar_scene* s, void* stream);
// Looks up the current device's address of the allocation-free traversal
// scratch (ar_trace_noalloc) once per device.
ar_status noalloc_setup_current_device();

// Host-side result of building a committed scene's acceleration structure
// and shading records (what ar_scene_commit uploads).
struct HostAccel {
  std::vector<WNode> nodes;
  std::vector<QNode> qnodes;
  std::vector<Tri> tris;
  // Tri records of the 4-wide trees when they are built with spatial splits
  // (AR_BUILDER_SAH_SPLIT): leaf order of those trees, a triangle possibly
  // more than once; empty otherwise (the 4-wide trees then use `tris`)
  std::vector<Tri> qtris;
  std::vector<Inst> insts;
  // the instance records as the 4-wide TLAS reads them: kLastBit also on
  // every record of a TLAS leaf of two to four instances (one slot each)
  std::vector<Inst> qinsts;
  std::vector<DObject> objs;
  std::vector<uint32_t> gidx, gslots, gmats, tri_obj;
  std::vector<float> gnrm;
  // per Tri record (leaf order), 3 float4: (n0.xyz, material id bits),
  // (n1.xyz, 1 if vertex normals else 0), (n2.xyz, object id bits); object
  // space vertex normals of the triangle, resolved from the object's index,
  // normal, slot and material tables
  std::vector<float4> gshade;
  std::vector<float4> guv;  // per Tri record: (u0, v0, u1, v1), (u2, v2, has_uv, -); empty if no mesh has uvs
  int tlas_root = 0, qtlas_root = 0;
  float tlas_box[6] = {0, 0, 0, 0, 0, 0};
  uint32_t max_depth = 0, tlas_depth = 0;    // binary trees
  uint32_t qmax_depth = 0, qtlas_depth = 0;  // 4-wide trees
  uint32_t wide_ok = 0;                      // see SceneView::wide_ok
};
ar_status build_host_accel(const ar_scene* s, HostAccel& h);
}
namespace ar {
void destroy_workspace(Workspace* ws);
}
