// Render workspace and integrator entry points (host side).
#pragma once
#include "ar_internal.h"

namespace ar {

// Per-path wavefront state, queues and counters (structure of arrays).
struct WaveBufs {
  float4* ray_o;    // (o.xyz, tmin) per path
  float4* ray_d;    // (d.xyz, tmax)
  float4* beta;     // (beta.xyz, prev_pdf)
  float4* L;        // (L.xyz, -): per-path radiance, reduced into the film
  float4* prev_p;   // (p.xyz, -)
  uint4* meta;      // (pixel, sample, dim, depth)
  uint32_t* q[2];   // active path queues (ping-pong)
  float4* hit;      // per queue slot: (t, pos, u, v)
  uint32_t* hit_inst;
  uint32_t* sh_path;  // shadow queue: owning path
  float4* sh_o;       // shadow ray origin, tmin
  float4* sh_d;       // shadow ray dir, tmax
  float4* sh_c;       // contribution if unoccluded
  uint32_t* cnt;      // [0],[1] queue counts, [2] shadow count
  unsigned long long* stats;  // [0] extension rays, [1] shadow rays
  uint32_t* work;     // walker counters of the traversal launches (kWalkWorkBytes; walk_items), passed
                      // to a launch whose configuration has balance set
};

// A chunk's launch sequence captured once in a HIP graph and replayed for
// every later chunk whose launches are identical.  The key holds everything
// the captured launches take by value (buffers, grids, launch configuration,
// scene views); what changes per chunk is read from device memory.
struct GraphCache {
  static constexpr int kKeyWords = 12;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool valid = false;
  uint64_t key[kKeyWords] = {};
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
  // Launches `record` on st through the cached graph, capturing it first when
  // the key differs from the cached one.
  template <class F>
  ar_status launch(const uint64_t (&k)[kKeyWords], hipStream_t st, const char* what, F&& record) {
    if (!valid || std::memcmp(k, key, sizeof(key)) != 0) {
      reset();
      AR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      record();
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (le != hipSuccess) return fail(AR_ERR_DEVICE, std::string(what) + " capture: " + hipGetErrorString(le));
      AR_HIP(ce);
      AR_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      std::memcpy(key, k, sizeof(key));
      valid = true;
    }
    AR_HIP(hipGraphLaunch(exec, st));
    return AR_OK;
  }
  ~GraphCache() { reset(); }
};

// Bytes of the device block that holds a chunk's integrator parameters.
constexpr size_t kParamsBytes = 1024;

struct ShadeView;
// FNV-1a over every field of a scene view, as held by value in captured launches.
uint64_t view_hash(const ShadeView& v);
// Closest-hit and any-hit launch configurations packed into key words.
inline uint64_t launch_key(const ar_trace_config& tc, const ar_trace_config& ta) {
  return ((uint64_t)tc.workgroup_size << 48) | ((uint64_t)tc.lds_stack << 40) | ((uint64_t)tc.balance << 38) |
         ((uint64_t)tc.park_slots << 32) | ((uint64_t)ta.workgroup_size << 16) | ((uint64_t)ta.lds_stack << 8) |
         ((uint64_t)ta.balance << 6) | ta.park_slots;
}
inline uint64_t launch_key_thresholds(const ar_trace_config& tc, const ar_trace_config& ta) {
  return ((uint64_t)tc.refill_min << 48) | ((uint64_t)tc.search_min << 32) | ((uint64_t)ta.refill_min << 16) |
         ta.search_min;
}

// Device memory of one render in flight (one per scene for ar_render, one
// per job slot for the server).  Grows as jobs need; trim() frees it.
struct Workspace {
  DevBuf buf;       // path state + queues + film accumulator + stack spill
  DevBuf out;       // staging for host output
  DevBuf aux;       // integrator-specific buffers (GDPT films, BDPT vertices)
  DevBuf pois;      // Poisson CG vectors
  WaveBufs B{};
  float4* acc = nullptr;  // per-pixel film accumulators (3 planes of npix)
  int2* spill = nullptr;  // traversal stack spill, [entry][thread]
  uint32_t capacity = 0;
  void* params = nullptr;  // device copy of the chunk's parameters (kParamsBytes), read by captured stages
  GraphCache pt_graph;     // path tracer / GDPT bounce loop of one chunk
  GraphCache bdpt_graph[2];  // BDPT chunk: camera and light subpath bounce loops
  ar_status reserve_paths(uint32_t cap, uint32_t npix, uint32_t spill_entries, uint32_t grid_threads,
                          hipStream_t st);
  size_t bytes() const { return buf.bytes + out.bytes + aux.bytes + pois.bytes; }
  void trim();  // frees every buffer (the caller has drained the stream that used them)
};

uint32_t stack_spill_entries(const ar_scene* s);

// Screened Poisson CG on device buffers (interleaved RGB films); the
// scratch grows to poisson_scratch_bytes(w, h) if smaller.
size_t poisson_scratch_bytes(uint32_t w, uint32_t h);
ar_status poisson_solve_device(const float* P_rgb, const float* gx_rgb, const float* gy_rgb, uint32_t w, uint32_t h,
                               float alpha, uint32_t iters, int32_t sign, float* out_rgb, double* final_residual,
                               DevBuf& scratch, hipStream_t st);

struct JobSpec;
// Samples per pixel issued per chunk (bounds in-flight path state).
uint32_t chunk_samples(const JobSpec& spec);

}  // namespace ar
