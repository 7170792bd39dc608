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
};

// Device memory of one render in flight (one per scene for ar_render, one
// per job slot for the server).  Grows monotonically, never shrinks.
struct Workspace {
  DevBuf buf;       // path state + queues + film accumulator + stack spill
  DevBuf out;       // staging for host output
  DevBuf aux;       // integrator-specific buffers (GDPT films, BDPT vertices)
  DevBuf pois;      // Poisson CG vectors
  WaveBufs B{};
  float4* acc = nullptr;  // per-pixel film accumulators (3 planes of npix)
  int2* spill = nullptr;  // traversal stack spill, [entry][thread]
  uint32_t capacity = 0;
  void* params = nullptr;  // device copy of the chunk's PtParams (read by the generate / shade stages)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  uint64_t graph_key[8] = {};  // what the captured bounce loop was recorded for
  ar_status reserve_paths(uint32_t cap, uint32_t npix, uint32_t spill_entries, uint32_t grid_threads,
                          hipStream_t st);
  size_t bytes() const { return buf.bytes + out.bytes + aux.bytes + pois.bytes; }
  void trim();  // frees every buffer (the caller has drained the stream that used them)
  ~Workspace();
};

uint32_t stack_spill_entries(const ar_scene* s);

// Screened Poisson CG on device buffers (interleaved RGB films).
ar_status poisson_solve_device(const float* P_rgb, const float* gx_rgb, const float* gy_rgb, uint32_t w, uint32_t h,
                               float alpha, uint32_t iters, int32_t sign, float* out_rgb, double* final_residual,
                               DevBuf& scratch, hipStream_t st);

struct JobSpec;
// Samples per pixel issued per chunk (bounds in-flight path state).
uint32_t chunk_samples(const JobSpec& spec);

}  // namespace ar
