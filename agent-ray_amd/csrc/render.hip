// ar_render: Integrator::render (agent-ray src/integrators/mod.rs:19-27) for
// the three integrators, as one job run to completion on the caller's stream.
#include <algorithm>
#include <chrono>
#include <memory>

#include "ar_job.h"
#include "ar_kernels.h"

using namespace ar;

extern "C" ar_status ar_render(const ar_scene* s, const ar_camera* cam, uint32_t width, uint32_t height,
                               const ar_render_config* cfg, float* out_rgb, float* out_aux, ar_render_stats* stats,
                               uint32_t flags, void* stream) {
  if (!s || !cam || !cfg || !out_rgb) return fail(AR_ERR_INVALID_ARG, "ar_render: NULL argument");
  JobSpec spec;
  spec.width = width;
  spec.height = height;
  spec.camera = *cam;
  spec.cfg = *cfg;
  ar_status e = validate_job(s, spec);
  if (e != AR_OK) return e;
  spec.chunk_target = s->chunk_target[cfg->integrator];
  auto t0 = std::chrono::steady_clock::now();
  ar_scene* ms = const_cast<ar_scene*>(s);
  hipStream_t st = pick_stream(s, stream);
  void* unused = nullptr;
  e = scratch_acquire(ms, 0, st, &unused);
  if (e != AR_OK) return e;
  if (!ms->ws) ms->ws = new Workspace();
  Workspace& ws = *ms->ws;
  size_t npix = (size_t)width * height;
  bool host = flags & AR_HOST_PTRS;
  float* out_dev = out_rgb;
  float* aux_dev = out_aux;
  if (host) {
    size_t need = npix * 3 * sizeof(float) * (out_aux ? 4 : 1);
    hipError_t he = hipSuccess;
    if (ws.out.bytes < need) {
      he = hipStreamSynchronize(st);
      if (he == hipSuccess) he = ws.out.alloc(need);
    }
    if (he != hipSuccess) {
      scratch_release(ms, st);
      return fail(AR_ERR_DEVICE, std::string("ar_render staging: ") + hipGetErrorString(he));
    }
    out_dev = ws.out.as<float>();
    aux_dev = out_aux ? out_dev + npix * 3 : nullptr;
  }
  std::unique_ptr<JobRun> job = make_job(s, spec, &ws, st, out_dev, aux_dev);
  e = job->begin();
  bool done = false;
  while (e == AR_OK && !done) e = job->step(&done);
  if (e == AR_OK) e = job->finish();
  if (e == AR_OK && host) {
    hipError_t he = hipMemcpyAsync(out_rgb, out_dev, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess && out_aux)
      he = hipMemcpyAsync(out_aux, aux_dev, npix * 9 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (he != hipSuccess) e = fail(AR_ERR_DEVICE, std::string("ar_render copy-out: ") + hipGetErrorString(he));
  }
  scratch_release(ms, st);
  if (e != AR_OK) return e;
  ar_render_stats local{};
  if (host || !(flags & AR_ASYNC) || stats) {
    AR_HIP(hipStreamSynchronize(st));
    if (stats) {
      e = job->read_stats(&local);
      if (e != AR_OK) return e;
    }
  }
  local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = local;
  return AR_OK;
}

namespace ar {
void destroy_workspace(Workspace* ws) { delete ws; }
}  // namespace ar

// Paths (BDPT: samples) per chunk of ar_render, tuned per scene and
// integrator.  A chunk pays the tail of its last bounces (few live paths, a
// chain of dependent loads each) once, so larger chunks amortise it, up to
// the point where the device is full; smaller chunks hold less path state.
// Each candidate is timed on one chunk of the job (its set-up and, for GDPT,
// the Poisson solve excluded) after a warm-up run that sizes the buffers;
// the smallest candidate within 3 % of the best rate is kept.
extern "C" ar_status ar_render_autotune_chunk(ar_scene* s, const ar_camera* cam, uint32_t width, uint32_t height,
                                              const ar_render_config* cfg, void* stream, uint64_t* best_target) {
  if (!s || !cam || !cfg) return fail(AR_ERR_INVALID_ARG, "ar_render_autotune_chunk: NULL argument");
  JobSpec spec;
  spec.width = width;
  spec.height = height;
  spec.camera = *cam;
  spec.cfg = *cfg;
  ar_status e = validate_job(s, spec);
  if (e != AR_OK) return e;
  const uint32_t integ = cfg->integrator;
  const uint64_t npix = (uint64_t)width * height;
  const uint64_t per_sample = npix * (integ == AR_INTEGRATOR_GDPT ? 3 : 1);
  const uint64_t lo = integ == AR_INTEGRATOR_BDPT ? 19 : 22;
  hipStream_t st = pick_stream(s, stream);
  DevBuf out;
  AR_HIP(out.alloc(npix * 3 * sizeof(float)));
  void* unused = nullptr;
  e = scratch_acquire(s, 0, st, &unused);
  if (e != AR_OK) return e;
  if (!s->ws) s->ws = new Workspace();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint64_t targets[5] = {};
  double rates[5] = {};
  int n = 0;
  uint32_t last_k = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    e = fail(AR_ERR_DEVICE, "hipEventCreate failed");
  for (uint64_t lg = lo; lg < lo + 5 && e == AR_OK; ++lg) {
    const uint64_t T = 1ull << lg;
    const uint32_t k = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cfg->spp, T / std::max<uint64_t>(per_sample, 1)));
    if (k == last_k) continue;
    last_k = k;
    JobSpec js = spec;
    js.cfg.spp = k;
    js.cfg.sample_chunk = k;
    float ms = 0.0f;
    for (int rep = 0; rep < 2 && e == AR_OK; ++rep) {
      std::unique_ptr<JobRun> job = make_job(s, js, s->ws, st, out.as<float>(), nullptr);
      e = job->begin();
      bool done = false;
      if (e == AR_OK && hipEventRecord(e0, st) != hipSuccess) e = fail(AR_ERR_DEVICE, "hipEventRecord failed");
      while (e == AR_OK && !done) e = job->step(&done);
      if (e == AR_OK && hipEventRecord(e1, st) != hipSuccess) e = fail(AR_ERR_DEVICE, "hipEventRecord failed");
      if (e == AR_OK) e = job->finish();
      if (e == AR_OK && hipStreamSynchronize(st) != hipSuccess) e = fail(AR_ERR_DEVICE, "stream failed");
      if (e == AR_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    }
    if (e == AR_OK) {
      targets[n] = (uint64_t)k * per_sample;
      rates[n] = (double)k * per_sample / std::max(1e-6, (double)ms);
      ++n;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  scratch_release(s, st);
  if (e != AR_OK) return e;
  double best = 0.0;
  for (int i = 0; i < n; ++i) best = std::max(best, rates[i]);
  uint64_t pick = 0;
  for (int i = 0; i < n && !pick; ++i)
    if (rates[i] >= 0.97 * best) pick = targets[i];
  s->chunk_target[integ] = pick;
  if (best_target) *best_target = pick;
  return AR_OK;
}
