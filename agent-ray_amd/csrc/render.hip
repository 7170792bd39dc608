// ar_render: Integrator::render (agent-ray src/integrators/mod.rs:19-27) for
// the three integrators, as one job run to completion on the caller's stream.
#include <chrono>

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
