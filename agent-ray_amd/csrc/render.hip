// ar_render: Integrator::render (agent-ray src/integrators/mod.rs:19-27) for
// the three integrators, with host/device buffer handling.
#include <chrono>

#include "ar_kernels.h"
#include "ar_render.h"

using namespace ar;

extern "C" ar_status ar_render(const ar_scene* s, const ar_camera* cam, uint32_t width, uint32_t height,
                               const ar_render_config* cfg, float* out_rgb, float* out_aux, ar_render_stats* stats,
                               uint32_t flags, void* stream) {
  if (!s || !cam || !cfg || !out_rgb) return fail(AR_ERR_INVALID_ARG, "ar_render: NULL argument");
  if (!s->committed) return fail(AR_ERR_STATE, "ar_render: scene not committed");
  if (width == 0 || height == 0 || cfg->spp == 0) return fail(AR_ERR_INVALID_ARG, "ar_render: empty image or spp");
  if ((uint64_t)width * height > (1ull << 26)) return fail(AR_ERR_INVALID_ARG, "ar_render: image too large");
  if (cfg->integrator > AR_INTEGRATOR_GDPT) return fail(AR_ERR_UNSUPPORTED_TYPE, "ar_render: unknown integrator");
  if (cfg->schedule > AR_SCHED_MEGAKERNEL) return fail(AR_ERR_UNSUPPORTED_TYPE, "ar_render: unknown schedule");
  auto t0 = std::chrono::steady_clock::now();
  ar_scene* ms = const_cast<ar_scene*>(s);
  hipStream_t st = pick_stream(s, stream);
  void* unused = nullptr;
  ar_status e = scratch_acquire(ms, 0, st, &unused);
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
  ar_render_stats local{};
  switch (cfg->integrator) {
    case AR_INTEGRATOR_PT: e = render_pt(s, ws, cam, width, height, cfg, out_dev, st, &local); break;
    case AR_INTEGRATOR_GDPT: e = render_gdpt(s, ws, cam, width, height, cfg, out_dev, aux_dev, st, &local); break;
    default: e = render_bdpt(s, ws, cam, width, height, cfg, out_dev, st, &local); break;
  }
  if (e == AR_OK && host) {
    hipError_t he = hipMemcpyAsync(out_rgb, out_dev, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess && out_aux)
      he = hipMemcpyAsync(out_aux, aux_dev, npix * 9 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (he != hipSuccess) e = fail(AR_ERR_DEVICE, std::string("ar_render copy-out: ") + hipGetErrorString(he));
  }
  scratch_release(ms, st);
  if (e != AR_OK) return e;
  if (host || !(flags & AR_ASYNC)) AR_HIP(hipStreamSynchronize(st));
  local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = local;
  return AR_OK;
}

namespace ar {
void destroy_workspace(Workspace* ws) { delete ws; }

}  // namespace ar
