// Job plumbing shared by ar_render and the server: chunk sizing, footprint
// estimate for admission, validation, and the integrator dispatch.
#include "ar_job.h"
#include "ar_kernels.h"

namespace ar {

uint32_t chunk_samples(const JobSpec& spec) {
  const ar_render_config& c = spec.cfg;
  uint64_t npix = (uint64_t)spec.width * spec.height;
  uint64_t per_sample = npix * (c.integrator == AR_INTEGRATOR_GDPT ? 3 : 1);
  // Paths per chunk: large, because every chunk pays the tail of its last
  // bounces (a handful of long paths whose traversal is one dependent chain
  // of loads) and HBM has room: 2^25 paths of wavefront state is ~6.6 GB,
  // 2^24 BDPT samples of vertices ~17 GB at depth 10.
  // The scene's tuned target (ar_render_autotune_chunk) overrides these.
  uint64_t target = c.integrator == AR_INTEGRATOR_BDPT ? (1ull << 24) : (1ull << 25);
  if (spec.chunk_target) target = spec.chunk_target;
  uint32_t k = c.sample_chunk;
  if (k == 0) k = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c.spp, target / std::max<uint64_t>(per_sample, 1)));
  return std::min(k, c.spp);
}

ar_status validate_job(const ar_scene* s, const JobSpec& spec) {
  const ar_render_config& c = spec.cfg;
  if (!s->committed) return fail(AR_ERR_STATE, "scene not committed");
  if (spec.width == 0 || spec.height == 0 || c.spp == 0) return fail(AR_ERR_INVALID_ARG, "empty image or spp");
  if ((uint64_t)spec.width * spec.height > (1ull << 26)) return fail(AR_ERR_INVALID_ARG, "image too large");
  if (c.integrator > AR_INTEGRATOR_GDPT) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown integrator");
  if (c.schedule > AR_SCHED_MEGAKERNEL) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown schedule");
  if (c.integrator == AR_INTEGRATOR_BDPT && c.max_depth > 64) return fail(AR_ERR_INVALID_ARG, "BDPT max_depth > 64");
  if (c.integrator == AR_INTEGRATOR_GDPT && !(c.alpha > 0.0f)) return fail(AR_ERR_INVALID_ARG, "GDPT alpha <= 0");
  if (c.mis_mode > 2) return fail(AR_ERR_UNSUPPORTED_TYPE, "unknown MIS mode");
  // Paths in flight per chunk index 32-bit path ids and size the per-path
  // buffers: bound them before anything is allocated or launched.
  const uint64_t per_chunk = (uint64_t)spec.width * spec.height * chunk_samples(spec) *
                             (c.integrator == AR_INTEGRATOR_GDPT ? 3u : 1u);
  if (per_chunk > kMaxChunkPaths)
    return fail(AR_ERR_INVALID_ARG, "sample_chunk too large: width * height * sample_chunk (x3 for GDPT) above 2^31");
  return AR_OK;
}

uint64_t bdpt_conn_slots(uint32_t max_depth) {
  // one t = 0 strategy (the camera subpath ends on an emitter), then every
  // (s, t) with s, t >= 1 that can emit a shadow ray
  const uint32_t V = max_depth + 1, limit = max_depth + 2;
  uint64_t K = 1;
  for (uint32_t t = 1; t <= V; ++t)
    for (uint32_t s = 1; s <= V; ++s)
      if (s + t <= limit && !(s == 1 && t < 2)) ++K;
  return K;
}

uint64_t bdpt_conn_pass(uint32_t max_depth, uint64_t cap, uint64_t limit) {
  const uint64_t per_path = bdpt_conn_slots(max_depth) * kConnSlotBytes + sizeof(uint32_t);
  uint64_t pass = std::min<uint64_t>(cap, kConnBudget / per_path);
  if (limit) pass = std::min<uint64_t>(pass, limit);
  return std::max<uint64_t>(64, pass);
}

size_t job_footprint(const ar_scene* s, const JobSpec& spec) {
  const ar_render_config& c = spec.cfg;
  uint64_t npix = (uint64_t)spec.width * spec.height;
  uint64_t k = chunk_samples(spec);
  uint64_t paths = npix * k * (c.integrator == AR_INTEGRATOR_GDPT ? 3 : 1);
  size_t bytes = paths * (5 * 16 + 16 + 8 + 16 + 4 + 4 + 48) + npix * 48 + (64ull << 20);  // state, queues, film, spill
  if (c.integrator == AR_INTEGRATOR_GDPT) bytes += npix * 12 * 3 + npix * 12 * 9 + (64 << 10);
  if (c.integrator == AR_INTEGRATOR_BDPT) {
    bytes += paths * (2ull * (c.max_depth + 1) * 56 + 8) + npix * 24;  // vertices, counts, splat film
    if (c.schedule != AR_SCHED_MEGAKERNEL)  // wavefront connection slots of one pass
      bytes += bdpt_conn_pass(c.max_depth, paths, spec.conn_pass) * (bdpt_conn_slots(c.max_depth) * kConnSlotBytes + 4) +
               256;
  }
  bytes += npix * 12 * 4;  // output staging
  (void)s;
  return bytes;
}

std::unique_ptr<JobRun> make_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st, float* out_dev,
                                 float* aux_dev) {
  if (spec.cfg.integrator == AR_INTEGRATOR_BDPT) return make_bdpt_job(s, spec, ws, st, out_dev);
  return make_pt_job(s, spec, ws, st, out_dev, aux_dev);
}

}  // namespace ar
