// A render in flight, issued chunk by chunk on one stream.  ar_render runs
// one job to completion; the server interleaves the chunks of many jobs on
// their own streams (continuous batching).  A job's film depends only on its
// own description: chunking and interleaving never change a bit.
#pragma once
#include <memory>

#include "ar_render.h"

namespace ar {

struct JobSpec {
  uint32_t width = 0, height = 0;
  ar_camera camera{};
  ar_render_config cfg{};
  uint64_t chunk_target = 0;  // paths (BDPT: samples) per chunk when cfg.sample_chunk is 0; 0 = default
  uint64_t conn_pass = 0;     // BDPT paths per connection pass at most; 0 = as many as the pass budget holds
};

class JobRun {
 public:
  virtual ~JobRun() {}
  // Reserve device memory and enqueue the set-up work.
  virtual ar_status begin() = 0;
  // Enqueue the next piece of work (a chunk, or a phase of one); *done is
  // set once every sample has been issued.
  virtual ar_status step(bool* done) = 0;
  // Share of a chunk's work the last step() issued.
  virtual double last_step_share() const { return 1.0; }
  // Enqueue the resolve (and the Poisson solve for GDPT) into out/aux.
  virtual ar_status finish() = 0;
  // After the stream has drained: rays/paths counters.
  virtual ar_status read_stats(ar_render_stats* st) = 0;

  const ar_scene* scene = nullptr;
  JobSpec spec;
  Workspace* ws = nullptr;
  hipStream_t stream = nullptr;
  float* out_dev = nullptr;  // width*height*3 floats
  float* aux_dev = nullptr;  // GDPT: primal, dx, dy films (optional)
};

std::unique_ptr<JobRun> make_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st, float* out_dev,
                                 float* aux_dev);
std::unique_ptr<JobRun> make_pt_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st,
                                    float* out_dev, float* aux_dev);
std::unique_ptr<JobRun> make_bdpt_job(const ar_scene* s, const JobSpec& spec, Workspace* ws, hipStream_t st,
                                      float* out_dev);
// Device bytes a job of this shape needs (admission control).
size_t job_footprint(const ar_scene* s, const JobSpec& spec);
// Largest chunk (paths in flight) a job may use: path ids are 32-bit.
constexpr uint64_t kMaxChunkPaths = 1ull << 31;
// BDPT wavefront connection: bytes of connection slots per pass, bytes per
// slot (ray origin, direction, contribution, occlusion flag, queue entry),
// slots per path at a depth, and paths per pass.
constexpr uint64_t kConnBudget = 12ull << 30;
constexpr uint64_t kConnSlotBytes = 3 * 16 + 1 + 4;
uint64_t bdpt_conn_slots(uint32_t max_depth);
uint64_t bdpt_conn_pass(uint32_t max_depth, uint64_t cap, uint64_t limit = 0);
ar_status validate_job(const ar_scene* s, const JobSpec& spec);

}  // namespace ar
