// Serving pool: path-tracer and gradient-domain jobs rendered together in
// one set of wavefront queues that are refilled every bounce (host side).
// No counterpart in the reference, whose PathTracer::render
// (path_tracer.rs:159-191) renders one image per call.
#pragma once
#include <memory>

#include "ar_job.h"

namespace ar {

constexpr uint32_t kPoolMaxJobs = 64;           // jobs sharing the pool at once
constexpr uint32_t kPoolDefaultPaths = 1u << 22;  // paths in flight
constexpr uint32_t kPoolWindowPaths = 1u << 24;   // a job's paths per accumulation window

class PathPool {
 public:
  virtual ~PathPool() {}
  // The pool's device memory: state of `capacity` paths in flight.  While
  // a job drains its last paths, the others keep at most `fill` alive.
  static ar_status create(const ar_scene* s, uint32_t capacity, uint32_t fill, std::unique_ptr<PathPool>* out);
  static size_t footprint(uint32_t capacity);
  // Device bytes a job needs besides the pool (window buffers, film, output).
  static size_t job_bytes(const JobSpec& spec);

  // Adds a PT or GDPT job.  Its buffers come from `ws` (film accumulator,
  // two window buffers, GDPT films, Poisson scratch); its finished film is
  // resolved into out_dev on slot_stream and copied to `pinned`
  // (width*height*3 floats).  `order` breaks priority ties (lower first).
  virtual ar_status add(const JobSpec& spec, Workspace* ws, float* out_dev, float* pinned, hipStream_t slot_stream,
                        uint64_t order, int* handle) = 0;
  // Stops generating the job's paths; it drains and ends as cancelled.
  virtual void cancel(int handle) = 0;
  // Retires finished iterations, accumulates completed windows, starts the
  // next windows and finished jobs' resolves, then issues iterations while
  // fewer than the in-flight limit are on the device, unless `issue` is
  // false.  Sets *progress when anything changed.
  virtual ar_status pump(bool issue, bool* progress) = 0;
  // Job state: 0 rendering, 1 film resolved and copied (after `done`
  // completes), 2 cancelled and drained.
  virtual int state(int handle) = 0;
  // Event recorded on the job's slot stream after its film copy (state 1).
  virtual hipEvent_t done_event(int handle) = 0;
  virtual void stats(int handle, ar_render_stats* st) = 0;
  // Frees the handle (after state() became 1 and done_event completed, or 2).
  virtual void release(int handle) = 0;
  // Paths not yet generated, over every job that still generates, and of one job.
  virtual uint64_t pending_paths() = 0;
  virtual uint64_t remaining_paths(int handle) = 0;
  virtual bool idle() = 0;       // no job in the pool
  virtual bool full() = 0;       // every handle taken
  virtual uint32_t in_flight() = 0;  // iterations issued and not yet retired
  virtual hipStream_t stream() = 0;
};

}  // namespace ar
