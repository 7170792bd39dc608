// Launch configuration shared by the kernel translation units.
#pragma once
#include "ar_internal.h"

namespace ar {

constexpr int kTraceBlock = 128;       // 2 waves per workgroup
constexpr int kTraceLdsStack = 16;     // LDS stack entries per lane (8 B each)
constexpr int kTraceBlocksPerCU = 8;   // persistent grid: 16 waves per CU
constexpr uint32_t kTraceRefillMin = 16;  // idle lanes a wave waits for before taking new rays
// Traversal kernels with an 8- or 12-entry LDS stack are compiled for at
// least 6 waves per SIMD (<= 80 VGPRs): measured faster than the
// unconstrained 5 waves / 85 VGPRs; the few values the register limit moves
// to scratch sit off the node loop.  At 6 waves per SIMD a CU holds six
// 256-thread workgroups, whose 12-entry stacks (24 KiB each) fit its 160 KiB
// of LDS.  A 16-entry stack is held to 5 waves by LDS anyway.
constexpr int trace_min_waves(int lds_stack) { return lds_stack <= 12 ? 6 : 1; }

uint32_t trace_grid(uint64_t n, int block);
uint32_t trace_grid_cfg(uint64_t n, const ar_trace_config& c);
int device_cus();
// Persistent grids hand every wave a static share of the work, so blocks
// that cannot be resident at once would only run after the others: clamps
// G to (resident blocks per CU of `kernel` at `block` threads) x CUs,
// rounded down to a multiple of 8.
uint32_t clamp_resident(uint32_t G, const void* kernel, int block);

// Traversal variants compiled: (workgroup size, LDS stack entries per lane).
#define AR_TRACE_VARIANTS(X) X(64, 8) X(64, 16) X(128, 8) X(128, 16) X(256, 8) X(256, 12)
bool trace_variant_supported(uint32_t block, uint32_t stack);
// The grid ar_trace (any = false) / ar_occluded (any = true) launches for n
// rays under c; 0 for an unsupported variant.
uint32_t trace_launch_grid(uint64_t n, const ar_trace_config& c, bool any, bool near);

// Work counters of walker launches (walk_items): kWalkWorkBytes of device
// memory that are zero when a launch starts and that its last wave sets back
// to zero.  Launches that run one after another may share them.  Newly
// allocated counters are cleared by clear_walk_work, a kernel on `st` (no
// memset node if the stream is being captured).
constexpr size_t kWalkWorkBytes = 256;
hipError_t clear_walk_work(void* p, hipStream_t st);

// Scene-level scratch: the walker counters of ar_trace / ar_occluded
// (kWalkWorkBytes) followed by the traversal stack spill.  acquire() orders
// the caller's stream after the previous user of the scratch (an event, no
// host wait) and grows the buffer to `bytes` (the counters included) if
// needed; release() records the new last use.
ar_status scratch_acquire(ar_scene* s, size_t bytes, hipStream_t st, void** out);
void scratch_release(ar_scene* s, hipStream_t st);

}  // namespace ar
