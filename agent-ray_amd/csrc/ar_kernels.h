// Launch configuration shared by the kernel translation units.
#pragma once
#include "ar_internal.h"

namespace ar {

constexpr int kTraceBlock = 128;       // 2 waves per workgroup
constexpr int kTraceLdsStack = 16;     // LDS stack entries per lane (8 B each)
constexpr int kTraceBlocksPerCU = 8;   // persistent grid: 16 waves per CU

uint32_t trace_grid(uint64_t n, int block);
uint32_t trace_grid_cfg(uint64_t n, const ar_trace_config& c);
int device_cus();

// Traversal variants compiled: (workgroup size, LDS stack entries per lane).
#define AR_TRACE_VARIANTS(X) X(64, 8) X(64, 16) X(128, 8) X(128, 16) X(256, 8)
bool trace_variant_supported(uint32_t block, uint32_t stack);

// Scene-level scratch (traversal stack spill).  acquire() orders the caller's
// stream after the previous user of the scratch (an event, no host wait) and
// grows the buffer if needed; release() records the new last use.
ar_status scratch_acquire(ar_scene* s, size_t bytes, hipStream_t st, void** out);
void scratch_release(ar_scene* s, hipStream_t st);

}  // namespace ar
