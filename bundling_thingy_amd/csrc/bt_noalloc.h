// Host side of the allocation-free entry point (bt_noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/bundling.h"

// Hash-table slots of the scratch: a power of two >= 4 slots per path.
#define BT_NOALLOC_MAX_SLOTS (4u * BT_NOALLOC_MAX_PATHS)

namespace bt {
// Resolves the scratch array's device address on `device` (current) once;
// throws bt::Error when HIP cannot.
void noalloc_setup(int device);
// Endpoint dedup of one job's current records `rec` (sorted order, n
// samples, n_paths paths) through the scratch of `device`, on stream `s`;
// blocks until the indices are on the host. *count receives the number
// found and min(capacity, count) indices are written.
void endpoints_noalloc(hipStream_t s, int device, const float* rec, uint32_t n, uint32_t n_paths,
                       uint32_t* indices, size_t capacity, uint32_t* count);
}  // namespace bt
