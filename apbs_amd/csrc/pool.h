// pool.h -- caching device-memory, pinned-scalar and stream pools.
//
// hipFree, hipHostFree and hipStreamDestroy can wait for the whole device,
// which would stall every other calculation running on its own stream.  The
// pools hand freed blocks to later allocations instead.  A block is returned
// only after the stream that used it is idle (every caller synchronises its
// stream before freeing), so reuse from another stream is safe.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace pmgpool {

void* dev_alloc(size_t bytes);
void dev_free(void* p);
double* pinned_scalars();            // 32 pinned doubles
void pinned_release(double* p);
hipStream_t stream_acquire();        // non-blocking, default priority
void stream_release(hipStream_t s);
// bytes in use, high-water mark of bytes in use, bytes cached
void stats(size_t* inuse, size_t* peak, size_t* cached);
void reset_peak();
// return every cached device block to the runtime (device must be idle)
void trim();

}  // namespace pmgpool
