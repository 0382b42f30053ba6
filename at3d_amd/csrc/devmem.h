// Device memory of solver states, kept by size for reuse.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <mutex>

namespace at3d {

// Serialises hipMalloc / hipFree and HIP-graph capture across host threads:
// a capture in progress on one state's stream must not overlap a call that
// can synchronise the device on another thread.
std::mutex& device_mutex();

// A serving loop creates and releases states of a few shapes over and over;
// hipMalloc and hipFree of gigabytes each time would stall it.  Freed blocks
// are kept by exact size and handed to the next request of that size; they
// go back to the runtime when an allocation fails for lack of memory
// (then it is retried once) or on dev_trim().  Blocks handed back must no
// longer be used by any queued work.
hipError_t dev_alloc(void** ptr, size_t bytes);
void dev_free(void* ptr, size_t bytes);
void dev_trim();
size_t dev_cached_bytes();

}  // namespace at3d
