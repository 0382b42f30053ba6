// Size-keyed cache of device blocks (see devmem.h).
#include "devmem.h"

#include <unordered_map>
#include <vector>

namespace at3d {

std::mutex& device_mutex() {
  static std::mutex m;
  return m;
}

namespace {
std::mutex cache_mutex;
std::unordered_multimap<size_t, void*> cache;
size_t cached = 0;

void trim_locked() {
  std::lock_guard<std::mutex> dl(device_mutex());
  for (auto& kv : cache) (void)hipFree(kv.second);
  cache.clear();
  cached = 0;
}
}  // namespace

hipError_t dev_alloc(void** ptr, size_t bytes) {
  if (bytes == 0) bytes = 1;
  {
    std::lock_guard<std::mutex> l(cache_mutex);
    auto it = cache.find(bytes);
    if (it != cache.end()) {
      *ptr = it->second;
      cache.erase(it);
      cached -= bytes;
      return hipSuccess;
    }
  }
  hipError_t e;
  {
    std::lock_guard<std::mutex> dl(device_mutex());
    e = hipMalloc(ptr, bytes);
  }
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    std::lock_guard<std::mutex> l(cache_mutex);
    trim_locked();
    std::lock_guard<std::mutex> dl(device_mutex());
    e = hipMalloc(ptr, bytes);
  }
  return e;
}

void dev_free(void* ptr, size_t bytes) {
  if (!ptr) return;
  if (bytes == 0) bytes = 1;
  std::lock_guard<std::mutex> l(cache_mutex);
  cache.emplace(bytes, ptr);
  cached += bytes;
}

void dev_trim() {
  std::lock_guard<std::mutex> l(cache_mutex);
  trim_locked();
}

size_t dev_cached_bytes() {
  std::lock_guard<std::mutex> l(cache_mutex);
  return cached;
}

}  // namespace at3d
