// Small HIP helpers shared by the builder, the sorter and the traversal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdexcept>
#include <string>
#include <utility>

namespace rtg {

struct hip_error : std::runtime_error {
  hipError_t code;
  hip_error(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

#define RTG_HIP(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw ::rtg::hip_error(e_, std::string(#call) + ": " + hipGetErrorString(e_));            \
  } while (0)

__host__ __device__ inline uint32_t fbits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
__host__ __device__ inline float bitsf(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// Monotone float <-> uint mapping for atomicMin/atomicMax on floats.
__host__ __device__ inline uint32_t f2o(float f) {
  uint32_t u = fbits(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t u) {
  return bitsf((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// reference halfArea: fma(d.x, d.y + d.z, d.y * d.z) (common/math/vec3fa.h:373)
__host__ __device__ inline float half_area(float dx, float dy, float dz) {
  return __builtin_fmaf(dx, dy + dz, dy * dz);
}

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    RTG_HIP(hipMalloc(&ptr, n < 256 ? 256 : n));
    bytes = n < 256 ? 256 : n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return (T*)ptr; }
  void swap(DevBuf& o) {
    std::swap(ptr, o.ptr);
    std::swap(bytes, o.bytes);
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Grow-only pinned host buffer (the source or target of an asynchronous
// copy that must not wait for the host).
struct PinnedBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    RTG_HIP(hipHostMalloc(&ptr, n, hipHostMallocDefault));
    bytes = n;
  }
  template <class T> T* as() const { return (T*)ptr; }
  ~PinnedBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
};

}  // namespace rtg
