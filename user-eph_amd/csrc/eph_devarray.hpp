// eph_devarray.hpp -- owning device buffer used by the host-side state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace eph {

// Owning device allocation, grown on demand (never shrunk).
template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // ensure room for `want` elements; contents are not preserved
  hipError_t reserve(size_t want) {
    if (want <= n && p) return hipSuccess;
    release();
    size_t m = want < 64 ? 64 : want;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), m * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = m;
    // The clearing runs on the null stream, which is not ordered with the
    // library's non-blocking streams: it must be finished before the caller
    // enqueues writes of its own to the buffer, or it can overwrite them.
    e = hipMemset(p, 0, m * sizeof(T));
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(nullptr);
  }
};

}  // namespace eph
