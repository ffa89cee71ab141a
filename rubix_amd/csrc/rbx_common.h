// Internal helpers shared by the rubix_amd HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/rubix_amd.h"

namespace rbx {

// Thread-local last-error text, read by rbx_last_error_string().
void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Check the launch status of the preceding kernel(s) on the host side.
int check_launch(const char* what);

inline int fail(int code, const char* msg) {
  set_error("%s", msg);
  return code;
}

template <class... A>
inline int failf(int code, const char* fmt, A... args) {
  set_error(fmt, args...);
  return code;
}

constexpr int kWave = 64;

// Round a byte count up to 256 so that workspace carving keeps alignment.
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Carves consecutive aligned regions out of a caller-provided workspace.
struct Carver {
  char* base;
  size_t cap;
  size_t used = 0;
  bool ok = true;
  Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <class T>
  T* take(size_t count) {
    size_t bytes = align256(count * sizeof(T));
    if (used + bytes > cap) {
      ok = false;
      return nullptr;
    }
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

// Size-only twin of Carver.
struct Sizer {
  size_t used = 0;
  template <class T>
  void take(size_t count) { used += align256(count * sizeof(T)); }
};

__device__ inline int64_t gtid() {
  return int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
}

// Wavefront-wide sum with a fixed butterfly order (deterministic).
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Exact searchsorted(a, x, side='right') on an ascending array of length n:
// the number of entries a[k] <= x.  NaN compares as larger than everything,
// matching numpy / jnp ordering.
__device__ inline int upper_bound_f(const float* __restrict__ a, int n, float x) {
  if (x != x) return n;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

}  // namespace rbx
