// Shared definitions for the gfx950 bispectrum kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define BSK_DEV __device__ __forceinline__

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
BSK_DEV cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
BSK_DEV cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
BSK_DEV cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
BSK_DEV cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }
template <typename T>
BSK_DEV cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
BSK_DEV cplx<T> mul_mi(cplx<T> a) {
  if (INV) return {-a.y, a.x};
  return {a.y, -a.x};
}

// Status codes, dtypes, formats and kinds come from the public header.
#include "../../include/bskhip.h"

void bsk_set_error(const char* fmt, ...);

#define BSK_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      bsk_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),      \
                    __FILE__, __LINE__);                                        \
      return BSK_EHIP;                                                          \
    }                                                                           \
  } while (0)

#define BSK_LAUNCH_CHECK()                                                      \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess) {                                                     \
      bsk_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),  \
                    __FILE__, __LINE__);                                        \
      return BSK_EHIP;                                                          \
    }                                                                           \
  } while (0)
