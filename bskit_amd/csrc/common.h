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

// Packed bf16 pairs for v_dot2_f32_bf16 (element 0 in the low half).  The
// bits reach the instruction through a register: an empty asm hides them from
// constant folding, so no packed-operand inline constant (whose bf16
// interpretation is not the 32-bit pattern) is ever emitted; being
// side-effect free, the move is hoisted out of loops.
typedef __bf16 bsk_bf16x2_t __attribute__((ext_vector_type(2)));
BSK_DEV bsk_bf16x2_t bf16_pair_bits(uint32_t bits) {
  asm("" : "+v"(bits));
  return __builtin_bit_cast(bsk_bf16x2_t, bits);
}
constexpr uint32_t kBf16NegOneLo = 0x0000bf80u;  // (-1, 0)
constexpr uint32_t kBf16NegOneHi = 0xbf800000u;  // (0, -1)
constexpr uint32_t kBf16OneLo = 0x00003f80u;     // (1, 0)
constexpr uint32_t kBf16OneHi = 0x3f800000u;     // (0, 1)

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
