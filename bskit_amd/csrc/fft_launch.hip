// Host launchers for the FFT passes: runtime (N, variant, precision) ->
// compile-time kernel instance.
#include "fft_passes.h"
#include "launch.h"

namespace {

constexpr int kW = 256;

// Pencils per workgroup for the strided x/y passes: 16 consecutive kz values
// (128 B in FP32) per grid point; FP64 and the longest axis use 8 to stay
// inside the LDS budget.
template <typename T, int N>
constexpr int pxy() {
  return (sizeof(T) == 8 || N >= 768) ? 8 : 16;
}
// Rows per workgroup for the contiguous z passes: about 4096 complex values.
template <typename T, int NH>
constexpr int pz() {
  int p = (sizeof(T) == 8 ? 2048 : 4096) / NH;
  return p < 1 ? 1 : (p > 64 ? 64 : p);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

#define BSK_SIZES(X) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)

template <typename T, int N, int V>
void fwd_z_n(const T* in, cplx<T>* out, const cplx<T>* tw, long long nrows, int MZ,
             hipStream_t s) {
  constexpr int NH = N / 2, P = pz<T, NH>();
  hipLaunchKernelGGL((k_fwd_z<T, NH, V, P, kW>), dim3(cdiv(nrows, P)), dim3(kW), 0, s, in, out,
                     tw, nrows, MZ);
}
template <typename T, int N, int V>
void fwd_y_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_fwd_y<T, N, V, P, kW>), dim3(N, cdiv(g.MZ, P)), dim3(kW), 0, s, in, out,
                     tw, g);
}
template <typename T, int N, int V>
void fwd_x_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, T scale, const T* comp,
             hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_fwd_x<T, N, V, P, kW>), dim3(g.MY, cdiv(g.MZ, P)), dim3(kW), 0, s, in,
                     out, tw, g, scale, comp);
}
template <typename T, int N, int V>
void inv_x_n(const cplx<T>* dk, cplx<T>* out, const cplx<T>* tw, Geom g, const int* lo,
             const int* hi, int shell0, int nsh, int kind, T kf, hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_inv_x_mask<T, N, V, P, kW>), dim3(g.MY, cdiv(g.MZ, P), nsh), dim3(kW),
                     0, s, dk, out, tw, g, lo, hi, shell0, kind, kf);
}
template <typename T, int N, int V>
void inv_y_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, int x0, int xs, int nsh,
             hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_inv_y<T, N, V, P, kW>), dim3(xs, cdiv(g.MZ, P), nsh), dim3(kW), 0, s, in,
                     out, tw, g, x0, xs);
}
template <typename T, int N, int V>
void inv_z_n(const cplx<T>* in, void* out, void* out_lo, const cplx<T>* tw, int MZ,
             long long nrows, long long cell0, long long stride, int fmt, int nsh, hipStream_t s) {
  constexpr int NH = N / 2, P = pz<T, NH>();
  hipLaunchKernelGGL((k_inv_z<T, NH, V, P, kW>), dim3(cdiv(nrows, P), nsh), dim3(kW), 0, s, in,
                     out, out_lo, tw, MZ, nrows, cell0, stride, fmt);
}

}  // namespace

bool bsk_fft_size_supported(int N) {
  switch (N) {
#define X(n) case n:
    BSK_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

template <typename T>
int bsk_fwd_z(int N, int V, const T* in, cplx<T>* out, const cplx<T>* tw, long long nrows, int MZ,
              hipStream_t s) {
  switch (N) {
#define X(n)                                                                      \
  case n:                                                                         \
    if constexpr (sizeof(T) == 4) { if (V == 1) { fwd_z_n<T, n, 1>(in, out, tw, nrows, MZ, s); break; } } \
    fwd_z_n<T, n, 0>(in, out, tw, nrows, MZ, s);                             \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_fwd_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g,
              hipStream_t s) {
  switch (N) {
#define X(n)                                                                      \
  case n:                                                                         \
    if constexpr (sizeof(T) == 4) { if (V == 1) { fwd_y_n<T, n, 1>(in, out, tw, g, s); break; } } \
    fwd_y_n<T, n, 0>(in, out, tw, g, s);                                     \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_fwd_x(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, T scale,
              const T* comp, hipStream_t s) {
  switch (N) {
#define X(n)                                                                      \
  case n:                                                                         \
    if constexpr (sizeof(T) == 4) { if (V == 1) { fwd_x_n<T, n, 1>(in, out, tw, g, scale, comp, s); break; } } \
    fwd_x_n<T, n, 0>(in, out, tw, g, scale, comp, s);                        \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_inv_x(int N, int V, const cplx<T>* dk, cplx<T>* out, const cplx<T>* tw, Geom g,
              const int* lo, const int* hi, int shell0, int nsh, int kind, T kf, hipStream_t s) {
  switch (N) {
#define X(n)                                                                             \
  case n:                                                                                \
    if constexpr (sizeof(T) == 4) { if (V == 1) { inv_x_n<T, n, 1>(dk, out, tw, g, lo, hi, shell0, nsh, kind, kf, s); break; } } \
    inv_x_n<T, n, 0>(dk, out, tw, g, lo, hi, shell0, nsh, kind, kf, s);             \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_inv_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, int x0,
              int xs, int nsh, hipStream_t s) {
  switch (N) {
#define X(n)                                                                      \
  case n:                                                                         \
    if constexpr (sizeof(T) == 4) { if (V == 1) { inv_y_n<T, n, 1>(in, out, tw, g, x0, xs, nsh, s); break; } } \
    inv_y_n<T, n, 0>(in, out, tw, g, x0, xs, nsh, s);                        \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_inv_z(int N, int V, const cplx<T>* in, void* out, void* out_lo, const cplx<T>* tw, int MZ,
              long long nrows, long long cell0, long long stride, int fmt, int nsh,
              hipStream_t s) {
  switch (N) {
#define X(n)                                                                             \
  case n:                                                                                \
    if constexpr (sizeof(T) == 4) { if (V == 1) { inv_z_n<T, n, 1>(in, out, out_lo, tw, MZ, nrows, cell0, stride, fmt, nsh, s); break; } } \
    inv_z_n<T, n, 0>(in, out, out_lo, tw, MZ, nrows, cell0, stride, fmt, nsh, s);   \
    break;
    BSK_SIZES(X)
#undef X
    default:
      bsk_set_error("unsupported FFT size %d", N);
      return BSK_EUNSUPPORTED;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

#define BSK_INST(T)                                                                        \
  template int bsk_fwd_z<T>(int, int, const T*, cplx<T>*, const cplx<T>*, long long, int,    \
                            hipStream_t);                                                  \
  template int bsk_fwd_y<T>(int, int, const cplx<T>*, cplx<T>*, const cplx<T>*, Geom,        \
                            hipStream_t);                                                  \
  template int bsk_fwd_x<T>(int, int, const cplx<T>*, cplx<T>*, const cplx<T>*, Geom, T,     \
                            const T*, hipStream_t);                                        \
  template int bsk_inv_x<T>(int, int, const cplx<T>*, cplx<T>*, const cplx<T>*, Geom,        \
                            const int*, const int*, int, int, int, T, hipStream_t);        \
  template int bsk_inv_y<T>(int, int, const cplx<T>*, cplx<T>*, const cplx<T>*, Geom, int,   \
                            int, int, hipStream_t);                                        \
  template int bsk_inv_z<T>(int, int, const cplx<T>*, void*, void*, const cplx<T>*, int,     \
                            long long, long long, long long, int, int, hipStream_t);
BSK_INST(float)
BSK_INST(double)
