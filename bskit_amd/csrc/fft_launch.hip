// Host launchers for the FFT passes: runtime (N, variant, precision) ->
// compile-time kernel instance.  The file is compiled once per pass
// (BSK_PASS = 0..5: forward z, y, x; inverse x, y, z) so the instances build
// in parallel.
#include "fft_passes.h"
#include "launch.h"

#ifndef BSK_PASS
#error "compile with -DBSK_PASS=0..5"
#endif

namespace {

constexpr int kW = 256;

template <typename T, int N>
constexpr int pxy() {
  return bsk_pxy(sizeof(T) == 8, N);
}
// The forward y/x passes from 512 up: 8 (more workgroups per CU)
template <typename T, int N>
constexpr int pfwd() {
  return N >= 512 ? 8 : pxy<T, N>();
}
// Rows per workgroup for the contiguous z passes: the power of two nearest
// below about 4096 (FP64: 2048) complex values per workgroup, so the
// butterfly -> (row, column) split of the LDS stages is a mask and a shift.
template <typename T, int NH>
constexpr int pz() {
  int p = (sizeof(T) == 8 ? 2048 : 4096) / NH;
  int q = 1;
  while (2 * q <= p && q < 64) q *= 2;
  return q;
}

[[maybe_unused]] inline unsigned cdiv(long long a, long long b) {
  return (unsigned)((a + b - 1) / b);
}

}  // namespace

#define BSK_SWITCH(T, N, V, CALL)                                                    \
  switch (N) {                                                                       \
    BSK_FFT_SIZES(CALL)                                                              \
    default:                                                                         \
      bsk_set_error("unsupported FFT size %d", N);                                   \
      return BSK_EUNSUPPORTED;                                                       \
  }                                                                                  \
  BSK_LAUNCH_CHECK();                                                                \
  return BSK_OK;

// one case: variant 1 exists for FP32 only
#define BSK_VCASE(n, EXPR0, EXPR1)                          \
  case n:                                                   \
    if constexpr (sizeof(T) == 4) {                         \
      if (V == 1) { EXPR1; break; }                         \
    }                                                       \
    EXPR0;                                                  \
    break;

#if BSK_PASS == 0
namespace {
template <typename T, int N, int V>
void fwd_z_n(const T* in, cplx<T>* out, const cplx<T>* tw, long long nrows, int MZ, int MZP,
             hipStream_t s) {
  constexpr int NH = N / 2, P = pz<T, NH>();
  hipLaunchKernelGGL((k_fwd_z<T, NH, V, P, kW>), dim3(cdiv(nrows, P)), dim3(kW), 0, s, in, out,
                     tw, nrows, MZ, MZP);
}
}  // namespace
template <typename T>
int bsk_fwd_z(int N, int V, const T* in, cplx<T>* out, const cplx<T>* tw, long long nrows, int MZ,
              int MZP, hipStream_t s) {
#define C(n) BSK_VCASE(n, (fwd_z_n<T, n, 0>(in, out, tw, nrows, MZ, MZP, s)), \
                       (fwd_z_n<T, n, 1>(in, out, tw, nrows, MZ, MZP, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_fwd_z<float>(int, int, const float*, cplx<float>*, const cplx<float>*, long long,
                              int, int, hipStream_t);
template int bsk_fwd_z<double>(int, int, const double*, cplx<double>*, const cplx<double>*,
                               long long, int, int, hipStream_t);
#endif

#if BSK_PASS == 1
namespace {
template <typename T, int N, int V>
void fwd_y_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, hipStream_t s) {
  constexpr int P = pfwd<T, N>();
  hipLaunchKernelGGL((k_fwd_y<T, N, V, P, kW>), dim3(cdiv(g.MZ, P), N), dim3(kW), 0, s, in, out,
                     tw, g);
}
}  // namespace
template <typename T>
int bsk_fwd_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g,
              hipStream_t s) {
#define C(n) BSK_VCASE(n, (fwd_y_n<T, n, 0>(in, out, tw, g, s)), (fwd_y_n<T, n, 1>(in, out, tw, g, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_fwd_y<float>(int, int, const cplx<float>*, cplx<float>*, const cplx<float>*, Geom,
                              hipStream_t);
template int bsk_fwd_y<double>(int, int, const cplx<double>*, cplx<double>*, const cplx<double>*,
                               Geom, hipStream_t);
#endif

#if BSK_PASS == 2
namespace {
template <typename T, int N, int V>
void fwd_x_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, T scale, const T* comp,
             hipStream_t s) {
  constexpr int P = pfwd<T, N>();
  hipLaunchKernelGGL((k_fwd_x<T, N, V, P, kW>), dim3(cdiv(g.MZ, P), g.MY), dim3(kW), 0, s, in,
                     out, tw, g, scale, comp);
}
}  // namespace
template <typename T>
int bsk_fwd_x(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, T scale,
              const T* comp, hipStream_t s) {
#define C(n) BSK_VCASE(n, (fwd_x_n<T, n, 0>(in, out, tw, g, scale, comp, s)), \
                       (fwd_x_n<T, n, 1>(in, out, tw, g, scale, comp, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_fwd_x<float>(int, int, const cplx<float>*, cplx<float>*, const cplx<float>*, Geom,
                              float, const float*, hipStream_t);
template int bsk_fwd_x<double>(int, int, const cplx<double>*, cplx<double>*, const cplx<double>*,
                               Geom, double, const double*, hipStream_t);
#endif

#if BSK_PASS == 3
namespace {
template <typename T, int N, int V>
void inv_x_n(const cplx<T>* dk, cplx<T>* out, const cplx<T>* tw, Geom g, int shell0,
             const int4* blocks, int nblocks, int kind, T kf, T kp, hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_inv_x_mask<T, N, V, P, kW>), dim3(nblocks), dim3(kW), 0, s, dk, out, tw,
                     g, shell0, blocks, kind, kf, kp);
}
}  // namespace
template <typename T>
int bsk_inv_x(int N, int V, const cplx<T>* dk, cplx<T>* out, const cplx<T>* tw, Geom g,
              int shell0, const int4* blocks, int nblocks, int kind, T kf, T kp, hipStream_t s) {
  if (nblocks <= 0) return BSK_OK;  // no shell holds a mode
#define C(n) BSK_VCASE(n, (inv_x_n<T, n, 0>(dk, out, tw, g, shell0, blocks, nblocks, kind, kf, kp, s)), \
                       (inv_x_n<T, n, 1>(dk, out, tw, g, shell0, blocks, nblocks, kind, kf, kp, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_inv_x<float>(int, int, const cplx<float>*, cplx<float>*, const cplx<float>*, Geom,
                              int, const int4*, int, int, float, float, hipStream_t);
template int bsk_inv_x<double>(int, int, const cplx<double>*, cplx<double>*, const cplx<double>*,
                               Geom, int, const int4*, int, int, double, double, hipStream_t);
#endif

#if BSK_PASS == 4
namespace {
template <typename T, int N, int V>
void inv_y_n(const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, int x0, int xs,
             const int4* blocks, int nblocks, int shell0, hipStream_t s) {
  constexpr int P = pxy<T, N>();
  hipLaunchKernelGGL((k_inv_y<T, N, V, P, kW>), dim3(nblocks, xs), dim3(kW), 0, s, in, out, tw, g,
                     x0, xs, blocks, shell0);
}
}  // namespace
template <typename T>
int bsk_inv_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, int x0,
              int xs, const int4* blocks, int nblocks, int shell0, hipStream_t s) {
  if (nblocks <= 0 || xs <= 0) return BSK_OK;  // no shell holds a mode
#define C(n) BSK_VCASE(n, (inv_y_n<T, n, 0>(in, out, tw, g, x0, xs, blocks, nblocks, shell0, s)), \
                       (inv_y_n<T, n, 1>(in, out, tw, g, x0, xs, blocks, nblocks, shell0, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_inv_y<float>(int, int, const cplx<float>*, cplx<float>*, const cplx<float>*, Geom,
                              int, int, const int4*, int, int, hipStream_t);
template int bsk_inv_y<double>(int, int, const cplx<double>*, cplx<double>*, const cplx<double>*,
                               Geom, int, int, const int4*, int, int, hipStream_t);
#endif

#if BSK_PASS == 5
namespace {
template <typename T, int N, int V>
void inv_z_n(const cplx<T>* in, void* out, void* out_lo, const cplx<T>* tw, int MZ, int MZP,
             long long nrows, long long cell0, long long stride, int fmt, int nsh,
             const int* n2hi, int shell0, int lv, const unsigned char* parts, hipStream_t s) {
  constexpr int NH = N / 2, P = pz<T, NH>();
  hipLaunchKernelGGL((k_inv_z<T, NH, V, P, kW>), dim3(cdiv(nrows, P), nsh), dim3(kW), 0, s, in,
                     out, out_lo, tw, MZ, MZP, nrows, cell0, stride, fmt, n2hi, shell0, lv, parts);
}
}  // namespace
template <typename T>
int bsk_inv_z(int N, int V, const cplx<T>* in, void* out, void* out_lo, const cplx<T>* tw, int MZ,
              int MZP, long long nrows, long long cell0, long long stride, int fmt, int nsh,
              const int* n2hi, int shell0, int lv, const unsigned char* parts, hipStream_t s) {
#define C(n) BSK_VCASE(n, (inv_z_n<T, n, 0>(in, out, out_lo, tw, MZ, MZP, nrows, cell0, stride, fmt, nsh, n2hi, shell0, lv, parts, s)), \
                       (inv_z_n<T, n, 1>(in, out, out_lo, tw, MZ, MZP, nrows, cell0, stride, fmt, nsh, n2hi, shell0, lv, parts, s)))
  BSK_SWITCH(T, N, V, C)
#undef C
}
template int bsk_inv_z<float>(int, int, const cplx<float>*, void*, void*, const cplx<float>*, int,
                              int, long long, long long, long long, int, int, const int*, int,
                              int, const unsigned char*, hipStream_t);
template int bsk_inv_z<double>(int, int, const cplx<double>*, void*, void*, const cplx<double>*,
                               int, int, long long, long long, long long, int, int, const int*, int,
                               int, const unsigned char*, hipStream_t);
#endif
