// Triple-product reductions sum_x A_i(x) B_j(x) C_k(x) over the N^3 cells.
//
// Three device paths:
//  * plain : one sweep over the cells per triangle, reading three shells,
//            FP32 (or FP64) products and per-thread accumulation, FP64
//            workgroup sums -- the structure of the reference's
//            np.sum(d_i*d_j*d_k) loop.
//  * tiled : cells staged through LDS in chunks, each thread owns a few
//            triangles; used in FP64 for the grid-info sums (N_tri and the
//            three k-mean numerators share their pair products).
//  * mfma  : split-BF16 MFMA GEMMs, in mfma.hip.
// All paths write per-workgroup partials that a second kernel reduces in a
// fixed order in FP64, so a result does not depend on what else runs.
#include "launch.h"

namespace {

constexpr int kW = 256;

template <typename T>
struct Vec;
template <>
struct Vec<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct Vec<double> {
  using type = double2;
  static constexpr int n = 2;
};

BSK_DEV float vget(const float4& v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
BSK_DEV double vget(const double2& v, int e) { return e == 0 ? v.x : v.y; }

BSK_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Fixed-order workgroup sum of one double per thread (W = 256).
BSK_DEV double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ plain
template <typename T>
__global__ __launch_bounds__(kW) void k_triple_plain(const T* __restrict__ a, const T* __restrict__ b,
                                                     const T* __restrict__ c, long long sa,
                                                     long long sb, long long sc, long long ncell,
                                                     const int* __restrict__ tri,
                                                     double* __restrict__ partial, int nchunk) {
  using V = typename Vec<T>::type;
  constexpr int VN = Vec<T>::n;
  __shared__ double sh[4];
  const int t = blockIdx.y, chunk = blockIdx.x;
  const int i = tri[3 * t], j = tri[3 * t + 1], k = tri[3 * t + 2];
  const long long nv = ncell / VN;
  const long long per = (nv + nchunk - 1) / nchunk;
  const long long v0 = (long long)chunk * per;
  const long long v1 = v0 + per < nv ? v0 + per : nv;
  const V* pa = reinterpret_cast<const V*>(a + i * sa);
  const V* pb = reinterpret_cast<const V*>(b + j * sb);
  const V* pc = reinterpret_cast<const V*>(c + k * sc);
  T acc = 0;
  for (long long v = v0 + threadIdx.x; v < v1; v += kW) {
    const V x = pa[v], y = pb[v], z = pc[v];
#pragma unroll
    for (int e = 0; e < VN; ++e) acc += vget(x, e) * vget(y, e) * vget(z, e);
  }
  const double s = block_sum((double)acc, sh);
  if (threadIdx.x == 0) partial[(long long)t * nchunk + chunk] = s;
}

// One 64-thread group per triangle; fixed summation order.
__global__ __launch_bounds__(64) void k_reduce_rows(const double* __restrict__ partial, int ncol,
                                                    double scale, double* __restrict__ out) {
  const long long row = blockIdx.x;
  double s = 0.0;
  for (int c = threadIdx.x; c < ncol; c += 64) s += partial[row * ncol + c];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[row] = s * scale;
}

// ------------------------------------------------------------------ tiled
// grid (nchunk, ceil(ntri / (kW*TPT))).  NSUM == 1: fields at tri[3t+0..2]
// (global shell indices).  NSUM == 4: tri holds unit-shell indices (i,j,k);
// the |k| shell of bin b is field nb + b where nb = nfield / 2.
template <typename T, int NSUM, int TPT, int CC>
__global__ __launch_bounds__(kW) void k_triple_tiled(const T* __restrict__ f, long long stride,
                                                     int nfield, long long ncell, int ntri,
                                                     const int* __restrict__ tri,
                                                     double* __restrict__ partial, int nchunk) {
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  constexpr int LDC = CC + 1;  // odd row length: distinct banks per field
  const int chunk = blockIdx.x;
  const int tbase = blockIdx.y * (kW * TPT);
  const int nb = nfield / 2;
  int fi[TPT][3];
  bool live[TPT];
#pragma unroll
  for (int q = 0; q < TPT; ++q) {
    const int t = tbase + q * kW + threadIdx.x;
    live[q] = t < ntri;
    const int tt = live[q] ? t : 0;
    fi[q][0] = tri[3 * tt];
    fi[q][1] = tri[3 * tt + 1];
    fi[q][2] = tri[3 * tt + 2];
  }
  double acc[TPT][NSUM];
#pragma unroll
  for (int q = 0; q < TPT; ++q)
#pragma unroll
    for (int s = 0; s < NSUM; ++s) acc[q][s] = 0.0;

  const long long ntile = ncell / CC;
  const long long per = (ntile + nchunk - 1) / nchunk;
  const long long t0 = (long long)chunk * per;
  const long long t1 = t0 + per < ntile ? t0 + per : ntile;
  for (long long tile = t0; tile < t1; ++tile) {
    __syncthreads();
    const long long cbase = tile * CC;
    for (int e = threadIdx.x; e < nfield * CC; e += kW) {
      const int fld = e / CC, cc = e - fld * CC;
      sm[fld * LDC + cc] = f[fld * stride + cbase + cc];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      if (!live[q]) continue;
      const T* ri = sm + fi[q][0] * LDC;
      const T* rj = sm + fi[q][1] * LDC;
      const T* rk = sm + fi[q][2] * LDC;
      if (NSUM == 1) {
        T s = 0;
        for (int cc = 0; cc < CC; ++cc) s += ri[cc] * rj[cc] * rk[cc];
        acc[q][0] += (double)s;
      } else {
        const T* ki = sm + (nb + fi[q][0]) * LDC;
        const T* kj = sm + (nb + fi[q][1]) * LDC;
        const T* kk = sm + (nb + fi[q][2]) * LDC;
        T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int cc = 0; cc < CC; ++cc) {
          const T ui = ri[cc], uj = rj[cc], uk = rk[cc];
          const T pjk = uj * uk;
          s0 += ui * pjk;
          s1 += ki[cc] * pjk;
          s2 += kj[cc] * (ui * uk);
          s3 += kk[cc] * (ui * uj);
        }
        acc[q][0] += (double)s0;
        acc[q][1] += (double)s1;
        acc[q][2] += (double)s2;
        acc[q][3] += (double)s3;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPT; ++q) {
    const int t = tbase + q * kW + threadIdx.x;
    if (!live[q]) continue;
#pragma unroll
    for (int s = 0; s < NSUM; ++s)
      partial[((long long)t * NSUM + s) * nchunk + chunk] = acc[q][s];
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <typename T>
int bsk_triple_plain_launch(const T* a, const T* b, const T* c, long long sa, long long sb,
                            long long sc, long long ncell, int ntri, const int* tri,
                            double* partial, int nchunk, hipStream_t s) {
  if (ncell % Vec<T>::n) {
    bsk_set_error("plain triple: cell count %lld not a multiple of %d", ncell, Vec<T>::n);
    return BSK_EINVAL;
  }
  if (ntri <= 0) return BSK_OK;
  if (ntri > 65535) {
    bsk_set_error("plain triple: at most 65535 triangles per call (got %d)", ntri);
    return BSK_EINVAL;
  }
  hipLaunchKernelGGL(k_triple_plain<T>, dim3(nchunk, ntri), dim3(kW), 0, s, a, b, c, sa, sb, sc,
                     ncell, tri, partial, nchunk);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
template int bsk_triple_plain_launch<float>(const float*, const float*, const float*, long long,
                                            long long, long long, long long, int, const int*,
                                            double*, int, hipStream_t);
template int bsk_triple_plain_launch<double>(const double*, const double*, const double*,
                                             long long, long long, long long, long long, int,
                                             const int*, double*, int, hipStream_t);

int bsk_reduce_plain_launch(const double* partial, int ntri, int nchunk, double scale, double* out,
                            hipStream_t s) {
  if (ntri <= 0) return BSK_OK;
  hipLaunchKernelGGL(k_reduce_rows, dim3(ntri), dim3(64), 0, s, partial, nchunk, scale, out);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <typename T>
int bsk_triple_tiled_launch(const T* f, long long stride, int nfield, long long ncell, int ntri,
                            const int* tri, int nsum, double* partial, int nchunk, hipStream_t s) {
  constexpr int CC = 32, TPT = 2;
  if (ncell % CC) {
    bsk_set_error("tiled triple: cell count %lld not a multiple of %d", ncell, CC);
    return BSK_EINVAL;
  }
  if (ntri <= 0) return BSK_OK;
  const size_t lds = (size_t)nfield * (CC + 1) * sizeof(T);
  if (lds > 160 * 1024) {
    bsk_set_error("tiled triple: %d fields exceed the LDS budget", nfield);
    return BSK_EUNSUPPORTED;
  }
  const dim3 grid(nchunk, (ntri + kW * TPT - 1) / (kW * TPT));
  if (nsum == 1) {
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_triple_tiled<T, 1, TPT, CC>, hipFuncAttributeMaxDynamicSharedMemorySize,
        160 * 1024);
    BSK_HIP_CHECK(attr);
    hipLaunchKernelGGL((k_triple_tiled<T, 1, TPT, CC>), grid, dim3(kW), lds, s, f, stride, nfield,
                       ncell, ntri, tri, partial, nchunk);
  } else if (nsum == 4) {
    static const hipError_t attr = hipFuncSetAttribute(
        (const void*)k_triple_tiled<T, 4, TPT, CC>, hipFuncAttributeMaxDynamicSharedMemorySize,
        160 * 1024);
    BSK_HIP_CHECK(attr);
    hipLaunchKernelGGL((k_triple_tiled<T, 4, TPT, CC>), grid, dim3(kW), lds, s, f, stride, nfield,
                       ncell, ntri, tri, partial, nchunk);
  } else {
    bsk_set_error("tiled triple: nsum must be 1 or 4");
    return BSK_EINVAL;
  }
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
template int bsk_triple_tiled_launch<float>(const float*, long long, int, long long, int,
                                            const int*, int, double*, int, hipStream_t);
template int bsk_triple_tiled_launch<double>(const double*, long long, int, long long, int,
                                             const int*, int, double*, int, hipStream_t);

int bsk_reduce_tiled_launch(const double* partial, int ntri, int nsum, int nchunk, double scale,
                            double* out, hipStream_t s) {
  if (ntri <= 0) return BSK_OK;
  hipLaunchKernelGGL(k_reduce_rows, dim3(ntri * nsum), dim3(64), 0, s, partial, nchunk, scale,
                     out);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

