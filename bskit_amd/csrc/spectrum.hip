// Binned power spectrum sums on the compact delta(k) cube.
//
// For bin b (modes with n2lo[b] <= |n|^2 <= n2hi[b]) the three sums
//   S_P = sum w |delta_k|^2,  S_N = sum w,  S_K = sum w k_f |n|
// run over the retained half spectrum [ry][rx][kz] with weight w = 1 on the
// kz = 0 plane (and the kz = N/2 plane) and 2 elsewhere, i.e. over every mode
// of the full spectrum.  By Parseval these equal the cell sums of squared
// shell fields used by a real-space estimator: sum_x d_b^2 = N^3 S_P,
// sum_x n_b^2 = N^3 S_N, sum_x (c2r of |k|^(1/2) mask)^2 = N^3 S_K.
// Grid (nchunk, nbin): each block sums its chunk of the cube for one bin
// with a fixed-order tree; a second kernel adds the chunks in order (FP64).
#include "launch.h"

namespace {

constexpr int kT = 256;

template <typename T>
__global__ __launch_bounds__(kT) void k_spectrum(const cplx<T>* __restrict__ dk, Geom g, int N,
                                                 const int* __restrict__ n2lo,
                                                 const int* __restrict__ n2hi, double kf,
                                                 int nchunk, double* __restrict__ partial) {
  __shared__ double red[3][kT];
  const int b = blockIdx.y, c = blockIdx.x;
  const long long total = (long long)g.MX * g.MY * g.MZ;
  const long long e0 = total * c / nchunk, e1 = total * (c + 1) / nchunk;
  const int lo = n2lo[b], hi = n2hi[b];
  double sp = 0.0, sn = 0.0, sk = 0.0;
  for (long long e = e0 + threadIdx.x; e < e1; e += kT) {
    const int kz = (int)(e % g.MZ);
    const long long r = e / g.MZ;
    const int rx = (int)(r % g.MX), ry = (int)(r / g.MX);
    const int nx = g.nlo + rx, ny = g.nlo + ry;
    const int n2 = nx * nx + ny * ny + kz * kz;
    if (n2 < lo || n2 > hi) continue;
    const double w = (kz == 0 || 2 * kz == N) ? 1.0 : 2.0;
    const cplx<T> v = dk[e];
    sp += w * ((double)v.x * v.x + (double)v.y * v.y);
    sn += w;
    sk += w * kf * sqrt((double)n2);
  }
  red[0][threadIdx.x] = sp;
  red[1][threadIdx.x] = sn;
  red[2][threadIdx.x] = sk;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[((long long)b * nchunk + c) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_spectrum_reduce(const double* __restrict__ partial,
                                                        int nchunk, double* __restrict__ out) {
  const int b = blockIdx.x;
  if (threadIdx.x >= 3) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += partial[((long long)b * nchunk + c) * 3 + threadIdx.x];
  out[b * 3 + threadIdx.x] = s;
}

}  // namespace

int bsk_spectrum_launch(int dtype, const void* dk, Geom g, int N, const int* n2lo,
                        const int* n2hi, int nbin, double kf, int nchunk, double* partial,
                        double* out, hipStream_t s) {
  if (nbin <= 0) return BSK_OK;
  if (nbin > 65535 || nchunk <= 0) {
    bsk_set_error("power spectrum: 1..65535 bins and a positive chunk count");
    return BSK_EINVAL;
  }
  const dim3 grid(nchunk, nbin);
  if (dtype == BSK_DT_F32)
    hipLaunchKernelGGL(k_spectrum<float>, grid, dim3(kT), 0, s, (const cplx<float>*)dk, g, N,
                       n2lo, n2hi, kf, nchunk, partial);
  else
    hipLaunchKernelGGL(k_spectrum<double>, grid, dim3(kT), 0, s, (const cplx<double>*)dk, g, N,
                       n2lo, n2hi, kf, nchunk, partial);
  BSK_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_spectrum_reduce, dim3(nbin), dim3(64), 0, s, partial, nchunk, out);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
