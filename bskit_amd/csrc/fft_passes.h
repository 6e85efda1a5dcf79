// The six FFT passes of the bispectrum path, as kernel templates.
//
// HBM layouts (complex elements are interleaved re,im; all offsets 64-bit):
//   real field   R [x][y][z]          N^3 reals
//   E1           [x][y][kz]           after the z r2c, kz < MZ kept
//   E2           [x][ry][kz]          after the y FFT, ry < MY kept
//   D0 (delta_k) [ry][rx][kz]         after the x FFT, compact cube, scaled 1/N^3
//   D1           [b][ry][x][kz]       after the masked inverse x FFT
//   D2           [b][xl][y][kz]       after the inverse y FFT (x slab local)
//   shells       [b][x][y][z]         after the z c2r, one N^3 array per bin
// Retained frequencies: along x and y the signed frequencies nlo..nlo+MX-1
// (nlo = -m, MX = 2m+1, or the whole axis when 2m+1 >= N); along z 0..MZ-1.
// Every pass keeps the kz index innermost, so the strided x and y passes read
// and write P consecutive kz values (P*8 or P*16 bytes) per grid point.
#pragma once
#include "fft.h"

struct Geom {
  int N;    // points per axis
  int MX;   // retained x frequencies (== MY)
  int MY;   // retained y frequencies
  int MZ;   // retained z frequencies
  int nlo;  // first retained signed x/y frequency
};

BSK_DEV int signed_freq(int idx, int N) { return idx < N / 2 ? idx : idx - N; }

// ------------------------------------------------------------------ F1: z r2c
// One workgroup = P rows (x,y) of N reals, viewed as NH = N/2 complex values.
template <typename T, int NH, int V, int P, int W>
__global__ __launch_bounds__(W) void k_fwd_z(const T* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg,
                                             long long nrows, int MZ) {
  constexpr int N = 2 * NH;
  constexpr int LD = pencil_ld<NH>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  const long long row0 = (long long)blockIdx.x * P;
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in);
  for (int e = tid; e < P * NH; e += W) {
    const int p = e / NH, n = e - p * NH;
    const long long row = row0 + p;
    cplx<T> v = {(T)0, (T)0};
    if (row < nrows) v = src[row * NH + n];
    buf[p * LD + n] = v;
  }
  __syncthreads();
  fft_lds<T, NH, false, P, W, LD, N>(buf, tw, typename Plan<NH, V>::seq{});
  for (int e = tid; e < P * MZ; e += W) {
    const int p = e / MZ, k = e - p * MZ;
    const long long row = row0 + p;
    if (row >= nrows) continue;
    cplx<T> zk = buf[p * LD + (k % NH)];
    cplx<T> zc = cconj(buf[p * LD + ((NH - k) % NH)]);
    cplx<T> xe = cscale(cadd(zk, zc), (T)0.5);
    cplx<T> d = csub(zk, zc);
    cplx<T> xo = {(T)0.5 * d.y, (T)-0.5 * d.x};  // (zk - zc) / (2i)
    out[row * MZ + k] = cadd(xe, cmul(tw[k], xo));
  }
}

// ------------------------------------------------------------------ F2: y fwd
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_fwd_y(const cplx<T>* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg, Geom g) {
  constexpr int LD = pencil_ld<N>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  const long long x = blockIdx.x;
  const int kz0 = blockIdx.y * P;
  for (int e = tid; e < P * N; e += W) {
    const int y = e / P, p = e - y * P, kz = kz0 + p;
    cplx<T> v = {(T)0, (T)0};
    if (kz < g.MZ) v = in[(x * N + y) * g.MZ + kz];
    buf[p * LD + y] = v;
  }
  __syncthreads();
  fft_lds<T, N, false, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  for (int e = tid; e < P * g.MY; e += W) {
    const int r = e / P, p = e - r * P, kz = kz0 + p;
    if (kz >= g.MZ) continue;
    const int idx = (g.nlo + r + N) % N;
    out[(x * g.MY + r) * g.MZ + kz] = buf[p * LD + idx];
  }
}

// ------------------------------------------------------------------ F3: x fwd
// Epilogue: scale by 1/N^3 and the optional separable window compensation
// comp[n + N/2] (n = signed frequency), applied along all three axes.
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_fwd_x(const cplx<T>* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg, Geom g,
                                             T scale, const T* __restrict__ comp) {
  constexpr int LD = pencil_ld<N>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  const long long ry = blockIdx.x;
  const int kz0 = blockIdx.y * P;
  for (int e = tid; e < P * N; e += W) {
    const int x = e / P, p = e - x * P, kz = kz0 + p;
    cplx<T> v = {(T)0, (T)0};
    if (kz < g.MZ) v = in[((long long)x * g.MY + ry) * g.MZ + kz];
    buf[p * LD + x] = v;
  }
  __syncthreads();
  fft_lds<T, N, false, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  const int ny = g.nlo + (int)ry;
  for (int e = tid; e < P * g.MX; e += W) {
    const int r = e / P, p = e - r * P, kz = kz0 + p;
    if (kz >= g.MZ) continue;
    const int nx = g.nlo + r;
    T s = scale;
    if (comp) s *= comp[nx + N / 2] * comp[ny + N / 2] * comp[kz + N / 2];
    out[(ry * g.MX + r) * g.MZ + kz] = cscale(buf[p * LD + (nx + N) % N], s);
  }
}

// ------------------------------------------------- S1: masked inverse x FFT
// grid (MY, ceil(MZ/P), nshell).  kind: delta / unit / |k|.
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_x_mask(const cplx<T>* __restrict__ dk,
                                                  cplx<T>* __restrict__ out,
                                                  const cplx<T>* __restrict__ twg, Geom g,
                                                  const int* __restrict__ n2lo,
                                                  const int* __restrict__ n2hi, int shell0,
                                                  int kind, T kf) {
  constexpr int LD = pencil_ld<N>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  const long long ry = blockIdx.x;
  const int kz0 = blockIdx.y * P;
  const int bl = blockIdx.z;
  const int lo = n2lo[shell0 + bl], hi = n2hi[shell0 + bl];
  const int ny = g.nlo + (int)ry;
  for (int e = tid; e < P * N; e += W) {
    const int xi = e / P, p = e - xi * P, kz = kz0 + p;
    const int nx = signed_freq(xi, N);
    const int rx = nx - g.nlo;
    cplx<T> v = {(T)0, (T)0};
    if (kz < g.MZ && rx >= 0 && rx < g.MX) {
      const int n2 = nx * nx + ny * ny + kz * kz;
      if (n2 >= lo && n2 <= hi) {
        if (kind == BSK_KIND_DELTA)
          v = dk[(ry * g.MX + rx) * g.MZ + kz];
        else if (kind == BSK_KIND_UNIT)
          v.x = (T)1;
        else
          v.x = kf * sqrt((T)n2);
      }
    }
    buf[p * LD + xi] = v;
  }
  __syncthreads();
  fft_lds<T, N, true, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  const long long obase = ((long long)bl * g.MY + ry) * N;
  for (int e = tid; e < P * N; e += W) {
    const int x = e / P, p = e - x * P, kz = kz0 + p;
    if (kz >= g.MZ) continue;
    out[(obase + x) * g.MZ + kz] = buf[p * LD + x];
  }
}

// -------------------------------------------------------- S2: inverse y FFT
// grid (xs, ceil(MZ/P), nshell); x = x0 + blockIdx.x.
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_y(const cplx<T>* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg, Geom g, int x0,
                                             int xs) {
  constexpr int LD = pencil_ld<N>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  const int xl = blockIdx.x;
  const long long x = x0 + xl;
  const int kz0 = blockIdx.y * P;
  const long long bl = blockIdx.z;
  for (int e = tid; e < P * N; e += W) {
    const int yi = e / P, p = e - yi * P, kz = kz0 + p;
    const int ry = signed_freq(yi, N) - g.nlo;
    cplx<T> v = {(T)0, (T)0};
    if (kz < g.MZ && ry >= 0 && ry < g.MY)
      v = in[((bl * g.MY + ry) * N + x) * g.MZ + kz];
    buf[p * LD + yi] = v;
  }
  __syncthreads();
  fft_lds<T, N, true, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  const long long obase = (bl * xs + xl) * N;
  for (int e = tid; e < P * N; e += W) {
    const int y = e / P, p = e - y * P, kz = kz0 + p;
    if (kz >= g.MZ) continue;
    out[(obase + y) * g.MZ + kz] = buf[p * LD + y];
  }
}

// --------------------------------------------------------- S3: z c2r + store
// grid (ceil(xs*N/P), nshell).  Rows are (xl, y) of the slab.
BSK_DEV uint32_t bf16_bits_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

template <typename T>
BSK_DEV void store_pair(int fmt, void* out, void* out_lo, long long off, T a, T b);

template <>
BSK_DEV void store_pair<float>(int fmt, void* out, void* out_lo, long long off, float a, float b) {
  if (fmt == BSK_FMT_BF16X2) {
    // hi = rne_bf16(v), lo = rne_bf16(v - hi) with v_cvt_pk_bf16_f32
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const f32x2 v = {a, b};
    const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    const f32x2 r = {a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u)};
    const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
    reinterpret_cast<uint32_t*>(out)[off >> 1] = hi;
    reinterpret_cast<uint32_t*>(out_lo)[off >> 1] = lo;
  } else {
    reinterpret_cast<float2*>(out)[off >> 1] = make_float2(a, b);
  }
}

template <>
BSK_DEV void store_pair<double>(int fmt, void* out, void*, long long off, double a, double b) {
  reinterpret_cast<double2*>(out)[off >> 1] = make_double2(a, b);
}

template <typename T, int NH, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_z(const cplx<T>* __restrict__ in, void* out,
                                             void* out_lo, const cplx<T>* __restrict__ twg,
                                             int MZ, long long nrows, long long cell0,
                                             long long shell_stride, int fmt) {
  constexpr int N = 2 * NH;
  constexpr int LD = pencil_ld<NH>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += W) tw[i] = twg[i];
  __syncthreads();
  const long long row0 = (long long)blockIdx.x * P;
  const long long bl = blockIdx.y;
  const cplx<T>* src = in + bl * nrows * MZ;
  for (int e = tid; e < P * NH; e += W) {
    const int p = e / NH, k = e - p * NH;
    const long long row = row0 + p;
    cplx<T> z = {(T)0, (T)0};
    if (row < nrows) {
      const int j = NH - k;  // 1..NH
      cplx<T> xk = {(T)0, (T)0}, xj = {(T)0, (T)0};
      if (k < MZ) xk = src[row * MZ + k];
      if (j < MZ) xj = src[row * MZ + j];
      cplx<T> xjc = cconj(xj);
      cplx<T> s = cadd(xk, xjc);
      cplx<T> d = csub(xk, xjc);
      cplx<T> w = cconj(tw[k]);  // exp(+2 pi i k / N)
      cplx<T> wd = cmul(w, d);
      z = {s.x - wd.y, s.y + wd.x};  // s + i * w * d
    }
    buf[p * LD + k] = z;
  }
  __syncthreads();
  fft_lds<T, NH, true, P, W, LD, N>(buf, tw, typename Plan<NH, V>::seq{});
  for (int e = tid; e < P * NH; e += W) {
    const int p = e / NH, n = e - p * NH;
    const long long row = row0 + p;
    if (row >= nrows) continue;
    const cplx<T> z = buf[p * LD + n];
    const long long off = bl * shell_stride + cell0 + row * N + 2 * n;
    store_pair<T>(fmt, out, out_lo, off, z.x, z.y);
  }
}
