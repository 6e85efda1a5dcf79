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
//  * mfma  : split-BF16 GEMM per i: sum_x (d_i d_j)(x) d_k(x) with the pair
//            products split to hi+lo bf16 and multiplied on
//            v_mfma_f32_16x16x32_bf16 (hi*hi + hi*lo + lo*hi, FP32 accumulate).
// All three write per-workgroup partials that a second kernel reduces in a
// fixed order in FP64, so a result does not depend on what else runs.
#include <type_traits>

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

// ------------------------------------------------------------------ mfma
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct U4 {
  uint32_t w[4];
};

BSK_DEV U4 ld16(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return U4{{v.x, v.y, v.z, v.w}};
}

BSK_DEV bf16x8 as_bf16x8(const U4& u) { return __builtin_bit_cast(bf16x8, u); }

// FP32 value hi+lo of packed element e (0..7).
BSK_DEV float unpack(const U4& h, const U4& l, int e) {
  const uint32_t wh = h.w[e >> 1], wl = l.w[e >> 1];
  const uint32_t bh = (e & 1) ? (wh & 0xffff0000u) : (wh << 16);
  const uint32_t bl = (e & 1) ? (wl & 0xffff0000u) : (wl << 16);
  return __uint_as_float(bh) + __uint_as_float(bl);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Split two FP32 values into packed bf16 hi (round to nearest even) and
// packed bf16 lo = rne(v - hi): v_cvt_pk_bf16_f32, shift/and, sub, cvt_pk.
BSK_DEV void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
  const f32x2 v = {a, b};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
  const f32x2 r = {a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

template <int II>
BSK_DEV float row_bcast(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + II, 0xf, 0xf, false));
}

template <int B, int E, typename F>
BSK_DEV void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <int MODE, int NB>
constexpr int ntiles_i(int ib) {
  return MODE == 0 ? (ib + 1) * (ib + 2) / 2 : (MODE == 1 ? (ib + 1) * NB : NB * NB);
}
template <int MODE, int NB>
constexpr int tile_in_i(int jb, int kb) {
  return MODE == 0 ? jb * (jb + 1) / 2 + kb : jb * NB + kb;
}
template <int MODE, int NB>
constexpr int tiles_before(int ib) {
  int s = 0;
  for (int b = 0; b < ib; ++b) s += ntiles_i<MODE, NB>(b);
  return s;
}
template <int MODE, int NB>
constexpr int jb_end(int ib) {
  return MODE == 2 ? NB : ib + 1;
}
template <int MODE, int NB>
constexpr int kb_end(int jb) {
  return MODE == 0 ? jb + 1 : NB;
}

template <int NB, int MODE, int WAVES, int U>
struct MfmaKernel {
  static constexpr int IPW = 16 / WAVES;
  static constexpr int NSLOT = IPW * tiles_before<MODE, NB>(NB);
  static constexpr int TILES = 16 * tiles_before<MODE, NB>(NB);
  static constexpr bool SEP_I = (MODE == 2);  // i-side set differs from j-side set

  // The work of one K step (32 cells) for the i-values of wave WV.  All
  // block/slot indices are compile-time so the accumulators stay in
  // registers and the DPP broadcast lane is an immediate.
  template <int WV>
  static BSK_DEV void step(f32x4* acc, const float (*dI)[8], const float (*dJ)[8],
                           const U4 (*kh)[1], const U4 (*kl)[1], int nI) {
    static_for<0, NB>([&](auto ibc) {
      constexpr int ib = decltype(ibc)::value;
      static_for<0, IPW>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int ii = WV * IPW + q;
        if (16 * ib + ii >= nI) return;
        static_for<0, jb_end<MODE, NB>(ib)>([&](auto jbc) {
          constexpr int jb = decltype(jbc)::value;
          // pair products d_i(x) * d_j(x) for this lane's j and 8 cells
          float p[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) p[e] = dJ[jb][e] * row_bcast<ii>(dI[ib][e]);
          U4 ah, al;
#pragma unroll
          for (int e = 0; e < 4; ++e) split2(p[2 * e], p[2 * e + 1], ah.w[e], al.w[e]);
          const bf16x8 Ah = as_bf16x8(ah), Al = as_bf16x8(al);
          static_for<0, kb_end<MODE, NB>(jb)>([&](auto kbc) {
            constexpr int kb = decltype(kbc)::value;
            constexpr int s = IPW * tiles_before<MODE, NB>(ib) + q * ntiles_i<MODE, NB>(ib) +
                              tile_in_i<MODE, NB>(jb, kb);
            const bf16x8 Bh = as_bf16x8(kh[kb][0]), Bl = as_bf16x8(kl[kb][0]);
            f32x4 c = acc[s];
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, c, 0, 0, 0);
            acc[s] = c;
          });
        });
      });
    });
  }

  template <int WV>
  static BSK_DEV void run(const MfmaArgs& a) {
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    f32x4 acc[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};

    const long long niter = a.ncell / (32 * U);
    const long long it0 = niter * blockIdx.x / a.nwg;
    const long long it1 = niter * (blockIdx.x + 1) / a.nwg;
    // per-set, per-block row pointers (this lane's shell), or null for padding
    const uint16_t* ph[3][NB];
    const uint16_t* pl[3][NB];
#pragma unroll
    for (int st = 0; st < 3; ++st)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int sh = 16 * b + col;
        const bool ok = sh < a.nsh[st];
        ph[st][b] = ok ? a.hi[st] + (long long)sh * a.stride[st] : nullptr;
        pl[st][b] = ok ? a.lo[st] + (long long)sh * a.stride[st] : nullptr;
      }
    const U4 zero = {{0u, 0u, 0u, 0u}};
    for (long long it = it0; it < it1; ++it) {
      const long long cbase = it * (32 * U) + grp * (8 * U);
      U4 rh[3][NB][U], rl[3][NB][U];
#pragma unroll
      for (int st = 0; st < 3; ++st) {
        if ((st == 0 && !SEP_I) || (st == 2 && MODE == 0)) continue;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int u = 0; u < U; ++u) {
            rh[st][b][u] = ph[st][b] ? ld16(ph[st][b] + cbase + 8 * u) : zero;
            rl[st][b][u] = pl[st][b] ? ld16(pl[st][b] + cbase + 8 * u) : zero;
          }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        float dJ[NB][8], dI[NB][8];
        U4 kh[NB][1], kl[NB][1];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
          for (int e = 0; e < 8; ++e) dJ[b][e] = unpack(rh[1][b][u], rl[1][b][u], e);
          if (SEP_I) {
#pragma unroll
            for (int e = 0; e < 8; ++e) dI[b][e] = unpack(rh[0][b][u], rl[0][b][u], e);
          }
          const int ks = (MODE == 0) ? 1 : 2;
          kh[b][0] = rh[ks][b][u];
          kl[b][0] = rl[ks][b][u];
        }
        step<WV>(acc, SEP_I ? dI : dJ, dJ, kh, kl, a.nsh[SEP_I ? 0 : 1]);
      }
    }
    // write the tiles of this wave: C[row][col], row = 4*grp + r (j), col = lane%16 (k)
    float* out = a.partial + (long long)blockIdx.x * TILES * 256;
#pragma unroll
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
      for (int q = 0; q < IPW; ++q)
#pragma unroll
        for (int jb = 0; jb < jb_end<MODE, NB>(ib); ++jb)
#pragma unroll
          for (int kb = 0; kb < kb_end<MODE, NB>(jb); ++kb) {
            const int s = IPW * tiles_before<MODE, NB>(ib) + q * ntiles_i<MODE, NB>(ib) +
                          tile_in_i<MODE, NB>(jb, kb);
            const int ii = WV * IPW + q;
            const int g = 16 * tiles_before<MODE, NB>(ib) + ii * ntiles_i<MODE, NB>(ib) +
                          tile_in_i<MODE, NB>(jb, kb);
#pragma unroll
            for (int r = 0; r < 4; ++r) out[g * 256 + (4 * grp + r) * 16 + col] = acc[s][r];
          }
  }
};

template <int NB, int MODE, int WAVES, int U>
__global__ __launch_bounds__(64 * WAVES) void k_triple_mfma(MfmaArgs a) {
  using K = MfmaKernel<NB, MODE, WAVES, U>;
  const int wv = threadIdx.x >> 6;
  if constexpr (WAVES == 4) {
    switch (wv) {
      case 0: K::template run<0>(a); break;
      case 1: K::template run<1>(a); break;
      case 2: K::template run<2>(a); break;
      default: K::template run<3>(a); break;
    }
  } else {
    switch (wv) {
      case 0: K::template run<0>(a); break;
      case 1: K::template run<1>(a); break;
      case 2: K::template run<2>(a); break;
      case 3: K::template run<3>(a); break;
      case 4: K::template run<4>(a); break;
      case 5: K::template run<5>(a); break;
      case 6: K::template run<6>(a); break;
      default: K::template run<7>(a); break;
    }
  }
}

__global__ __launch_bounds__(64) void k_reduce_mfma(const float* __restrict__ partial, int nwg,
                                                    long long wg_stride, const int* __restrict__ slot,
                                                    double scale, double* __restrict__ out) {
  const int t = blockIdx.x;
  const long long e = slot[t];
  double s = 0.0;
  for (int w = threadIdx.x; w < nwg; w += 64) s += (double)partial[w * wg_stride + e];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[t] = s * scale;
}

template <int NB, int MODE>
constexpr int waves_for() {
  return (NB >= 3 || (NB == 2 && MODE != 0)) ? 8 : 4;
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
    BSK_HIP_CHECK(hipFuncSetAttribute((const void*)k_triple_tiled<T, 1, TPT, CC>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_triple_tiled<T, 1, TPT, CC>), grid, dim3(kW), lds, s, f, stride, nfield,
                       ncell, ntri, tri, partial, nchunk);
  } else if (nsum == 4) {
    BSK_HIP_CHECK(hipFuncSetAttribute((const void*)k_triple_tiled<T, 4, TPT, CC>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
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

int bsk_mfma_tiles(int nblk, int mode) {
  int s = 0;
  for (int ib = 0; ib < nblk; ++ib)
    s += mode == 0 ? (ib + 1) * (ib + 2) / 2 : (mode == 1 ? (ib + 1) * nblk : nblk * nblk);
  return 16 * s;
}

constexpr int kMfmaU = 2;

int bsk_triple_mfma_launch(const MfmaArgs& a, hipStream_t s) {
  if (a.ncell % (32 * kMfmaU)) {
    bsk_set_error("mfma triple: cell count %lld not a multiple of %d", a.ncell, 32 * kMfmaU);
    return BSK_EINVAL;
  }
  if (a.nwg <= 0) {
    bsk_set_error("mfma triple: nwg must be positive");
    return BSK_EINVAL;
  }
#define BSK_MF(NB, MODE)                                                                     \
  if (a.nblk == NB && a.mode == MODE) {                                                      \
    constexpr int W = waves_for<NB, MODE>();                                                 \
    hipLaunchKernelGGL((k_triple_mfma<NB, MODE, W, kMfmaU>), dim3(a.nwg), dim3(64 * W), 0, s, \
                       a);                                                                   \
    BSK_LAUNCH_CHECK();                                                                      \
    return BSK_OK;                                                                           \
  }
  BSK_MF(1, 0) BSK_MF(2, 0) BSK_MF(3, 0)
  BSK_MF(1, 1) BSK_MF(2, 1) BSK_MF(3, 1)
  BSK_MF(1, 2) BSK_MF(2, 2) BSK_MF(3, 2)
#undef BSK_MF
  bsk_set_error("mfma triple: unsupported block count %d / mode %d", a.nblk, a.mode);
  return BSK_EUNSUPPORTED;
}

int bsk_reduce_mfma_launch(const float* partial, int nwg, int tiles_per_wg, int ntri,
                           const int* slot, double scale, double* out, hipStream_t s) {
  if (ntri <= 0) return BSK_OK;
  hipLaunchKernelGGL(k_reduce_mfma, dim3(ntri), dim3(64), 0, s, partial, nwg,
                     (long long)tiles_per_wg * 256, slot, scale, out);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
