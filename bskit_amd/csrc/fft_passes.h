// The six FFT passes of the bispectrum path, as kernel templates.
//
// HBM layouts (complex elements are interleaved re,im; all offsets 64-bit):
//   real field   R [x][y][z]          N^3 reals
//   E1           [x][y][kz]           after the z r2c, kz < MZ kept (stride MZP)
//   E2           [x][ry][kz]          after the y FFT, ry < MY kept (stride MZP)
//   D0 (delta_k) [ry][rx][kz]         after the x FFT, compact cube, scaled 1/N^3
//   D1           [b][x][ry][kz]       after the masked inverse x FFT (stride MZP)
//   D2           [b][xl][y][kz]       after the inverse y FFT, x slab local (MZP)
//   shells       [b][x][y][z]         after the z c2r, one N^3 array per bin
// Retained frequencies: along x and y the signed frequencies nlo..nlo+MX-1
// (nlo = -m, MX = 2m+1, or the whole axis when 2m+1 >= N); along z 0..MZ-1.
// Every pass keeps the kz index innermost, so the strided x and y passes read
// and write P consecutive kz values (P*8 or P*16 bytes) per grid point; the
// kz block is the fastest logical grid dimension and the XCD remap below keeps
// the blocks of one kz row on one XCD, so they run together in one L2.
#pragma once
#include "fft.h"

struct Geom {
  int N;    // points per axis
  int MX;   // retained x frequencies (== MY)
  int MY;   // retained y frequencies
  int MZ;   // retained z frequencies
  int nlo;  // first retained signed x/y frequency
  int MZP;  // kz stride of the FFT intermediates E1, E2, D1, D2: MZ rounded
            // up to 16, so the P = 16 (FP32) kz blocks of the strided passes
            // are whole 128-B lines (the padding is never read or written)
};

BSK_DEV int signed_freq(int idx, int N) { return idx < N / 2 ? idx : idx - N; }

// Largest r with r*r <= v (-1 for v < 0).
BSK_DEV int isqrt_floor(int v) {
  if (v < 0) return -1;
  int r = (int)sqrtf((float)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

// Shell b holds only modes with |n|^2 <= n2hi[b], so after the inverse x
// transform the pencil (ny, kz) of D1 is zero unless ny^2 + kz^2 <= n2hi[b],
// and after the inverse y transform the kz values of D2 are zero from
// mz_b = min(MZ, floor(sqrt(n2hi[b])) + 1) on.  The passes write and read
// only those parts (readers take the rest as zero): the same bits, and for a
// set of shells of radii up to m about a quarter of the D1 and half of the D2
// traffic of the full rectangles.
BSK_DEV int shell_mz(int MZ, int hi) {
  const int r = isqrt_floor(hi) + 1;
  return r < MZ ? r : MZ;
}

// The dispatcher deals workgroups round-robin to the 8 XCDs, each with its own
// L2.  Neighbouring kz blocks share the cache lines at their edges (a kz row
// is MZ*8 bytes, rarely a multiple of 128), so the linear block id is
// remapped to give each XCD one contiguous range of logical blocks, walked in
// dispatch order: the blocks of one row then meet in one L2.  The map is a
// bijection of [0, total): XCDs below `tall` take `per` blocks, the rest one
// fewer.
BSK_DEV int xcd_linear(int pid, int total) {
  const int per = (total + 7) >> 3;
  const int tall = total - (per - 1) * 8;
  const int xcd = pid & 7, local = pid >> 3;
  return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

struct Blk {
  int x, y, z;
};

// Logical (x, y, z) block of this workgroup under the XCD remap.
BSK_DEV Blk xcd_block() {
  const int gx = gridDim.x, gy = gridDim.y;
  const int pid = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int l = xcd_linear(pid, gx * gy * gridDim.z);
  const int r = l / gx;
  return Blk{l - r * gx, r % gy, r / gy};
}

// Global <-> LDS staging with the accesses of a batch in flight together.
// Thread t handles the elements t, t + W, ...; a batch issues up to 16
// loads (or LDS reads) into registers before the first is consumed, so a
// workgroup waits for the memory latency once per batch, not per element.
template <typename T, int TOTAL, int W, typename LoadF, typename StoreF>
BSK_DEV void stage_in(LoadF&& load, StoreF&& store) {
  constexpr int CNT = (TOTAL + W - 1) / W;
  constexpr int B = CNT < 16 ? CNT : 16;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q0 = 0; q0 < CNT; q0 += B) {
    cplx<T> v[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int e = tid + (q0 + q) * W;
      v[q] = (q0 + q < CNT && (TOTAL % W == 0 || e < TOTAL)) ? load(e) : cplx<T>{(T)0, (T)0};
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int e = tid + (q0 + q) * W;
      if (q0 + q < CNT && (TOTAL % W == 0 || e < TOTAL)) store(e, v[q]);
    }
  }
}

// Element e of a workgroup-uniform base through a 32-bit byte offset (the
// SGPR base + VGPR offset address form, no 64-bit address arithmetic per
// lane); every offset below stays under 4 GB.
template <typename V>
BSK_DEV V get_u32(const V* base, unsigned e) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) +
                                     e * (unsigned)sizeof(V));
}

// The same load, non-temporal: for data read once (the input field), so
// it does not displace cache lines that other kernels in flight reuse.
template <typename T>
BSK_DEV cplx<T> get_nt(const cplx<T>* base, unsigned e) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(
      reinterpret_cast<const char*>(base) + e * (unsigned)sizeof(cplx<T>)));
  return {v.x, v.y};
}

template <typename V>
BSK_DEV void set_u32(V* base, unsigned e, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + e * (unsigned)sizeof(V)) = v;
}

// Twiddles in two steps: loads into registers (issued before a pass's data
// loads, so both latencies overlap) and the LDS stores after them.
template <typename T, int N, int W>
struct TwiddleRegs {
  cplx<T> v[(N + W - 1) / W];
};
template <typename T, int N, int W>
BSK_DEV void fetch_twiddles(TwiddleRegs<T, N, W>& r, const cplx<T>* __restrict__ twg) {
#pragma unroll
  for (int i = 0; i < (N + W - 1) / W; ++i) {
    const int t = (int)threadIdx.x + i * W;
    r.v[i] = (N % W == 0 || t < N) ? twg[t] : cplx<T>{(T)0, (T)0};
  }
}
template <typename T, int N, int W>
BSK_DEV void put_twiddles(const TwiddleRegs<T, N, W>& r, cplx<T>* tw) {
#pragma unroll
  for (int i = 0; i < (N + W - 1) / W; ++i) {
    const int t = (int)threadIdx.x + i * W;
    if (N % W == 0 || t < N) tw[t] = r.v[i];
  }
}

template <typename T, int N, int W>
BSK_DEV void load_twiddles(cplx<T>* tw, const cplx<T>* __restrict__ twg) {
  stage_in<T, N, W>([&](int i) { return twg[i]; }, [&](int i, cplx<T> v) { tw[i] = v; });
}

// ------------------------------------------------------------------ F1: z r2c
// One workgroup = P rows (x,y) of N reals, viewed as NH = N/2 complex values.
template <typename T, int NH, int V, int P, int W>
__global__ __launch_bounds__(W) void k_fwd_z(const T* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg,
                                             long long nrows, int MZ, int MZP) {
  constexpr int N = 2 * NH;
  constexpr int LD = pencil_ld<NH>();
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[N];
  const int tid = threadIdx.x;
  TwiddleRegs<T, N, W> twr;
  fetch_twiddles<T, N, W>(twr, twg);
  const long long row0 = (long long)blockIdx.x * P;
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in) + row0 * NH;
  const int nr = (int)(nrows - row0 < P ? nrows - row0 : P);
  stage_in<T, P * NH, W>(
      [&](int e) {
        return e < nr * NH ? get_nt(src, (unsigned)e) : cplx<T>{(T)0, (T)0};
      },
      [&](int e, cplx<T> v) {
        const int p = e / NH;
        buf[p * LD + (e - p * NH)] = v;
      });
  put_twiddles<T, N, W>(twr, tw);
  __syncthreads();
  fft_lds<T, NH, false, P, W, LD, N>(buf, tw, typename Plan<NH, V>::seq{});
  cplx<T>* const dst = out + row0 * MZP;
  for (int e = tid; e < P * MZ; e += W) {
    const int p = e / MZ, k = e - p * MZ;
    if (p >= nr) continue;
    cplx<T> zk = buf[p * LD + (k % NH)];
    cplx<T> zc = cconj(buf[p * LD + ((NH - k) % NH)]);
    cplx<T> xe = cscale(cadd(zk, zc), (T)0.5);
    cplx<T> d = csub(zk, zc);
    cplx<T> xo = {(T)0.5 * d.y, (T)-0.5 * d.x};  // (zk - zc) / (2i)
    set_u32(dst, (unsigned)(p * MZP + k), cadd(xe, cmul(tw[k], xo)));
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
  TwiddleRegs<T, N, W> twr;
  fetch_twiddles<T, N, W>(twr, twg);
  const Blk bk = xcd_block();
  const long long x = bk.y;
  const int kz0 = bk.x * P;
  const int np = g.MZ - kz0 < P ? g.MZ - kz0 : P;  // valid pencils
  const cplx<T>* src = in + x * N * g.MZP + kz0;
  stage_in<T, P * N, W>(
      [&](int e) {
        const int y = e / P, p = e - y * P;
        return p < np ? get_u32(src, (unsigned)(y * g.MZP + p)) : cplx<T>{(T)0, (T)0};
      },
      [&](int e, cplx<T> v) {
        const int y = e / P, p = e - y * P;
        buf[p * LD + y] = v;
      });
  put_twiddles<T, N, W>(twr, tw);
  __syncthreads();
  fft_lds<T, N, false, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  cplx<T>* const dst = out + x * g.MY * g.MZP + kz0;
  for (int e = threadIdx.x; e < P * g.MY; e += W) {
    const int r = e / P, p = e - r * P;
    if (p >= np) continue;
    const int idx = (g.nlo + r + N) % N;
    set_u32(dst, (unsigned)(r * g.MZP + p), buf[p * LD + idx]);
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
  TwiddleRegs<T, N, W> twr;
  fetch_twiddles<T, N, W>(twr, twg);
  const Blk bk = xcd_block();
  const long long ry = bk.y;
  const int kz0 = bk.x * P;
  const int np = g.MZ - kz0 < P ? g.MZ - kz0 : P;
  const cplx<T>* src = in + ry * g.MZP + kz0;
  const int xstride = g.MY * g.MZP;
  stage_in<T, P * N, W>(
      [&](int e) {
        const int x = e / P, p = e - x * P;
        return p < np ? get_u32(src, (unsigned)(x * xstride + p)) : cplx<T>{(T)0, (T)0};
      },
      [&](int e, cplx<T> v) {
        const int x = e / P, p = e - x * P;
        buf[p * LD + x] = v;
      });
  put_twiddles<T, N, W>(twr, tw);
  __syncthreads();
  fft_lds<T, N, false, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  const int ny = g.nlo + (int)ry;
  cplx<T>* const dst = out + ry * g.MX * g.MZ + kz0;
  for (int e = threadIdx.x; e < P * g.MX; e += W) {
    const int r = e / P, p = e - r * P, kz = kz0 + p;
    if (p >= np) continue;
    const int nx = g.nlo + r;
    T s = scale;
    if (comp) s *= comp[nx + N / 2] * comp[ny + N / 2] * comp[kz + N / 2];
    set_u32(dst, (unsigned)(r * g.MZ + p), cscale(buf[p * LD + (nx + N) % N], s));
  }
}

// ------------------------------------------------- S1: masked inverse x FFT
// grid (nblocks): block entries (shell, ry << 16 | kz0) of the launch's
// shells, only the P-pencil blocks that meet the shell's disk ny^2 + kz^2
// <= n2hi (the rest of D1 is never read).  kind: delta / unit / |k| / |k|^kp.
// The XCD remap gives each XCD a contiguous range of entries (neighbouring
// kz blocks of one row).
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_x_mask(const cplx<T>* __restrict__ dk,
                                                  cplx<T>* __restrict__ out,
                                                  const cplx<T>* __restrict__ twg, Geom g,
                                                  int shell0, const int4* __restrict__ blocks,
                                                  int kind, T kf, T kp) {
  constexpr int LD = pencil_ld<N>();
  // the entries of the length-N table the plan reads (157 of 200 at N = 200,
  // which keeps the FP32 workgroup under 1/6 of the CU's LDS)
  constexpr int TWS = tw_span<N>(typename Plan<N, V>::seq{});
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[TWS];
  TwiddleRegs<T, TWS, W> twr;
  fetch_twiddles<T, TWS, W>(twr, twg);
  const int4 ent = blocks[xcd_linear(blockIdx.x, gridDim.x)];
  const long long ry = ent.y >> 16;
  const int kz0 = ent.y & 0xffff;
  const int bl = ent.x - shell0;
  const int lo = ent.z, hi = ent.w;  // the shell's n2lo, n2hi
  const int ny = g.nlo + (int)ry;
  const cplx<T>* src = dk + ry * g.MX * g.MZ;
  stage_in<T, P * N, W>(
      [&](int e) {
        const int xi = e / P, p = e - xi * P, kz = kz0 + p;
        const int nx = signed_freq(xi, N);
        const int rx = nx - g.nlo;
        cplx<T> v = {(T)0, (T)0};
        if (kz < g.MZ && rx >= 0 && rx < g.MX) {
          const int n2 = nx * nx + ny * ny + kz * kz;
          if (n2 >= lo && n2 <= hi) {
            if (kind == BSK_KIND_DELTA)
              v = get_u32(src, (unsigned)(rx * g.MZ + kz));
            else if (kind == BSK_KIND_UNIT)
              v.x = (T)1;
            else if (kind == BSK_KIND_KMAG)
              v.x = kf * sqrt((T)n2);
            else
              v.x = pow(kf * sqrt((T)n2), kp);
          }
        }
        return v;
      },
      [&](int e, cplx<T> v) {
        const int xi = e / P, p = e - xi * P;
        buf[p * LD + xi] = v;
      });
  put_twiddles<T, TWS, W>(twr, tw);
  __syncthreads();
  fft_lds<T, N, true, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  cplx<T>* const dst = out + ((long long)bl * N * g.MY + ry) * g.MZP + kz0;
  const int xstride = g.MY * g.MZP;
  for (int e = threadIdx.x; e < P * N; e += W) {
    const int x = e / P, p = e - x * P, kz = kz0 + p;
    if (kz >= g.MZ) continue;
    set_u32(dst, (unsigned)(x * xstride + p), buf[p * LD + x]);
  }
}

// -------------------------------------------------------- S2: inverse y FFT
// grid (nblocks, xs): block entries (shell, kz0) of the launch's shells,
// only the kz blocks below the shell's mz_b (the rest of D2 is zero and
// never read); x = x0 + logical block y.  Reads only the shell's disk of D1.
template <typename T, int N, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_y(const cplx<T>* __restrict__ in,
                                             cplx<T>* __restrict__ out,
                                             const cplx<T>* __restrict__ twg, Geom g, int x0,
                                             int xs, const int4* __restrict__ blocks,
                                             int shell0) {
  constexpr int LD = pencil_ld<N>();
  // the entries of the length-N table the plan reads (157 of 200 at N = 200,
  // which keeps the FP32 workgroup under 1/6 of the CU's LDS)
  constexpr int TWS = tw_span<N>(typename Plan<N, V>::seq{});
  __shared__ cplx<T> buf[P * LD];
  __shared__ cplx<T> tw[TWS];
  TwiddleRegs<T, TWS, W> twr;
  fetch_twiddles<T, TWS, W>(twr, twg);
  const Blk bk = xcd_block();
  const int xl = bk.y;
  const long long x = x0 + xl;
  const int4 ent = blocks[bk.x];
  const int kz0 = ent.y;
  const long long bl = ent.x - shell0;
  const int hi = ent.z;  // the shell's n2hi
  const int np = g.MZ - kz0 < P ? g.MZ - kz0 : P;
  const cplx<T>* src = in + (bl * N + x) * g.MY * g.MZP + kz0;
  stage_in<T, P * N, W>(
      [&](int e) {
        const int yi = e / P, p = e - yi * P;
        const int ny = signed_freq(yi, N), ry = ny - g.nlo, kz = kz0 + p;
        return (p < np && ry >= 0 && ry < g.MY && ny * ny + kz * kz <= hi)
                   ? get_u32(src, (unsigned)(ry * g.MZP + p))
                   : cplx<T>{(T)0, (T)0};
      },
      [&](int e, cplx<T> v) {
        const int yi = e / P, p = e - yi * P;
        buf[p * LD + yi] = v;
      });
  put_twiddles<T, TWS, W>(twr, tw);
  __syncthreads();
  fft_lds<T, N, true, P, W, LD, N>(buf, tw, typename Plan<N, V>::seq{});
  cplx<T>* const dst = out + (bl * xs + xl) * N * g.MZP + kz0;
  for (int e = threadIdx.x; e < P * N; e += W) {
    const int y = e / P, p = e - y * P;
    if (p >= np) continue;
    set_u32(dst, (unsigned)(y * g.MZP + p), buf[p * LD + y]);
  }
}

// --------------------------------------------------------- S3: z c2r + store
// grid (ceil(xs*N/P), nshell).  Rows are (xl, y) of the slab.
//
// c2r of length N = 2 NH by one NH-point complex transform: with X the half
// spectrum (kz < MZ stored, the rest zero), s = X_k + conj(X_{NH-k}),
// d = X_k - conj(X_{NH-k}) and w = exp(+2 pi i k / N),
//   z_k = s + i w d,   z_{NH-k} = conj(s) + i conj(w d),
// so one thread builds both from the same two loads; the transform of z gives
// x_{2n} = Re z_n, x_{2n+1} = Im z_n.
typedef float bsk_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bsk_bf16x2 __attribute__((ext_vector_type(2)));

// hi = rne_bf16(v), lo = rne_bf16(v - hi) of two values, packed per word.
// (v - hi exactly by v_dot2_f32_bf16: hi.x * -1 + hi.y * 0 + v; the bf16
// pairs (-1, 0) and (0, -1) come in registers, made once per thread.)
struct SplitConsts {
  bsk_bf16x2 neg_lo, neg_hi;
};

BSK_DEV SplitConsts split_consts() {
  return SplitConsts{bf16_pair_bits(kBf16NegOneLo), bf16_pair_bits(kBf16NegOneHi)};
}

BSK_DEV void bf16_split2(float a, float b, uint32_t& hi, uint32_t& lo, const SplitConsts& c) {
  const bsk_f32x2 v = {a, b};
  const bsk_bf16x2 h = __builtin_convertvector(v, bsk_bf16x2);
  const bsk_f32x2 r = {__builtin_amdgcn_fdot2_f32_bf16(h, c.neg_lo, a, false),
                       __builtin_amdgcn_fdot2_f32_bf16(h, c.neg_hi, b, false)};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bsk_bf16x2));
}

// The two z values of the mirrored pair (k, NH - k) of one row from
// X_k = xk and X_{NH-k} = xj (zero where not retained); wk = exp(+2 pi i k / N).
template <typename T>
BSK_DEV void c2r_pair(cplx<T>* row, int k, int NH, cplx<T> xk, cplx<T> xj, cplx<T> wk) {
  const int j = NH - k;
  const cplx<T> xjc = cconj(xj);
  const cplx<T> s = cadd(xk, xjc);
  const cplx<T> wd = cmul(wk, csub(xk, xjc));
  if (k < NH) row[k] = {s.x - wd.y, s.y + wd.x};            // s + i w d
  if (j != k && j < NH) row[j] = {s.x + wd.y, wd.x - s.y};  // conj(s) + i conj(w d)
}

// Stores into output planes addressed from a workgroup-uniform base with a
// 32-bit byte offset, which the compiler turns into the SGPR base + VGPR
// offset address form (no 64-bit address arithmetic per lane).  e is an
// element offset from the base; hi holds the values (F32/F64) or the bf16 hi
// plane, lo the bf16 lo plane.
template <typename V>
BSK_DEV void put_u32(char* base, unsigned byte_off, V v) {
  *reinterpret_cast<V*>(base + byte_off) = v;
}

// Bytes per stored element.
template <typename T>
BSK_DEV int out_esize(int fmt) {
  return sizeof(T) == 8 ? 8 : (fmt == BSK_FMT_BF16X2 ? 2 : 4);
}

// Four consecutive reals at element offset e (a multiple of 4).
template <typename T>
BSK_DEV void store_quad(int fmt, char* hi, char* lo, unsigned e, cplx<T> u, cplx<T> v,
                        const SplitConsts& c) {
  if constexpr (sizeof(T) == 8) {
    put_u32(hi, e * 8u, make_double2(u.x, u.y));
    put_u32(hi, e * 8u + 16u, make_double2(v.x, v.y));
  } else if (fmt == BSK_FMT_BF16X2) {
    uint32_t h0, l0, h1, l1;
    bf16_split2(u.x, u.y, h0, l0, c);
    bf16_split2(v.x, v.y, h1, l1, c);
    put_u32(hi, e * 2u, make_uint2(h0, h1));
    put_u32(lo, e * 2u, make_uint2(l0, l1));
  } else {
    put_u32(hi, e * 4u, make_float4(u.x, u.y, v.x, v.y));
  }
}

// Two consecutive reals at element offset e (even), and one zero.
template <typename T>
BSK_DEV void store_pair(int fmt, char* hi, char* lo, unsigned e, T a, T b,
                        const SplitConsts& c) {
  if constexpr (sizeof(T) == 8) {
    put_u32(hi, e * 8u, make_double2(a, b));
  } else if (fmt == BSK_FMT_BF16X2) {
    uint32_t h, l;
    bf16_split2(a, b, h, l, c);
    put_u32(hi, e * 2u, h);
    put_u32(lo, e * 2u, l);
  } else {
    put_u32(hi, e * 4u, make_float2(a, b));
  }
}

// Four reals a, b, c, d at element offsets e .. e + 3 (e a multiple of 4).
template <typename T>
BSK_DEV void store_four(int fmt, char* hi, char* lo, unsigned e, T a, T b, T c, T d,
                        const SplitConsts& k) {
  if constexpr (sizeof(T) == 8) {
    put_u32(hi, e * 8u, make_double2(a, b));
    put_u32(hi, e * 8u + 16u, make_double2(c, d));
  } else if (fmt == BSK_FMT_BF16X2) {
    uint32_t h0, l0, h1, l1;
    bf16_split2(a, b, h0, l0, k);
    bf16_split2(c, d, h1, l1, k);
    put_u32(hi, e * 2u, make_uint2(h0, h1));
    put_u32(lo, e * 2u, make_uint2(l0, l1));
  } else {
    put_u32(hi, e * 4u, make_float4(a, b, c, d));
  }
}

template <typename T>
BSK_DEV void store_zero(int fmt, char* hi, char* lo, unsigned e) {
  if constexpr (sizeof(T) == 8) {
    put_u32(hi, e * 8u, 0.0);
  } else if (fmt == BSK_FMT_BF16X2) {
    put_u32(hi, e * 2u, (uint16_t)0);
    put_u32(lo, e * 2u, (uint16_t)0);
  } else {
    put_u32(hi, e * 4u, 0.f);
  }
}

constexpr int pow2_at_least(int v) {
  int q = 1;
  while (q < v) q *= 2;
  return q;
}

// Sublattice level layout of one shell of the S^3 mesh (S = 2 NH): the
// cells in four parts, each zero-padded to a multiple of 128 cells,
//   P0 = C0 minus E,  P1 = E,  P2 = O,  P3 = C1 minus O,
// with C0 / C1 the cells of even / odd x+y+z and E / O the cells with x, y,
// z all even / all odd (E lies in C0, O in C1).  A cell (x, y, z) of a part
// sits at ((c NH + x/2) NH + y/2) NH + z/2, c its row class: 0 in E and O,
// (x%2) 2 + y%2 - 1 in P0 (rows whose x, y are not both even), (x%2) 2 + y%2
// in P3 (rows whose x, y are not both odd).  The sum over a sublattice H of
// the mesh is |H|/S^3 of the full sum whenever no nonzero vector of its dual
// lattice is a frequency of the product, and every level is one contiguous
// range: the whole mesh P0..P3 (index 1), C0 = P0 P1 (index 2, dual vectors
// (S/2)(+-1,+-1,+-1) + S Z^3), E and O = P1 P2 (index 4, dual (S/2)(+-1,+-1,0)
// and permutations + S Z^3) and E = P1 (index 8, dual (S/2) Z^3).
template <int NH>
struct ShellLevels {
  static constexpr long long E = (long long)NH * NH * NH, NE = (E + 127) / 128 * 128;
  static constexpr long long C = 3 * E, NC = (C + 127) / 128 * 128;  // P0, P3
  static constexpr long long OFF[4] = {0, NC, NC + NE, NC + 2 * NE};
  static constexpr long long CNT[4] = {C, E, E, C};
};

// Lanes map to (row, k) with a power-of-two lane count per row (KQ for the
// KH mirrored pairs, QQ for the NH/4 output groups of four complex values), so
// no index divides.
//
// lv != 0 (whole mesh: nrows = N^2, cell0 = 0): each shell is written in the
// sublattice level layout instead of C order (ShellLevels: a permutation of
// the cells into four parts, zero-padded), only the parts p with bit p of
// parts[shell] set (all parts without the array).
template <typename T, int NH, int V, int P, int W>
__global__ __launch_bounds__(W) void k_inv_z(const cplx<T>* __restrict__ in, void* out,
                                             void* out_lo, const cplx<T>* __restrict__ twg,
                                             int MZ, int MZP, long long nrows, long long cell0,
                                             long long shell_stride, int fmt,
                                             const int* __restrict__ n2hi, int shell0, int lv,
                                             const unsigned char* __restrict__ parts) {
  constexpr int N = 2 * NH;
  constexpr int LD = pencil_ld<NH>();
  constexpr int KH = NH / 2 + 1;  // k = 0..NH/2 (each with its mirror NH - k)
  // lanes per row for the pairs: at most the workgroup, each lane then
  // taking KS pairs of its row (KS = 1 unless NH/2 + 1 > W)
  constexpr int KQ = pow2_at_least(KH) < W ? pow2_at_least(KH) : W;
  constexpr int KS = (KH + KQ - 1) / KQ, RK = W / KQ, NPK = P / RK;
  constexpr int QQ = pow2_at_least(NH / 4) > W / P ? pow2_at_least(NH / 4) : W / P;
  constexpr int RQ = W / QQ, NPQ = P / RQ;
  using Lay = ShellLevels<NH>;
  static_assert(NH % 4 == 0, "z c2r stores need N divisible by 8");
  static_assert(P % RK == 0 && P % RQ == 0, "row mapping of the z c2r");
  __shared__ cplx<T> buf[P * LD];
  // the length-NH transform's table, tw[t] = twg[2t] = exp(-2 pi i t / NH),
  // only the entries its plan reads (77 of the 200 of twg at N = 200, which
  // keeps the FP32 workgroup under 1/6 of the CU's LDS: six workgroups per CU
  // instead of five); the mirrored-pair twiddles exp(+2 pi i k / N) come
  // from twg itself
  constexpr int TWS = tw_span<NH>(typename Plan<NH, V>::seq{});
  __shared__ cplx<T> tw[TWS];
  const int tid = threadIdx.x;
  stage_in<T, TWS, W>([&](int i) { return twg[2 * i]; }, [&](int i, cplx<T> v) { tw[i] = v; });
  __syncthreads();
  const long long row0 = (long long)blockIdx.x * P;
  const long long bl = blockIdx.y;
  const int mzb = shell_mz(MZ, n2hi[shell0 + bl]);  // kz from mzb on are zero
  const cplx<T>* src = in + (bl * nrows + row0) * MZP;
  const int esz = out_esize<T>(fmt);
  // uniform base of the level layout: this shell's row
  char* const shi = static_cast<char*>(out) + bl * shell_stride * esz;
  char* const slo = out_lo ? static_cast<char*>(out_lo) + bl * shell_stride * esz : nullptr;
  const int pm = parts ? parts[bl] : 0xf;  // parts written of this shell (level layout)
  if (lv && blockIdx.x == 0)  // zero padding of the parts written
#pragma unroll
    for (int part = 0; part < 4; ++part) {
      if (!(pm >> part & 1)) continue;
      const long long end = part + 1 < 4 ? Lay::OFF[part + 1] : Lay::OFF[3] + Lay::NC;
      for (long long e = Lay::OFF[part] + Lay::CNT[part] + tid; e < end; e += W)
        store_zero<T>(fmt, shi, slo, (unsigned)e);
    }
  const int nr = (int)(nrows - row0 < P ? nrows - row0 : P);
  // Pencil p transforms row row0 + prow(p), p < np.  All nr rows, or, for a
  // shell whose parts hold only E and O, the rows with x + y even (x + y odd
  // feeds neither; y = row mod 2 as N is even): n1 of them in the x plane of
  // row0 from r1, the rest in the next plane from r2 (at most one plane
  // boundary per workgroup as N >= P), gathered into the first pencils.
  const bool gather = lv && !(pm & 0x9) && N >= P;
  int np = nr, n1 = nr, r1 = 0, r2 = 0, step = 1;
  if (gather) {
    const long long x0 = row0 / N, b = (x0 + 1) * N, end = row0 + nr;
    const long long e1 = b < end ? b : end;
    const int f1 = (int)((x0 + row0) & 1), f2 = (int)((x0 + 1 + b) & 1);
    n1 = e1 - row0 > f1 ? (int)((e1 - row0 - f1 + 1) / 2) : 0;
    const int n2 = end - b > f2 ? (int)((end - b - f2 + 1) / 2) : 0;
    r1 = f1;
    r2 = (int)(b - row0) + f2;
    np = n1 + n2;
    step = 2;
    if (np == 0) return;  // uniform
  }
  auto prow = [&](int p) { return p < n1 ? r1 + step * p : r2 + step * (p - n1); };
  // (k, NH - k) pairs: lane k of rows r0, r0 + RK, ...; up to 8 rows (16
  // loads) in flight per batch
  // with at least a wave of lanes per row every lane of a wave has the same
  // row: make that explicit so the row index math runs in scalar registers
  const int r0 = KQ >= 64 ? __builtin_amdgcn_readfirstlane(tid / KQ) : tid / KQ;
  constexpr int B = NPK < 8 ? NPK : 8;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = (tid & (KQ - 1)) + ks * KQ, j = NH - k;
    const bool kact = k < KH;
    const cplx<T> wk = kact ? cconj(twg[k]) : cplx<T>{(T)0, (T)0};  // exp(+2 pi i k / N)
#pragma unroll
    for (int q0 = 0; q0 < NPK; q0 += B) {
      cplx<T> xk[B], xj[B];
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const int p = r0 + (q0 + q) * RK;
        const bool ok = q0 + q < NPK && kact && p < np;
        const unsigned r = (unsigned)(prow(p) * MZP);
        xk[q] = (ok && k < mzb) ? get_u32(src, r + k) : cplx<T>{(T)0, (T)0};
        xj[q] = (ok && j < mzb) ? get_u32(src, r + j) : cplx<T>{(T)0, (T)0};
      }
#pragma unroll
      for (int q = 0; q < B; ++q)
        if (q0 + q < NPK && kact)
          c2r_pair<T>(buf + (r0 + (q0 + q) * RK) * LD, k, NH, xk[q], xj[q], wk);
    }
  }
  __syncthreads();
  if (gather)
    fft_lds_n<T, NH, true, P, W, LD, NH>(
        buf, tw, np, div_magic(np), tid,
        typename Plan<NH, V>::seq{});
  else
    fft_lds<T, NH, true, P, W, LD, NH>(buf, tw, typename Plan<NH, V>::seq{});
  // four complex values (eight reals, z = 8q .. 8q + 7) per lane and store:
  // in the level layout the four even z and the four odd z are each four
  // consecutive cells of one part
  const SplitConsts sc = split_consts();
  const long long obase = (bl * shell_stride + cell0 + row0 * N) * esz;
  char* const ohi = static_cast<char*>(out) + obase;
  char* const olo = out_lo ? static_cast<char*>(out_lo) + obase : nullptr;
  const int q = tid & (QQ - 1);
  const int rq = QQ >= 64 ? __builtin_amdgcn_readfirstlane(tid / QQ) : tid / QQ;  // see r0
  if (q < NH / 4) {
#pragma unroll
    for (int i = 0; i < NPQ; ++i) {
      const int p = rq + i * RQ;
      if (p >= np) continue;
      const cplx<T>* b = buf + p * LD + 4 * q;
      const cplx<T> v0 = b[0], v1 = b[1], v2 = b[2], v3 = b[3];
      if (!lv) {
        store_quad<T>(fmt, ohi, olo, (unsigned)(p * N + 8 * q), v0, v1, sc);
        store_quad<T>(fmt, ohi, olo, (unsigned)(p * N + 8 * q + 4), v2, v3, sc);
        continue;
      }
      // z = 8q + 2t (vt.x) and 8q + 2t + 1 (vt.y), t < 4: z/2 = 4q + t in the
      // part of each z parity
      const int row = (int)row0 + prow(p), x = row / N, y = row - x * N;
      const int xy = (x & 1) * 2 + (y & 1);  // 0: both even, 3: both odd
      const unsigned pos = (unsigned)(((x >> 1) * NH + (y >> 1)) * NH + 4 * q);
#pragma unroll
      for (int zp = 0; zp < 2; ++zp) {
        int part;
        unsigned off;
        if (((x + y + zp) & 1) == 0) {  // C0: E if all even, else P0
          part = (xy == 0 && zp == 0) ? 1 : 0;
          off = part ? (unsigned)Lay::OFF[1]
                     : (unsigned)(Lay::OFF[0] + (long long)(xy - 1) * Lay::E);
        } else {  // C1: O if all odd, else P3
          part = (xy == 3 && zp == 1) ? 2 : 3;
          off = part == 2 ? (unsigned)Lay::OFF[2]
                          : (unsigned)(Lay::OFF[3] + (long long)xy * Lay::E);
        }
        if (pm >> part & 1) {
          if (zp)
            store_four<T>(fmt, shi, slo, off + pos, v0.y, v1.y, v2.y, v3.y, sc);
          else
            store_four<T>(fmt, shi, slo, off + pos, v0.x, v1.x, v2.x, v3.x, sc);
        }
      }
    }
  }
}
