// Split-BF16 MFMA triple sums:  B[i][j][k] = sum_x (d_i d_j)(x) d_k(x).
//
// For each i the sum over the cells is the GEMM (D o d_i) . D^T: rows j of
// the A operand are the pair products d_i(x) d_j(x), columns k of the B
// operand the shells d_k(x), K = cells.  Operands are split into bf16
// hi + lo (hi = rne(v), lo = rne(v - hi)) and multiplied on
// v_mfma_f32_16x16x32_bf16 as hi*hi + hi*lo + lo*hi.
//
// Workgroup structure (one code path for every wave):
//  * the workgroup owns a contiguous range of cell chunks (CK cells each);
//    per chunk all shells' hi/lo planes are loaded from HBM once, coalesced,
//    into LDS as FP32 d = hi + lo (A side) and as the raw bf16 planes
//    (B side); padding rows (shell index >= nshell) are zero;
//  * a 16x16 tile (i, jb, kb) holds C[j][k] for j in block jb, k in block
//    kb; wave w owns the i-rows 16*ib + w*R(ib) + q (q < R(ib)) and every
//    tile of those rows (auto: jb <= ib, kb <= jb; <AAB>: jb <= ib, all kb;
//    <ABC>: all jb, kb).  Lane (col = lane%16, grp = lane/16) works on shell
//    16*b + col and cells 8*grp..8*grp+7 of each 32-cell K step, so its A
//    row j and B column k are its own LDS values; d_i comes from a broadcast
//    LDS read;
//  * MFMA accumulates in FP32 for at most F K steps; the tile is then added
//    into an FP64 (or FP32) second-level accumulator.  Long FP32 chains lose
//    accuracy because per-region partial sums are spatially coherent and
//    large compared with the globally cancelling total.
// Per-workgroup tiles are written as FP32 partials [wg][tile][16][16]; a
// second kernel sums them over workgroups in a fixed order in FP64.
#include <type_traits>

#include "launch.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

template <int B, int E, typename F>
BSK_DEV void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// hi = rne_bf16(v), lo = rne_bf16(v - hi), packed two values per word.
BSK_DEV void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
  const f32x2 v = {a, b};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
  const f32x2 r = {a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

BSK_DEV float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
BSK_DEV float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

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

// Second-level accumulation of the FP32 MFMA chains.
enum { ACC_F64 = 0, ACC_F32X2 = 1, ACC_F32 = 2 };

template <int NB, int MODE, int WAVES, int RT, int CK, int F, int ACC, int XS>
struct MfmaK {
  static constexpr int NT = 64 * WAVES;
  static constexpr int RF = 16 / WAVES;   // i-rows per wave in a full block
  static constexpr int NSH = 16 * NB;     // padded shell rows per set (block part)
  static constexpr int NSHF = NSH + (XS ? 16 : 0);  // FP32 rows incl. extra i-rows
  static constexpr int LDF = CK + 4;      // FP32 row stride (floats)
  static constexpr int LDH = CK + 8;      // bf16 row stride (elements)
  static constexpr bool SEP_I = MODE == 2;
  static constexpr int FS = SEP_I ? 2 : 1;
  static constexpr int KSTEPS = CK / 32;
  static constexpr int SEGS = CK / 8;
  static constexpr size_t LDS_BYTES =
      (size_t)FS * NSHF * LDF * 4 + (size_t)2 * NSH * LDH * 2;
  static constexpr int rows(int ib) { return ib == NB - 1 ? RT : RF; }
  static constexpr int slot_base(int ib) {
    int s = 0;
    for (int b = 0; b < ib; ++b) s += rows(b) * ntiles_i<MODE, NB>(b);
    return s;
  }
  static constexpr int NSLOT = slot_base(NB);
  static constexpr int TILES = 16 * tiles_before<MODE, NB>(NB);
  using acc2_t = typename std::conditional<ACC == ACC_F64, double, float>::type;
  static constexpr int NSLOT2 = ACC == ACC_F32 ? 1 : NSLOT;  // single level: unused

  static BSK_DEV void zero_rows(float* f, int r0, int ld) {
    for (int e = threadIdx.x; e < (NSHF - r0) * ld; e += NT) f[r0 * ld + e] = 0.f;
  }
  static BSK_DEV void zero_rows_h(uint16_t* h, int r0) {
    for (int e = threadIdx.x; e < (NSH - r0) * LDH; e += NT) h[r0 * LDH + e] = 0;
  }

  // Per-thread share of one staged shell set (16-byte segments of the hi and
  // lo planes): loaded from HBM one chunk ahead, stored to LDS afterwards.
  static constexpr int PV = (NSHF * SEGS + NT - 1) / NT;
  struct Pref {
    uint4 h[PV], l[PV];
  };

  static BSK_DEV void fetch(Pref& p, const uint16_t* __restrict__ gh,
                            const uint16_t* __restrict__ gl, long long stride, int n,
                            long long cbase) {
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int v = threadIdx.x + k * NT;
      if (v < n * SEGS) {
        const int row = v / SEGS, sg = v - row * SEGS;
        const long long g = (long long)row * stride + cbase + sg * 8;
        p.h[k] = *reinterpret_cast<const uint4*>(gh + g);
        p.l[k] = *reinterpret_cast<const uint4*>(gl + g);
      }
    }
  }

  // FP32 values to f (if non-null) and the raw planes of rows < nh to h/l
  // (if non-null).
  static BSK_DEV void put(const Pref& p, int n, float* f, uint16_t* h, uint16_t* l, int nh) {
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int v = threadIdx.x + k * NT;
      if (v < n * SEGS) {
        const int row = v / SEGS, sg = v - row * SEGS;
        const uint4 hv = p.h[k], lv = p.l[k];
        if (f) {
          float4* dst = reinterpret_cast<float4*>(f + row * LDF + sg * 8);
          dst[0] = make_float4(bf_lo(hv.x) + bf_lo(lv.x), bf_hi(hv.x) + bf_hi(lv.x),
                               bf_lo(hv.y) + bf_lo(lv.y), bf_hi(hv.y) + bf_hi(lv.y));
          dst[1] = make_float4(bf_lo(hv.z) + bf_lo(lv.z), bf_hi(hv.z) + bf_hi(lv.z),
                               bf_lo(hv.w) + bf_lo(lv.w), bf_hi(hv.w) + bf_hi(lv.w));
        }
        if (h && row < nh) {
          *reinterpret_cast<uint4*>(h + row * LDH + sg * 8) = hv;
          *reinterpret_cast<uint4*>(l + row * LDH + sg * 8) = lv;
        }
      }
    }
  }

  static BSK_DEV void run(const MfmaArgs& a, unsigned char* smem) {
    float* fJ = reinterpret_cast<float*>(smem);
    float* fI = SEP_I ? fJ + NSHF * LDF : fJ;
    uint16_t* hK = reinterpret_cast<uint16_t*>(fJ + FS * NSHF * LDF);
    uint16_t* lK = hK + NSH * LDH;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    const int nJ = a.nsh[1];
    const int nI = SEP_I ? a.nsh[0] : nJ;
    const int nK = MODE == 0 ? (nJ < NSH ? nJ : NSH) : a.nsh[2];

    zero_rows(fJ, nJ, LDF);
    if (SEP_I) zero_rows(fI, nI, LDF);
    zero_rows_h(hK, nK);
    zero_rows_h(lK, nK);

    f32x4 acc[NSLOT];
    acc2_t acc2[NSLOT2][4];
    f32x4 xacc[XS ? XS : 1];
    acc2_t xacc2[XS ? XS : 1][4];
#pragma unroll
    for (int s = 0; s < (XS ? XS : 1); ++s) {
      xacc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) xacc2[s][r] = 0;
    }
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NSLOT2; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc2[s][r] = 0;

    const long long nch = a.ncell / CK;
    const long long ch0 = nch * blockIdx.x / a.nwg;
    const long long ch1 = nch * (blockIdx.x + 1) / a.nwg;
    int since_flush = 0;
    Pref pJ, pI, pK;
    if (ch0 < ch1) {
      fetch(pJ, a.hi[1], a.lo[1], a.stride[1], nJ, ch0 * CK);
      if (SEP_I) fetch(pI, a.hi[0], a.lo[0], a.stride[0], nI, ch0 * CK);
      if (MODE != 0) fetch(pK, a.hi[2], a.lo[2], a.stride[2], nK, ch0 * CK);
    }
    for (long long ch = ch0; ch < ch1; ++ch) {
      __syncthreads();  // every wave is done with the previous chunk
      put(pJ, nJ, fJ, MODE == 0 ? hK : nullptr, MODE == 0 ? lK : nullptr, nK);
      if (SEP_I) put(pI, nI, fI, nullptr, nullptr, 0);
      if (MODE != 0) put(pK, nK, nullptr, hK, lK, nK);
      __syncthreads();
      if (ch + 1 < ch1) {  // next chunk's HBM loads overlap this chunk's math
        const long long nb = (ch + 1) * CK;
        fetch(pJ, a.hi[1], a.lo[1], a.stride[1], nJ, nb);
        if (SEP_I) fetch(pI, a.hi[0], a.lo[0], a.stride[0], nI, nb);
        if (MODE != 0) fetch(pK, a.hi[2], a.lo[2], a.stride[2], nK, nb);
      }
#pragma unroll 1
      for (int ks = 0; ks < KSTEPS; ++ks) {
        const int c = ks * 32 + grp * 8;
        float dJ[NB][8];
        uint4 kh[NB], kl[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const float4* src = reinterpret_cast<const float4*>(fJ + (16 * b + col) * LDF + c);
          const float4 x0 = src[0], x1 = src[1];
          dJ[b][0] = x0.x; dJ[b][1] = x0.y; dJ[b][2] = x0.z; dJ[b][3] = x0.w;
          dJ[b][4] = x1.x; dJ[b][5] = x1.y; dJ[b][6] = x1.z; dJ[b][7] = x1.w;
          kh[b] = *reinterpret_cast<const uint4*>(hK + (16 * b + col) * LDH + c);
          kl[b] = *reinterpret_cast<const uint4*>(lK + (16 * b + col) * LDH + c);
        }
        static_for<0, NB>([&](auto ibc) {
          constexpr int ib = decltype(ibc)::value;
          static_for<0, rows(ib)>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            // rows i >= nI are zero in LDS and contribute nothing; no branch
            // here keeps the accumulators in place across the whole K step
            const int i = 16 * ib + wave * rows(ib) + q;
            const float4* srci = reinterpret_cast<const float4*>(fI + i * LDF + c);
            const float4 y0 = srci[0], y1 = srci[1];
            const float di[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
            static_for<0, jb_end<MODE, NB>(ib)>([&](auto jbc) {
              constexpr int jb = decltype(jbc)::value;
              uint32_t ah[4], al[4];
#pragma unroll
              for (int e = 0; e < 4; ++e)
                split2(dJ[jb][2 * e] * di[2 * e], dJ[jb][2 * e + 1] * di[2 * e + 1], ah[e], al[e]);
              const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah);
              const bf16x8 Al = __builtin_bit_cast(bf16x8, al);
              static_for<0, kb_end<MODE, NB>(jb)>([&](auto kbc) {
                constexpr int kb = decltype(kbc)::value;
                constexpr int s = slot_base(ib) + q * ntiles_i<MODE, NB>(ib) +
                                  tile_in_i<MODE, NB>(jb, kb);
                const bf16x8 Bh = __builtin_bit_cast(bf16x8, kh[kb]);
                const bf16x8 Bl = __builtin_bit_cast(bf16x8, kl[kb]);
                f32x4 t = acc[s];
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, t, 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, t, 0, 0, 0);
                acc[s] = t;
              });
            });
          });
        });
        // extra tiles (rows i beyond the blocks): operands by runtime index
        static_for<0, XS>([&](auto sc) {
          constexpr int xs = decltype(sc)::value;
          const int t = wave + WAVES * xs;
          if (t >= a.nextra) return;  // uniform
          const int i = a.xt[t][0], jb = a.xt[t][1], kb = a.xt[t][2];
          const float4* srci = reinterpret_cast<const float4*>(fI + i * LDF + c);
          const float4* srcj = reinterpret_cast<const float4*>(fJ + (16 * jb + col) * LDF + c);
          const float4 y0 = srci[0], y1 = srci[1], z0 = srcj[0], z1 = srcj[1];
          uint32_t ah[4], al[4];
          split2(z0.x * y0.x, z0.y * y0.y, ah[0], al[0]);
          split2(z0.z * y0.z, z0.w * y0.w, ah[1], al[1]);
          split2(z1.x * y1.x, z1.y * y1.y, ah[2], al[2]);
          split2(z1.z * y1.z, z1.w * y1.w, ah[3], al[3]);
          const uint4 bh = *reinterpret_cast<const uint4*>(hK + (16 * kb + col) * LDH + c);
          const uint4 bl = *reinterpret_cast<const uint4*>(lK + (16 * kb + col) * LDH + c);
          const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
          const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
          f32x4 u = xacc[xs];
          u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, u, 0, 0, 0);
          u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, u, 0, 0, 0);
          u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, u, 0, 0, 0);
          xacc[xs] = u;
        });
        if (ACC != ACC_F32 && ++since_flush == F) {
          since_flush = 0;
#pragma unroll
          for (int s = 0; s < XS; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xacc2[s][r] += (acc2_t)xacc[s][r];
            xacc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
          }
#pragma unroll
          for (int s = 0; s < NSLOT; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc2[s % NSLOT2][r] += (acc2_t)acc[s][r];
            acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
          }
        }
      }
    }
    if (ACC != ACC_F32) {
#pragma unroll
      for (int s = 0; s < NSLOT; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc2[s % NSLOT2][r] += (acc2_t)acc[s][r];
#pragma unroll
      for (int s = 0; s < XS; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) xacc2[s][r] += (acc2_t)xacc[s][r];
    }

    // write the tiles: C[row][col] with row = 4*grp + r (j), col = lane%16 (k)
    float* out = a.partial + (long long)blockIdx.x * (TILES + a.nextra) * 256;
    static_for<0, XS>([&](auto sc) {
      constexpr int xs = decltype(sc)::value;
      const int t = wave + WAVES * xs;
      if (t >= a.nextra) return;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[(long long)(TILES + t) * 256 + (4 * grp + r) * 16 + col] =
            ACC == ACC_F32 ? xacc[xs][r] : (float)xacc2[xs][r];
    });
    static_for<0, NB>([&](auto ibc) {
      constexpr int ib = decltype(ibc)::value;
      static_for<0, rows(ib)>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int ii = wave * rows(ib) + q;
        if (ii >= 16) return;
        static_for<0, jb_end<MODE, NB>(ib)>([&](auto jbc) {
          constexpr int jb = decltype(jbc)::value;
          static_for<0, kb_end<MODE, NB>(jb)>([&](auto kbc) {
            constexpr int kb = decltype(kbc)::value;
            constexpr int s = slot_base(ib) + q * ntiles_i<MODE, NB>(ib) +
                              tile_in_i<MODE, NB>(jb, kb);
            const int g = 16 * tiles_before<MODE, NB>(ib) + ii * ntiles_i<MODE, NB>(ib) +
                          tile_in_i<MODE, NB>(jb, kb);
#pragma unroll
            for (int r = 0; r < 4; ++r)
              out[(long long)g * 256 + (4 * grp + r) * 16 + col] =
                  ACC == ACC_F32 ? acc[s][r] : (float)acc2[s % NSLOT2][r];
          });
        });
      });
    });
  }
};

template <int NB, int MODE, int WAVES, int RT, int CK, int F, int ACC, int XS>
__global__ __launch_bounds__(64 * WAVES) void k_triple_mfma(MfmaArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  MfmaK<NB, MODE, WAVES, RT, CK, F, ACC, XS>::run(a, smem);
}

__global__ __launch_bounds__(64) void k_reduce_mfma(const float* __restrict__ partial, int nwg,
                                                    long long wg_stride, const int* __restrict__ slot,
                                                    double scale, double* __restrict__ out) {
  const int t = blockIdx.x;
  const long long e = slot[t];
  double s = 0.0;
  for (int w = threadIdx.x; w < nwg; w += 64) s += (double)partial[w * wg_stride + e];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) out[t] = s * scale;
}

constexpr int kCK = 128;  // cells per staged chunk
constexpr int kF = 8;     // K steps per FP32 accumulation chain

// Waves per workgroup: 8 (two per SIMD, 256 registers each) up to two
// 16-shell blocks; 4 (one per SIMD, 512 registers) for three blocks.
template <int NB>
constexpr int waves_of() {
  return NB <= 2 ? 8 : 4;
}

// Accumulation scheme that fits the register file for this tile count.
template <int NB, int MODE, int RT>
constexpr int acc_of() {
  constexpr int W = waves_of<NB>();
  constexpr int n = MfmaK<NB, MODE, W, RT, kCK, kF, ACC_F64, 0>::NSLOT;
  constexpr int budget = W == 8 ? 256 : 512;
  constexpr int base = W == 8 ? 72 : 200;  // operands, addresses, prefetch registers
  return (12 * n + base <= budget) ? ACC_F64 : ((8 * n + base <= budget) ? ACC_F32X2 : ACC_F32);
}

template <int NB, int MODE, int RT, int XS = 0>
int launch_rt(const MfmaArgs& a, hipStream_t s) {
  constexpr int W = waves_of<NB>();
  constexpr int ACC = acc_of<NB, MODE, RT>();
  using K = MfmaK<NB, MODE, W, RT, kCK, kF, ACC, XS>;
  auto fn = k_triple_mfma<NB, MODE, W, RT, kCK, kF, ACC, XS>;
  // once per instance (not per launch: keeps launches capturable and cheap)
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)K::LDS_BYTES);
  BSK_HIP_CHECK(attr);
  hipLaunchKernelGGL(fn, dim3(a.nwg), dim3(64 * W), K::LDS_BYTES, s, a);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

}  // namespace

int bsk_mfma_tiles(int nblk, int mode) {
  int s = 0;
  for (int ib = 0; ib < nblk; ++ib)
    s += mode == 0 ? (ib + 1) * (ib + 2) / 2 : (mode == 1 ? (ib + 1) * nblk : nblk * nblk);
  return 16 * s;
}

int bsk_mfma_supported(int nblk, int mode) {
  return nblk >= 1 && nblk <= 3 && mode >= 0 && mode <= 2 && !(nblk == 3 && mode == 2);
}

int bsk_triple_mfma_launch(const MfmaArgs& a, hipStream_t s) {
  if (a.ncell % kCK) {
    bsk_set_error("mfma triple: cell count %lld not a multiple of %d", a.ncell, kCK);
    return BSK_EINVAL;
  }
  if (a.nwg <= 0) {
    bsk_set_error("mfma triple: nwg must be positive");
    return BSK_EINVAL;
  }
  for (int st = 0; st < 3; ++st)
    if (a.stride[st] % 8) {
      bsk_set_error("mfma triple: shell strides must be multiples of 8 elements");
      return BSK_EINVAL;
    }
  if (a.nextra < 0 || a.nextra > kMfmaMaxExtra || (a.nextra && a.mode != 0)) {
    bsk_set_error("mfma triple: 0..%d extra tiles, auto mode only", kMfmaMaxExtra);
    return BSK_EINVAL;
  }
  // rows of the last (possibly partial) 16-shell block per wave
  const int nI = a.mode == 2 ? a.nsh[0] : a.nsh[1];
  const int tail = (a.nextra ? 16 * a.nblk : nI) - 16 * (a.nblk - 1);
  if (a.nextra) {  // the blocks are full; the extra i-rows ride on extra tiles
    if (a.nblk == 1) return launch_rt<1, 0, 2, 1>(a, s);
    if (a.nblk == 2) return launch_rt<2, 0, 2, 1>(a, s);
    bsk_set_error("mfma triple: extra tiles need 1 or 2 full blocks");
    return BSK_EUNSUPPORTED;
  }
#define BSK_MF(NB, MODE)                                                     \
  if (a.nblk == NB && a.mode == MODE) {                                      \
    constexpr int W = waves_of<NB>();                                        \
    const int rt = (tail + W - 1) / W;                                       \
    if (rt <= 1) return launch_rt<NB, MODE, 1>(a, s);                        \
    if (rt <= 2) return launch_rt<NB, MODE, 2>(a, s);                        \
    return launch_rt<NB, MODE, 16 / W>(a, s);                                \
  }
  BSK_MF(1, 0) BSK_MF(2, 0) BSK_MF(3, 0)
  BSK_MF(1, 1) BSK_MF(2, 1) BSK_MF(3, 1)
  BSK_MF(1, 2) BSK_MF(2, 2)
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
