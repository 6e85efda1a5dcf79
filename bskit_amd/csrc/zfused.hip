// Split-BF16 MFMA triple sums fed straight from the inverse y output: the
// z c2r of every shell runs inside the triple-sum workgroup, so the shell
// fields never go through HBM.
//
// A workgroup owns a contiguous range of (x, y) rows of the S^3 shell mesh.
// Per row:
//  * the D2 row of every shell (kz < mz_b complex values, prefetched into
//    registers one row ahead) becomes the NH-point complex input of the
//    mirrored-pair c2r (the same arithmetic as the standalone z pass), one
//    pencil per shell in the FP32 rows of LDS;
//  * one batched Stockham transform over the shells (pencil count known at
//    run time) leaves the S real values of each shell row in those FP32
//    rows, which are the pair-side operands as they stand;
//  * a split pass writes the bf16 hi/lo planes of the third side, each
//    32-cell K step permuted as the MFMA core expects;
//  * ceil(S/32) K steps of the wave item lists (pair_mma.h); the cells past
//    S are zero in both operands.
// Only auto spectra (one field, pairs and third side from the same shells),
// FP32 shells, S <= 256, at most kZMaxShells shells.  Partials and their
// reduction are those of k_triple_pairs.
#include "pair_mma.h"

#ifndef BSK_ZVARIANT
#error "compile with -DBSK_ZVARIANT=0 or 1 (radix plan of the z transform)"
#endif

namespace {

using namespace pairmma;

constexpr int kZT = 64 * kPairWaves;  // threads per workgroup

template <int S>
struct ZGeo {
  static constexpr int NH = S / 2;              // z transform length
  static constexpr int KH = NH / 2 + 1;         // mirrored pairs (k, NH - k) per row
  static constexpr int KS = (S + 31) / 32;      // K steps per row
  static constexpr int CELLS = 32 * KS;
  static constexpr int LDA = CELLS + 8;         // FP32 row stride (floats)
  static constexpr int LDC = LDA / 2;           // the same rows as complex pencils
  static constexpr int LDH = CELLS + 16;        // bf16 row stride (elements)
  static constexpr int QP = (kZMaxShells * KH + kZT - 1) / kZT;         // pair items / thread
  static constexpr int QS = (kZMaxShells * (S / 4) + kZT - 1) / kZT;    // split quads / thread
  static_assert(NH % 2 == 0, "S divisible by 4");
};

// cell c (a multiple of 4) of a K step -> its position in the permuted bf16
// row: cells 4m..4m+3 go to 8m (m < 4) or 8(m-4)+4 (m >= 4)
BSK_DEV int permuted_pos(int c) {
  const int m = (c & 31) >> 2;
  return (c & ~31) + (m < 4 ? 8 * m : 8 * (m - 4) + 4);
}

template <int NI, int S>
__global__ __launch_bounds__(kZT, 4) void k_triple_zrows(ZArgs a) {
  using G = ZGeo<S>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int nsh = a.nsh;
  const int krows = 16 * a.nkb;
  float* fA = reinterpret_cast<float*>(smem);  // (nsh + 1) rows; row nsh stays zero
  cplx<float>* zc = reinterpret_cast<cplx<float>*>(fA);
  uint16_t* hK = reinterpret_cast<uint16_t*>(fA + (nsh + 1) * G::LDA);
  uint16_t* lK = hK + krows * G::LDH;
  cplx<float>* tw = reinterpret_cast<cplx<float>*>(lK + krows * G::LDH);
  int* mzs = reinterpret_cast<int*>(tw + S);

  const int tid = threadIdx.x;
  const PairConsts kc = pair_consts();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, col = lane & 15, grp = lane >> 4;

  // zero once: padding cells of every row, the zero pair row, third-side
  // rows beyond nsh; the transforms and splits only ever write cells < S
  for (int e = tid; e < (nsh + 1) * G::LDA; e += kZT) fA[e] = 0.f;
  for (int e = tid; e < krows * G::LDH; e += kZT) {
    hK[e] = 0;
    lK[e] = 0;
  }
  for (int e = tid; e < S; e += kZT) tw[e] = a.tw[e];
  for (int e = tid; e < nsh; e += kZT) mzs[e] = shell_mz(a.MZ, a.n2hi[a.shell0 + e]);

  int desc[NI], oij[NI], opos[NI];
  setup<NI, G::LDA>(desc, oij, opos, a.items, a.outpos, a.pairs, wave, col, a.lo_row, nsh);
  State<NI> st;
  init(st);

  const long long r0 = a.nrows * blockIdx.x / a.nwg;
  const long long r1 = a.nrows * (blockIdx.x + 1) / a.nwg;
  const unsigned mdiv = div_magic(nsh);
  __syncthreads();

  // this thread's mirrored-pair items (shell b, k) are the same for all
  // rows; their index math is redone per row from an opaque copy of the
  // thread index (t), so that it is not held in registers across the row
  cplx<float> xk[G::QP], xj[G::QP];
  auto fetch_row = [&](long long r, int t) {
#pragma unroll
    for (int q = 0; q < G::QP; ++q) {
      const int i = t + q * kZT;
      const int b = i / G::KH, k = i - b * G::KH, j = G::NH - k;
      cplx<float> u = {0.f, 0.f}, w = {0.f, 0.f};
      if (b < nsh) {
        const int mz = mzs[b];
        const cplx<float>* src = a.d2 + ((long long)b * a.nrows + r) * a.MZP;
        if (k < mz) u = src[k];
        if (j < mz) w = src[j];
      }
      xk[q] = u;
      xj[q] = w;
    }
  };
  if (r0 < r1) fetch_row(r0, tid);
  for (long long r = r0; r < r1; ++r) {
    int t = tid;
    asm volatile("" : "+v"(t));
    __syncthreads();  // the previous row's K steps are done with fA, hK, lK
#pragma unroll
    for (int q = 0; q < G::QP; ++q) {
      const int i = t + q * kZT;
      const int b = i / G::KH, k = i - b * G::KH;
      if (b < nsh) c2r_pair<float>(zc + b * G::LDC, k, G::NH, xk[q], xj[q], cconj(tw[k]));
    }
    if (r + 1 < r1) fetch_row(r + 1, t);  // in flight during this row's transform and sums
    __syncthreads();
    fft_lds_n<float, G::NH, true, kZMaxShells, kZT, G::LDC, S>(
        zc, tw, nsh, mdiv, t, typename Plan<G::NH, BSK_ZVARIANT>::seq{});
#pragma unroll
    for (int q = 0; q < G::QS; ++q) {
      const int i = t + q * kZT;
      const int b = i / (S / 4), c = 4 * (i - b * (S / 4));
      if (b < nsh) {
        const float4 v = *reinterpret_cast<const float4*>(fA + b * G::LDA + c);
        uint32_t h0, l0, h1, l1;
        split2(v.x, v.y, h0, l0, kc);
        split2(v.z, v.w, h1, l1, kc);
        const int pos = (b + a.koff) * G::LDH + permuted_pos(c);  // third-side row of shell b
        *reinterpret_cast<uint2*>(hK + pos) = make_uint2(h0, h1);
        *reinterpret_cast<uint2*>(lK + pos) = make_uint2(l0, l1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < G::KS; ++ks)
      kstep<NI, G::LDH>(st, desc, oij, fA, fA, hK, lK, ks * 32 + grp * 4, ks * 32 + grp * 8,
                        col, kc);
  }
  store_at(st, a.partial + (long long)blockIdx.x * a.item_total * 256, opos, false, grp, col);
}

template <int NI, int S>
int launch_zrows(const ZArgs& a, hipStream_t s) {
  auto fn = k_triple_zrows<NI, S>;
  const size_t lds = bsk_zrows_lds_bytes(S, a.nsh, a.nkb);
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  BSK_HIP_CHECK(attr);
  hipLaunchKernelGGL(fn, dim3(a.nwg), dim3(kZT), lds, s, a);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

template <int S>
int launch_zrows_ni(const ZArgs& a, int ni, hipStream_t s) {
  switch (ni) {
    case 1: return launch_zrows<1, S>(a, s);
    case 2: return launch_zrows<2, S>(a, s);
    case 3: return launch_zrows<3, S>(a, s);
    case 4: return launch_zrows<4, S>(a, s);
    default:
      bsk_set_error("fused z/triple sums: %d items per wave not compiled", ni);
      return BSK_EUNSUPPORTED;
  }
}

}  // namespace

#if BSK_ZVARIANT == 0
size_t bsk_zrows_lds_bytes(int S, int nsh, int nkb) {
  const size_t cells = (size_t)32 * ((S + 31) / 32);
  return (size_t)(nsh + 1) * (cells + 8) * 4 + (size_t)2 * 16 * nkb * (cells + 16) * 2 +
         (size_t)S * 8 + (size_t)nsh * 4;
}
#define BSK_ZROWS_FN bsk_triple_zrows_launch_v0
#else
#define BSK_ZROWS_FN bsk_triple_zrows_launch_v1
#endif

int BSK_ZROWS_FN(const ZArgs& a, int S, int ni, hipStream_t s) {
  switch (S) {
#define C(n) \
  case n:    \
    return launch_zrows_ni<n>(a, ni, s);
    BSK_ZFUSED_SIZES(C)
#undef C
    default:
      bsk_set_error("fused z/triple sums: no kernel for shell mesh %d", S);
      return BSK_EUNSUPPORTED;
  }
}
