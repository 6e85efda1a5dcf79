// Split-BF16 MFMA core shared by the pair kernels (mfma.hip, zfused.hip):
// the product splits, one K step of a wave's item list and the partial tiles.
//
// A wave runs a fixed list of NI items (descriptors in SGPRs: tile | kb << 16
// | padding << 29 | new-tile << 31).  Lane (col = lane%16, grp = lane/16)
// forms the products of pair row col over its eight cells of a 32-cell K
// step: cells c..c+3 and c+16..c+19 of the FP32 pair-side rows (c = 32ks +
// 4grp), skipped when the item continues the previous item's tile; it reads
// its third-side column 16*kb + col from the bf16 planes, where each K step
// is stored permuted so that lane group g finds the same eight cells
// contiguous at 8g (cell 4g+e at 8g+e, cell 16+4g+e at 8g+4+e); then three
// MFMAs hi*hi + hi*lo + lo*hi.  FP32 chains run for at most kChain K
// steps and are then added into FP64 accumulators.
#pragma once
#include <type_traits>

#include "launch.h"

namespace pairmma {

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

// The bf16 pairs (-1, 0), (0, -1) as dot2 operands for the product splits,
// made once per thread and kept in registers (as inline constants the packed
// pairs are not encodable).
struct PairConsts {
  bf16x2 neg_lo, neg_hi;
};

BSK_DEV PairConsts pair_consts() {
  return PairConsts{bf16_pair_bits(kBf16NegOneLo), bf16_pair_bits(kBf16NegOneHi)};
}

// hi = rne_bf16(v), lo = rne_bf16(v - hi), packed two values per word.
// v - hi comes from v_dot2_f32_bf16 (hi.x * -1 + hi.y * 0 + v): one op per
// value instead of an unpack and a subtraction, and exact (v - hi is
// representable in FP32).
BSK_DEV void split2(float a, float b, uint32_t& hi, uint32_t& lo, const PairConsts& k) {
  const f32x2 v = {a, b};
  const bf16x2 h = __builtin_convertvector(v, bf16x2);
  const f32x2 r = {__builtin_amdgcn_fdot2_f32_bf16(h, k.neg_lo, a, false),
                   __builtin_amdgcn_fdot2_f32_bf16(h, k.neg_hi, b, false)};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

constexpr int kChain = 8;  // K steps per FP32 accumulation chain

template <int NI>
struct State {
  f32x4 acc[NI];
  double acc2[NI][4];
  int since_flush;
};

template <int NI>
BSK_DEV void init(State<NI>& st) {
#pragma unroll
  for (int s = 0; s < NI; ++s) {
    st.acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 4; ++r) st.acc2[s][r] = 0.0;
  }
  st.since_flush = 0;
}

// Descriptors and partial positions of wave `wave` and, per item, the LDS
// offsets (floats, both < 2^16) of this lane's pair rows in FP32 rows of LDF
// floats: table rows are shells, staged at row shell - lo; 0xffff (padding)
// is the zero row zrow.
template <int NI, int LDF>
BSK_DEV void setup(int (&desc)[NI], int (&oij)[NI], int (&opos)[NI], const int* items,
                   const int* outpos, const int* pairs, int wave, int col, int lo, int zrow) {
#pragma unroll
  for (int s = 0; s < NI; ++s) {
    desc[s] = __builtin_amdgcn_readfirstlane(items[wave * NI + s]);
    opos[s] = __builtin_amdgcn_readfirstlane(outpos[wave * NI + s]);
    const int tile = desc[s] & 0xffff;
    const int pr = pairs[tile * 16 + col];
    const int ra = pr & 0xffff, rb = (pr >> 16) & 0xffff;
    const int ia = ra == 0xffff ? zrow : ra - lo, ib = rb == 0xffff ? zrow : rb - lo;
    oij[s] = (ia * LDF) | ((ib * LDF) << 16);
  }
}

// One K step: c = 32ks + 4grp into the FP32 rows, cb = 32ks + 8grp into the
// permuted bf16 rows of LDH elements.
template <int NI, int LDH>
BSK_DEV void kstep(State<NI>& st, const int (&desc)[NI], const int (&oij)[NI], const float* fA,
                   const float* fB, const uint16_t* hK, const uint16_t* lK, int c, int cb,
                   int col, const PairConsts& kc) {
  bf16x8 Ah, Al;
  static_for<0, NI>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    if (desc[s] & (1 << 29)) return;  // uniform: padding item
    if (s == 0 || desc[s] < 0) {      // uniform: new tile -> new pair products
      const float4* pi = reinterpret_cast<const float4*>(fA + (oij[s] & 0xffff) + c);
      const float4* pj = reinterpret_cast<const float4*>(fB + ((unsigned)oij[s] >> 16) + c);
      const float4 x0 = pi[0], x1 = pi[4], y0 = pj[0], y1 = pj[4];
      uint32_t ah[4], al[4];
      split2(x0.x * y0.x, x0.y * y0.y, ah[0], al[0], kc);
      split2(x0.z * y0.z, x0.w * y0.w, ah[1], al[1], kc);
      split2(x1.x * y1.x, x1.y * y1.y, ah[2], al[2], kc);
      split2(x1.z * y1.z, x1.w * y1.w, ah[3], al[3], kc);
      Ah = __builtin_bit_cast(bf16x8, ah);
      Al = __builtin_bit_cast(bf16x8, al);
    }
    const int kb = (desc[s] >> 16) & 0xff;
    const int ko = (16 * kb + col) * LDH + cb;
    const bf16x8 Bh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(hK + ko));
    const bf16x8 Bl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(lK + ko));
    f32x4 t = st.acc[s];
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, t, 0, 0, 0);
    st.acc[s] = t;
  });
  if (++st.since_flush == kChain) {
    st.since_flush = 0;
#pragma unroll
    for (int s = 0; s < NI; ++s) {
#pragma unroll
      for (int r = 0; r < 4; ++r) st.acc2[s][r] += (double)st.acc[s][r];
      st.acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  }
}

// Partial tiles of this wave: C[row][col], row = 4*grp + r (pair row of the
// tile), col = lane%16 (third side), at out[s][16 * row + col].
template <int NI>
BSK_DEV void store(State<NI>& st, float* out, int grp, int col) {
#pragma unroll
  for (int s = 0; s < NI; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double v = st.acc2[s][r] + (double)st.acc[s][r];
      out[s * 256 + (4 * grp + r) * 16 + col] = (float)v;
    }
}

// The same tiles at item positions pos[s] of the workgroup's partials
// (out + 256 pos[s]; pos < 0: padding, nothing stored), written or added to
// what an earlier launch over other cells stored there.
template <int NI>
BSK_DEV void store_at(State<NI>& st, float* out, const int (&pos)[NI], bool add, int grp,
                      int col) {
#pragma unroll
  for (int s = 0; s < NI; ++s) {
    if (pos[s] < 0) continue;  // uniform
    float* o = out + (long long)pos[s] * 256;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double v = st.acc2[s][r] + (double)st.acc[s][r];
      float& d = o[(4 * grp + r) * 16 + col];
      d = add ? (float)((double)d + v) : (float)v;
    }
  }
}

}  // namespace pairmma
