// Split-BF16 MFMA triple sums over a list of shell pairs:
//   B[i][j][k] = sum_x (d_i d_j)(x) d_k(x).
//
// The host plan (capi.hip, bsk_pairs_create) turns a triangle list into the
// distinct pairs (i, j) it needs, packs them into 16-row tiles and, per tile,
// the 16-shell blocks kb of the third side it needs.  Each (tile, kb) is an
// item: a 16x16 output tile C[pair row][k in block kb] computed as the GEMM
// over the cells of A = (d_i d_j) (16 pair rows) and B = D_kb (16 shells),
// operands split into bf16 hi = rne(v), lo = rne(v - hi) and multiplied on
// v_mfma_f32_16x16x32_bf16 as hi*hi + hi*lo + lo*hi.  Only pairs that close
// a requested triangle are formed, so the VALU work (pair products and their
// splits) is proportional to the number of distinct pairs.
//
// Workgroup structure (kPairWaves waves, one code path):
//  * the workgroup owns a contiguous range of CK-cell chunks; per chunk the
//    hi/lo planes of every shell of each role set are read from HBM once
//    (coalesced, prefetched into registers one chunk ahead) and stored in
//    LDS as FP32 d = hi + lo (pair sides) and as raw bf16 planes (third
//    side); one zero FP32 row per pair set serves padding rows;
//  * wave w runs a fixed list of NI items (descriptors in SGPRs): lane
//    (col = lane%16, grp = lane/16) forms the products of pair row col over
//    its eight cells of each 32-cell K step (skipped when the item reuses
//    the previous item's tile), loads its B column 16*kb + col, and issues
//    three MFMAs; padding items do nothing (keeping the third-side fragment
//    live across items for reuse costs more in registers than it saves);
//  * FP32 MFMA chains run for at most kChain K steps and are then added into an
//    FP64 accumulator (long FP32 chains lose accuracy: per-region partial
//    sums are spatially coherent and large against the cancelling total).
// Per-workgroup tiles are written as FP32 partials [wg][item][16][16]; a
// second kernel sums them over workgroups in a fixed order in FP64.
#include <algorithm>

#include "pair_mma.h"

namespace {

using namespace pairmma;

// hi + lo of the low / high bf16 of a pair of words, as one dot2 each (the
// (1, 0) / (0, 1) operands are made per staged chunk, not kept live).
BSK_DEV float bf_sum_lo(uint32_t h, uint32_t l, bf16x2 one_lo) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, h), one_lo,
                                         __uint_as_float(l << 16), false);
}
BSK_DEV float bf_sum_hi(uint32_t h, uint32_t l, bf16x2 one_hi) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, h), one_hi,
                                         __uint_as_float(l & 0xffff0000u), false);
}

constexpr int kW = kPairWaves;   // waves per workgroup
constexpr int kNT = 64 * kW;
constexpr int kCK = kPairChunk;  // cells per staged chunk
// LDS layout (strides chosen by measured ds_read_b128 throughput on gfx950):
//  * FP32 pair-side rows of 136 floats (544 B), cells in natural order; in a
//    K step lane group g reads cells 4g..4g+3 and 16+4g..16+4g+3, the pattern
//    that serves 16 arbitrary rows with the fewest bank conflicts;
//  * bf16 third-side rows of 144 elements (288 B), each 32-cell K step stored
//    permuted so that lane group g finds the same eight cells contiguous at
//    8g: cell 4g+e at 8g+e and cell 16+4g+e at 8g+4+e (e < 4).  The MFMA K
//    order is then the same permutation on both operands.
constexpr int kLDF = kCK + 8;    // FP32 row stride (floats)
constexpr int kLDH = kCK + 16;   // bf16 row stride (elements)
constexpr int kSEGS = kCK / 8;   // 16-byte segments per row and plane
constexpr int kKSTEPS = kCK / 32;

// Per-thread share of one staged shell set: PV 16-byte segments of the hi
// and lo planes, loaded one chunk ahead and stored to LDS afterwards.
template <int PV>
struct Pref {
  uint4 h[PV], l[PV];
};

template <int PV>
BSK_DEV void fetch(Pref<PV>& p, const uint16_t* __restrict__ gh, const uint16_t* __restrict__ gl,
                   long long stride, int n, long long cbase) {
#pragma unroll
  for (int k = 0; k < PV; ++k) {
    const int v = threadIdx.x + k * kNT;
    if (v < n * kSEGS) {
      const int row = v / kSEGS, sg = v - row * kSEGS;
      const long long g = (long long)row * stride + cbase + sg * 8;
      p.h[k] = *reinterpret_cast<const uint4*>(gh + g);
      p.l[k] = *reinterpret_cast<const uint4*>(gl + g);
    }
  }
}

// The same loads issued by every thread (threads past the staged rows
// re-read row 0), so a wave always has the same number of loads in flight
// and a wait for an older chunk's loads can leave a newer chunk's pending.
template <int PV>
BSK_DEV void fetch_all(Pref<PV>& p, const uint16_t* __restrict__ gh,
                       const uint16_t* __restrict__ gl, long long stride, int n, long long cbase) {
#pragma unroll
  for (int k = 0; k < PV; ++k) {
    int v = threadIdx.x + k * kNT;
    v = v < n * kSEGS ? v : 0;
    const int row = v / kSEGS, sg = v - row * kSEGS;
    const long long g = (long long)row * stride + cbase + sg * 8;
    p.h[k] = *reinterpret_cast<const uint4*>(gh + g);
    p.l[k] = *reinterpret_cast<const uint4*>(gl + g);
  }
}

// FP32 values d = hi + lo to f (if non-null) and raw planes to h/l (if
// non-null) for rows < n.
template <int PV>
BSK_DEV void put(const Pref<PV>& p, int n, float* f, uint16_t* h, uint16_t* l) {
  const bf16x2 olo = bf16_pair_bits(kBf16OneLo), ohi = bf16_pair_bits(kBf16OneHi);
#pragma unroll
  for (int k = 0; k < PV; ++k) {
    const int v = threadIdx.x + k * kNT;
    if (v < n * kSEGS) {
      const int row = v / kSEGS, sg = v - row * kSEGS;
      const uint4 hv = p.h[k], lv = p.l[k];
      if (f) {
        float4* dst = reinterpret_cast<float4*>(f + row * kLDF + sg * 8);
        dst[0] = make_float4(bf_sum_lo(hv.x, lv.x, olo), bf_sum_hi(hv.x, lv.x, ohi),
                             bf_sum_lo(hv.y, lv.y, olo), bf_sum_hi(hv.y, lv.y, ohi));
        dst[1] = make_float4(bf_sum_lo(hv.z, lv.z, olo), bf_sum_hi(hv.z, lv.z, ohi),
                             bf_sum_lo(hv.w, lv.w, olo), bf_sum_hi(hv.w, lv.w, ohi));
      }
      if (h) {
        // cells 8s+4q..8s+4q+3 of the K step (s = sg % 4, q = 0, 1) go to
        // 16(s & 1) + 8q + 4(s >> 1)
        const int base = row * kLDH + (sg & ~3) * 8 + 16 * (sg & 1) + 4 * ((sg >> 1) & 1);
        *reinterpret_cast<uint2*>(h + base) = make_uint2(hv.x, hv.y);
        *reinterpret_cast<uint2*>(h + base + 8) = make_uint2(hv.z, hv.w);
        *reinterpret_cast<uint2*>(l + base) = make_uint2(lv.x, lv.y);
        *reinterpret_cast<uint2*>(l + base + 8) = make_uint2(lv.z, lv.w);
      }
    }
  }
}

// SETS: 1 = one field (pairs and third side from set 0), 2 = pairs from set
// 0, third side from set 2 (<AAB>), 3 = pairs from sets 0 and 1, third side
// from set 2 (<ABC>).
// Waves per SIMD the register allocation must allow: 4 (two workgroups per
// CU, at most 128 VGPRs) where the item, prefetch and role-set registers fit
// without spilling, else 2; 8 for one item and one prefetch segment (at most
// 64 VGPRs), whose launches stage few shells per chunk and need more
// workgroups in flight.
template <int NI, int PV, int SETS, bool D2>
constexpr int pair_occupancy() {
  if (NI == 1 && PV == 1 && SETS == 1) return 8;  // thin launches: 4 workgroups per CU
  if (NI == 2 && PV == 1 && SETS == 1) return 6;  // three workgroups per CU
  constexpr int ni_max = SETS == 1 ? 4 : (SETS == 2 ? 3 : 1);
  return (NI <= ni_max && PV <= 2) ? 4 : 2;
}

// D2 (one field, one prefetch segment per thread): two chunks in flight
// per workgroup instead of one, for segments whose few items leave the
// staging loads' latency exposed (replicated wave lists).
template <int NI, int PV, int SETS, bool D2>
__global__ __launch_bounds__(kNT, (pair_occupancy<NI, PV, SETS, D2>())) void k_triple_pairs(PairArgs a) {
  static_assert(!D2 || (SETS == 1 && PV == 1), "two-chunk prefetch: one field, one segment");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int n0 = a.nsh[0], n1 = a.nsh[1], n2 = a.nsh[2];
  float* fA = reinterpret_cast<float*>(smem);         // (n0 + 1) rows
  float* fB = SETS == 3 ? fA + (n0 + 1) * kLDF : fA;  // (n1 + 1) rows for <ABC>
  const int frows = (n0 + 1) + (SETS == 3 ? n1 + 1 : 0);
  uint16_t* hK = reinterpret_cast<uint16_t*>(fA + frows * kLDF);
  const int krows = 16 * a.nkb;
  uint16_t* lK = hK + krows * kLDH;

  const PairConsts kc = pair_consts();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
  // rep replicas of the item lists (waves w and w + kW/rep hold the same
  // list); replica rgrp runs the K steps ks with ks % rep == rgrp
  const int rep = a.rep, rgrp = wave / (kW / rep);
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;

  // zero rows: FP32 padding rows and the third-side rows outside the staged
  // shells [koff, koff + n2)
  for (int e = threadIdx.x; e < kLDF; e += kNT) {
    fA[n0 * kLDF + e] = 0.f;
    if (SETS == 3) fB[n1 * kLDF + e] = 0.f;
  }
  for (int e = threadIdx.x; e < krows * kLDH; e += kNT) {
    const int row = e / kLDH;
    if (row < a.koff || row >= a.koff + n2) {
      hK[e] = 0;
      lK[e] = 0;
    }
  }
  uint16_t* const hKs = hK + a.koff * kLDH;  // row of staged third-side shell 0
  uint16_t* const lKs = lK + a.koff * kLDH;

  // the cell range (lw) and item-list pass of this workgroup
  int lw = blockIdx.x, pass = 0;
  if (a.npass > 1) {
    pass = (blockIdx.x >> 3) % a.npass;
    lw = (blockIdx.x / (8 * a.npass)) * 8 + (blockIdx.x & 7);
  }
  // item descriptors and partial positions of this wave and the LDS rows of
  // this lane's pair
  int desc[NI], oij[NI], opos[NI];
  setup<NI, kLDF>(desc, oij, opos, a.items + pass * a.pass_stride,
                  a.outpos + pass * a.pass_stride, a.pairs, wave, col, a.lo_row, n0);
  State<NI> st;
  init(st);

  const long long nch = a.ncell / kCK;
  const long long ch0 = nch * lw / a.nwg;
  const long long ch1 = nch * (lw + 1) / a.nwg;
  auto ksteps = [&]() {
#pragma unroll  // constant LDS offsets per K step
    for (int ks = 0; ks < kKSTEPS; ++ks)
      if (ks % rep == rgrp)  // uniform: this replica's K steps
        kstep<NI, kLDH>(st, desc, oij, fA, fB, hK, lK, ks * 32 + grp * 4, ks * 32 + grp * 8, col,
                        kc);
  };
  if constexpr (D2) {
    // chunks alternate between two register sets, each refilled right after
    // it is stored to LDS with the chunk two ahead (the last chunk again
    // past the end: a read that hits L2, keeping every path's loads alike)
    if (ch0 < ch1) {
      const long long last = ch1 - 1;
      Pref<PV> qa, qb;
      fetch_all(qa, a.hi[0], a.lo[0], a.stride, n0, ch0 * kCK);
      fetch_all(qb, a.hi[0], a.lo[0], a.stride, n0, (ch0 + 1 < last ? ch0 + 1 : last) * kCK);
      long long ch = ch0;
      for (; ch + 1 < ch1; ch += 2) {  // chunk pairs
        __syncthreads();  // every wave is done with the previous chunk
        put(qa, n0, fA, hKs, lKs);
        __syncthreads();
        fetch_all(qa, a.hi[0], a.lo[0], a.stride, n0, (ch + 2 < last ? ch + 2 : last) * kCK);
        ksteps();
        __syncthreads();
        put(qb, n0, fA, hKs, lKs);
        __syncthreads();
        fetch_all(qb, a.hi[0], a.lo[0], a.stride, n0, (ch + 3 < last ? ch + 3 : last) * kCK);
        ksteps();
      }
      if (ch < ch1) {  // odd count: the last chunk
        __syncthreads();
        put(qa, n0, fA, hKs, lKs);
        __syncthreads();
        ksteps();
      }
    }
  } else {
    Pref<PV> p0, p1, p2;
    if (ch0 < ch1) {
      fetch(p0, a.hi[0], a.lo[0], a.stride, n0, ch0 * kCK);
      if (SETS == 3) fetch(p1, a.hi[1], a.lo[1], a.stride, n1, ch0 * kCK);
      if (SETS >= 2) fetch(p2, a.hi[2], a.lo[2], a.stride, n2, ch0 * kCK);
    }
    for (long long ch = ch0; ch < ch1; ++ch) {
      __syncthreads();  // every wave is done with the previous chunk
      put(p0, n0, fA, SETS == 1 ? hKs : nullptr, SETS == 1 ? lKs : nullptr);
      if (SETS == 3) put(p1, n1, fB, nullptr, nullptr);
      if (SETS >= 2) put(p2, n2, nullptr, hKs, lKs);
      __syncthreads();
      if (ch + 1 < ch1) {  // next chunk's HBM loads overlap this chunk's math
        const long long cb = (ch + 1) * kCK;
        fetch(p0, a.hi[0], a.lo[0], a.stride, n0, cb);
        if (SETS == 3) fetch(p1, a.hi[1], a.lo[1], a.stride, n1, cb);
        if (SETS >= 2) fetch(p2, a.hi[2], a.lo[2], a.stride, n2, cb);
      }
      ksteps();
    }
  }
  float* const out = a.partial + (long long)lw * a.item_total * 256;
  if (rep == 1) {
    store_at(st, out, opos, a.accumulate != 0, grp, col);
    return;
  }
  // replicas of a wave's list ran other K steps: the first adds theirs
  // (through LDS, FP64) and stores
  const int wpr = kW / rep, wl = wave % wpr;  // waves per replica, wave within it
  double* x = reinterpret_cast<double*>(smem);
  __syncthreads();  // every wave is done with the staged chunk
  if (rgrp > 0)
#pragma unroll
    for (int s = 0; s < NI; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        x[(((rgrp - 1) * wpr + wl) * NI + s) * 256 + r * 64 + lane] =
            st.acc2[s][r] + (double)st.acc[s][r];
  __syncthreads();
  if (rgrp == 0) {
    for (int g = 1; g < rep; ++g)
#pragma unroll
      for (int s = 0; s < NI; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) st.acc2[s][r] += x[(((g - 1) * wpr + wl) * NI + s) * 256 + r * 64 + lane];
    store_at(st, out, opos, a.accumulate != 0, grp, col);
  }
}

// Reduction of the per-workgroup FP32 partials in two coalesced steps, in
// a fixed order (deterministic): k_reduce_slots sums, per partial slot, the
// workgroups of each of kRedSplit ranges in FP64; k_reduce_tri adds, per
// triangle, the ranges of its slots (one per launch segment of its level,
// kPairSlots entries, -1 for none) and applies the scale and the level factor.
constexpr int kRedSplit = 16;

__global__ __launch_bounds__(256) void k_reduce_slots(const float* __restrict__ partial, int nwg,
                                                      int nslot, double* __restrict__ sums) {
  const int e = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (e >= nslot) return;
  const int w0 = (int)((long long)nwg * c / kRedSplit), w1 = (int)((long long)nwg * (c + 1) / kRedSplit);
  // four interleaved chains (four loads in flight), added in a fixed order
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int w = w0;
  for (; w + 4 <= w1; w += 4)
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] += (double)partial[(long long)(w + k) * nslot + e];
  for (; w < w1; ++w) s[0] += (double)partial[(long long)w * nslot + e];
  sums[(long long)c * nslot + e] = (s[0] + s[1]) + (s[2] + s[3]);
}

static_assert(kPairSlots == 4, "k_reduce_tri reads a triangle's slots as one int4");

__global__ __launch_bounds__(256) void k_reduce_tri(const double* __restrict__ sums, int nslot,
                                                    int ntri, const int* __restrict__ slot,
                                                    double scale, double* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntri) return;
  const int4 v = *reinterpret_cast<const int4*>(slot + (long long)kPairSlots * t);
  const int vs[kPairSlots] = {v.x, v.y, v.z, v.w};
  // bits 28-29: sublattice level L, whose cells are 2^-L of the mesh (the
  // first slot is always present)
  const double f = (double)(1 << ((v.x >> 28) & 3));
  double s = 0.0;
#pragma unroll
  for (int g = 0; g < kPairSlots; ++g) {
    if (vs[g] < 0) continue;
    const int e = vs[g] & 0x0fffffff;
#pragma unroll
    for (int c = 0; c < kRedSplit; ++c) s += sums[(long long)c * nslot + e];
  }
  out[t] = s * scale * f;
}

template <int NI, int PV, int SETS, bool D2>
int launch_pairs(const PairArgs& a, size_t lds, hipStream_t s) {
  auto fn = k_triple_pairs<NI, PV, SETS, D2>;
  // once per instance (not per launch: keeps launches capturable and cheap)
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  BSK_HIP_CHECK(attr);
  hipLaunchKernelGGL(fn, dim3(a.nwg * a.npass), dim3(kNT), lds, s, a);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}

// Replicated wave lists (few items per workgroup) with one field and one
// prefetch segment per thread run the two-chunk prefetch.
template <int NI, int PV, int SETS>
int launch_d2(const PairArgs& a, size_t lds, hipStream_t s) {
  if constexpr (SETS == 1 && PV == 1 && NI <= 4 && NI != 2)  // NI = 2: no registers to spare
    if (a.rep > 1) return launch_pairs<NI, PV, SETS, true>(a, lds, s);
  return launch_pairs<NI, PV, SETS, false>(a, lds, s);
}

template <int PV, int SETS>
int launch_ni(const PairArgs& a, int ni, size_t lds, hipStream_t s) {
  switch (ni) {
    case 1: return launch_d2<1, PV, SETS>(a, lds, s);
    case 2: return launch_d2<2, PV, SETS>(a, lds, s);
    case 3: return launch_d2<3, PV, SETS>(a, lds, s);
    case 4: return launch_d2<4, PV, SETS>(a, lds, s);
    case 6: return launch_pairs<6, PV, SETS, false>(a, lds, s);
    case 8: return launch_pairs<8, PV, SETS, false>(a, lds, s);
    default:
      bsk_set_error("mfma triple: %d items per wave not compiled", ni);
      return BSK_EUNSUPPORTED;
  }
}

template <int SETS>
int launch_pv(const PairArgs& a, int ni, int pv, size_t lds, hipStream_t s) {
  switch (pv) {
    case 1: return launch_ni<1, SETS>(a, ni, lds, s);
    case 2: return launch_ni<2, SETS>(a, ni, lds, s);
    case 4: return launch_ni<4, SETS>(a, ni, lds, s);
    default:
      bsk_set_error("mfma triple: %d prefetch segments per thread not compiled", pv);
      return BSK_EUNSUPPORTED;
  }
}

}  // namespace

int bsk_pair_items_round(int ni) {
  if (ni <= 4) return ni < 1 ? 1 : ni;
  if (ni <= 6) return 6;
  if (ni <= 8) return 8;
  return 0;
}

int bsk_pair_item_cap(int sets, int nshell) {
  // with 4 prefetch segments per set, 8 (<AAB>) or 6 and 8 (<ABC>) items per
  // wave exceed the register file at two waves per SIMD
  if (bsk_pair_prefetch(nshell) == 4 && sets >= 2) return sets == 2 ? 6 : 4;
  return kPairMaxItems;
}

int bsk_pair_prefetch(int nshell) {
  const int pv = (nshell * kSEGS + kNT - 1) / kNT;
  return pv <= 1 ? 1 : (pv <= 2 ? 2 : (pv <= 4 ? 4 : 0));
}

size_t bsk_pair_lds_bytes(int sets, const int nsh[3], int nkb) {
  const size_t frows = (size_t)(nsh[0] + 1) + (sets == 3 ? (size_t)(nsh[1] + 1) : 0);
  return frows * kLDF * 4 + (size_t)2 * 16 * nkb * kLDH * 2;
}

int bsk_triple_pairs_launch(const PairArgs& a, int sets, int ni, hipStream_t s) {
  if (a.ncell % kCK) {
    bsk_set_error("mfma triple: cell count %lld not a multiple of %d", a.ncell, kCK);
    return BSK_EINVAL;
  }
  if (a.nsh[0] < 1 || a.nsh[1] < 1 || a.nsh[2] < 1) {
    bsk_set_error("mfma triple: every role set needs a shell");
    return BSK_EINVAL;
  }
  if (a.nwg <= 0 || a.stride % 8) {
    bsk_set_error("mfma triple: nwg > 0 and shell strides in multiples of 8 elements");
    return BSK_EINVAL;
  }
  if (a.npass < 1 || (a.npass > 1 && (a.nwg % 8 || a.rep != 1))) {
    bsk_set_error("mfma triple: folded passes need nwg a multiple of 8 and no replicas");
    return BSK_EINVAL;
  }
  int nmax = a.nsh[0];
  for (int r = 1; r < 3; ++r) nmax = a.nsh[r] > nmax ? a.nsh[r] : nmax;
  const int pv = bsk_pair_prefetch(nmax);
  if (a.rep != 1 && a.rep != 2 && a.rep != 4) {
    bsk_set_error("mfma triple: %d replicas (1, 2 or 4)", a.rep);
    return BSK_EINVAL;
  }
  // LDS: the staging, and the replicas' tiles (FP64) at the end
  const size_t lds = std::max(bsk_pair_lds_bytes(sets, a.nsh, a.nkb),
                              (size_t)(a.rep - 1) * (kW / a.rep) * ni * 256 * 8);
  if (!pv || lds > 160 * 1024) {
    bsk_set_error("mfma triple: %d shells exceed the kernel's staging (LDS %zu bytes)", nmax,
                  lds);
    return BSK_EUNSUPPORTED;
  }
  if (sets == 1) return launch_pv<1>(a, ni, pv, lds, s);
  if (sets == 2) return launch_pv<2>(a, ni, pv, lds, s);
  if (sets == 3) return launch_pv<3>(a, ni, pv, lds, s);
  bsk_set_error("mfma triple: sets must be 1, 2 or 3");
  return BSK_EINVAL;
}

size_t bsk_reduce_mfma_bytes(int items_per_wg) {
  return sizeof(double) * kRedSplit * 256 * (size_t)items_per_wg;
}

int bsk_reduce_mfma_launch(const float* partial, int nwg, int items_per_wg, int ntri,
                           const int* slot, double scale, double* out, double* sums,
                           hipStream_t s) {
  if (ntri <= 0) return BSK_OK;
  const int nslot = items_per_wg * 256;
  hipLaunchKernelGGL(k_reduce_slots, dim3((nslot + 255) / 256, kRedSplit), dim3(256), 0, s,
                     partial, nwg, nslot, sums);
  BSK_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_tri, dim3((ntri + 255) / 256), dim3(256), 0, s, sums, nslot, ntri,
                     slot, scale, out);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
