// Batched 1-D complex FFTs on pencils held in LDS (Stockham autosort).
//
// A workgroup of W threads transforms P pencils of length L that sit in LDS
// at stride LD complex elements.  Each stage of radix R performs P*L/R
// butterflies; a thread loads all of its butterflies' inputs into registers,
// the workgroup synchronises, and the thread writes the outputs back in
// place at the Stockham (self-sorting) positions, so no second LDS buffer is
// needed and the result is in natural order.  Twiddles come from a table
// tw[t] = exp(-2 pi i t / TWN) in LDS; a length-L sub-transform indexes it
// with stride TWN/L (TWN is a multiple of L).
#pragma once
#include "common.h"

template <int... Rs>
struct Radices {};

// ---------------------------------------------------------------- butterflies
template <typename T, int R, bool INV>
struct Dft;

template <typename T, bool INV>
struct Dft<T, 2, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, bool INV>
struct Dft<T, 3, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    const T s = (T)0.86602540378443864676;  // sin(2 pi / 3)
    cplx<T> a = v[0], b = v[1], c = v[2];
    cplx<T> t1 = cadd(b, c);
    cplx<T> t2 = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
    cplx<T> u = cscale(csub(b, c), s);
    cplx<T> iu = mul_mi<T, INV>(u);  // -i*u forward, +i*u inverse
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, iu);
    v[2] = csub(t2, iu);
  }
};

template <typename T, bool INV>
struct Dft<T, 4, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    cplx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cplx<T> s13 = cadd(v[1], v[3]), d13 = mul_mi<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <typename T, bool INV>
struct Dft<T, 8, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    const T h = (T)0.70710678118654752440;
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, INV>::run(e);
    Dft<T, 4, INV>::run(o);
    // o[k] *= w8^k, w8 = exp(-+ 2 pi i / 8)
    cplx<T> o1 = o[1], o2 = o[2], o3 = o[3];
    if (!INV) {
      o[1] = {h * (o1.x + o1.y), h * (o1.y - o1.x)};
      o[2] = {o2.y, -o2.x};
      o[3] = {h * (o3.y - o3.x), -h * (o3.x + o3.y)};
    } else {
      o[1] = {h * (o1.x - o1.y), h * (o1.x + o1.y)};
      o[2] = {-o2.y, o2.x};
      o[3] = {-h * (o3.x + o3.y), h * (o3.x - o3.y)};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

template <typename T, bool INV>
struct Dft<T, 5, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    const T c1 = (T)0.30901699437494742410;   // cos(2 pi / 5)
    const T c2 = (T)-0.80901699437494742410;  // cos(4 pi / 5)
    const T s1 = (T)0.95105651629515357212;   // sin(2 pi / 5)
    const T s2 = (T)0.58778525229247312917;   // sin(4 pi / 5)
    const cplx<T> x0 = v[0];
    const cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const cplx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const cplx<T> a1 = {x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y};
    const cplx<T> a2 = {x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y};
    const cplx<T> b1 = mul_mi<T, INV>({s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
    const cplx<T> b2 = mul_mi<T, INV>({s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
    v[0] = cadd(x0, cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// 10 = 2 x 5: E = DFT5 of the even inputs, O = DFT5 of the odd ones, then
// X[k] = E[k] + w^k O[k] and X[k + 5] = E[k] - w^k O[k], w = exp(-+2 pi i / 10).
template <typename T, bool INV>
struct Dft<T, 10, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    const T c1 = (T)0.80901699437494742410;  // cos(2 pi / 10)
    const T s1 = (T)0.58778525229247312917;  // sin(2 pi / 10)
    const T c2 = (T)0.30901699437494742410;  // cos(4 pi / 10)
    const T s2 = (T)0.95105651629515357212;  // sin(4 pi / 10)
    cplx<T> e[5] = {v[0], v[2], v[4], v[6], v[8]};
    cplx<T> o[5] = {v[1], v[3], v[5], v[7], v[9]};
    Dft<T, 5, INV>::run(e);
    Dft<T, 5, INV>::run(o);
    const T sg = INV ? (T)1 : (T)-1;  // sign of the twiddles' imaginary parts
    const cplx<T> w[5] = {{(T)1, (T)0}, {c1, sg * s1}, {c2, sg * s2}, {-c2, sg * s2},
                          {-c1, sg * s1}};
    v[0] = cadd(e[0], o[0]);
    v[5] = csub(e[0], o[0]);
#pragma unroll
    for (int k = 1; k < 5; ++k) {
      const cplx<T> t = cmul(o[k], w[k]);
      v[k] = cadd(e[k], t);
      v[k + 5] = csub(e[k], t);
    }
  }
};

// 20 = 4 x 5: for s < 4 the DFT5 of v[4m + s] (m < 5), twiddled by
// w^(s k), w = exp(-+2 pi i / 20), then X[k + 5q] = DFT4 over s (k < 5).
template <typename T, bool INV>
struct Dft<T, 20, INV> {
  static BSK_DEV void run(cplx<T>* v) {
    const T a = (T)0.95105651629515357212;  // cos(pi / 10) = sin(2 pi / 5)
    const T b = (T)0.30901699437494742410;  // sin(pi / 10) = cos(2 pi / 5)
    const T c = (T)0.80901699437494742410;  // cos(pi / 5)
    const T d = (T)0.58778525229247312917;  // sin(pi / 5)
    // cos and sin of 2 pi j / 20 for the exponents j = s k used (s, k < 5)
    const T wc[13] = {(T)1, a, c, d, b, 0, -b, 0, -c, -a, 0, 0, -c};
    const T ws[13] = {0, b, d, c, a, 0, a, 0, d, b, 0, 0, -d};
    const T sg = INV ? (T)1 : (T)-1;
    cplx<T> x[4][5];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int m = 0; m < 5; ++m) x[s][m] = v[4 * m + s];
      Dft<T, 5, INV>::run(x[s]);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      cplx<T> y[4] = {x[0][k], x[1][k], x[2][k], x[3][k]};
#pragma unroll
      for (int s = 1; s < 4; ++s)
        if (k > 0) y[s] = cmul(y[s], cplx<T>{wc[s * k], sg * ws[s * k]});
      Dft<T, 4, INV>::run(y);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[k + 5 * q] = y[q];
    }
  }
};

// Pencil stride in LDS: the next odd length >= L (complex elements), so
// consecutive pencils start two (FP32) or four (FP64) banks apart and the
// lanes of a wave, which walk consecutive pencils, hit distinct banks.
template <int L>
constexpr int pencil_ld() {
  return L | 1;
}

// ---------------------------------------------------------------- one stage
// Butterfly b of the stage: pencil p = b % P (fastest across lanes), column
// j = b / P.  Loads read j + r*NB, stores write (j/NS)*NS*R + k + r*NS of the
// pencil: one base address plus immediate offsets per butterfly, and lanes
// with the same j share their twiddles (broadcast reads).
template <typename T, int L, int R, int NS, bool INV, int P, int W, int LD, int TWN>
BSK_DEV void stockham_stage(cplx<T>* buf, const cplx<T>* tw) {
  constexpr int NB = L / R;
  constexpr int TOT = P * NB;
  constexpr int BPT = (TOT + W - 1) / W;
  const int tid = threadIdx.x;
  cplx<T> v[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * W;
    if ((TOT % W) == 0 || b < TOT) {
      const int p = b % P, j = b / P;
      const cplx<T>* src = buf + p * LD + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * NB];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * W;
    if ((TOT % W) == 0 || b < TOT) {
      const int p = b % P, j = b / P;
      const int k = j % NS;
      if (NS > 1) {
        const int ks = k * (TWN / (NS * R));
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> w = tw[ks * r];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      Dft<T, R, INV>::run(v[q]);
      cplx<T>* dst = buf + p * LD + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, int L, bool INV, int P, int W, int LD, int TWN, int NS, int R, int... Rest>
BSK_DEV void stockham_run(cplx<T>* buf, const cplx<T>* tw) {
  stockham_stage<T, L, R, NS, INV, P, W, LD, TWN>(buf, tw);
  if constexpr (sizeof...(Rest) > 0)
    stockham_run<T, L, INV, P, W, LD, TWN, NS * R, Rest...>(buf, tw);
}

template <typename T, int L, bool INV, int P, int W, int LD, int TWN, int... Rs>
BSK_DEV void fft_lds(cplx<T>* buf, const cplx<T>* tw, Radices<Rs...>) {
  stockham_run<T, L, INV, P, W, LD, TWN, 1, Rs...>(buf, tw);
}

// Entries of the twiddle table a radix sequence reads: the stage of radix R
// after NS points reads tw[k * (TWN / (NS * R)) * r] for k < NS, r < R, so a
// table of this many entries (often well short of TWN) serves the transform.
template <int TWN, int NS, int R, int... Rest>
constexpr int tw_span_run() {
  constexpr int m = NS > 1 ? (NS - 1) * (TWN / (NS * R)) * (R - 1) + 1 : 1;
  if constexpr (sizeof...(Rest) > 0) {
    constexpr int n = tw_span_run<TWN, NS * R, Rest...>();
    return m > n ? m : n;
  } else {
    return m;
  }
}
template <int TWN, int... Rs>
constexpr int tw_span(Radices<Rs...>) {
  return tw_span_run<TWN, 1, Rs...>();
}

// ------------------------------------------- runtime pencil count (np <= PMAX)
// Quotient by a divisor n known only at run time: multiply-high with
// m = ceil(2^32 / n), exact for dividends below 2^32 / n; n = 1 (whose m
// does not fit 32 bits) is m = 0, the quotient the dividend itself.
BSK_DEV unsigned div_magic(int n) {
  return n > 1 ? (unsigned)((0x100000000ull + (unsigned)n - 1) / (unsigned)n) : 0u;
}
BSK_DEV int div_by_magic(unsigned b, unsigned m) { return m ? (int)__umulhi(b, m) : (int)b; }

// The same stages over np pencils known only at run time: butterfly b maps to
// column j = b / np and pencil p = b - j*np, the quotient from
// mdiv = div_magic(np) (exact for b < 2^32 / np, far above PMAX * L).
template <typename T, int L, int R, int NS, bool INV, int PMAX, int W, int LD, int TWN>
BSK_DEV void stockham_stage_n(cplx<T>* buf, const cplx<T>* tw, int np, unsigned mdiv, int tid) {
  constexpr int NB = L / R;
  constexpr int BPT = (PMAX * NB + W - 1) / W;
  const int tot = np * NB;
  cplx<T> v[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * W;
    if (b < tot) {
      const int j = div_by_magic((unsigned)b, mdiv), p = b - j * np;
      const cplx<T>* src = buf + p * LD + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * NB];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * W;
    if (b < tot) {
      const int j = div_by_magic((unsigned)b, mdiv), p = b - j * np;
      const int k = j % NS;
      if (NS > 1) {
        const int ks = k * (TWN / (NS * R));
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> w = tw[ks * r];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      Dft<T, R, INV>::run(v[q]);
      cplx<T>* dst = buf + p * LD + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, int L, bool INV, int PMAX, int W, int LD, int TWN, int NS, int R,
          int... Rest>
BSK_DEV void stockham_run_n(cplx<T>* buf, const cplx<T>* tw, int np, unsigned mdiv, int tid) {
  stockham_stage_n<T, L, R, NS, INV, PMAX, W, LD, TWN>(buf, tw, np, mdiv, tid);
  if constexpr (sizeof...(Rest) > 0)
    stockham_run_n<T, L, INV, PMAX, W, LD, TWN, NS * R, Rest...>(buf, tw, np, mdiv, tid);
}

// tid: the thread's index in the workgroup (passed in so a caller inside a
// loop can keep the butterfly index math from being held across iterations)
template <typename T, int L, bool INV, int PMAX, int W, int LD, int TWN, int... Rs>
BSK_DEV void fft_lds_n(cplx<T>* buf, const cplx<T>* tw, int np, unsigned mdiv, int tid,
                       Radices<Rs...>) {
  stockham_run_n<T, L, INV, PMAX, W, LD, TWN, 1, Rs...>(buf, tw, np, mdiv, tid);
}

// ---------------------------------------------------------------- radix plans
// Two variants per length: 0 = radix-8 heavy, 1 = radix-4 heavy (100 and
// 200: two stages of radix 10 and 20, one LDS round trip fewer).
template <int L, int V>
struct Plan;
#define BSK_PLAN(L, V, ...) \
  template <>               \
  struct Plan<L, V> {       \
    using seq = Radices<__VA_ARGS__>; \
  };
BSK_PLAN(8, 0, 8) BSK_PLAN(8, 1, 4, 2)
BSK_PLAN(12, 0, 4, 3) BSK_PLAN(12, 1, 2, 2, 3)
BSK_PLAN(16, 0, 4, 4) BSK_PLAN(16, 1, 8, 2)
BSK_PLAN(20, 0, 4, 5) BSK_PLAN(20, 1, 2, 2, 5)
BSK_PLAN(24, 0, 8, 3) BSK_PLAN(24, 1, 4, 2, 3)
BSK_PLAN(32, 0, 8, 4) BSK_PLAN(32, 1, 4, 4, 2)
BSK_PLAN(36, 0, 4, 3, 3) BSK_PLAN(36, 1, 2, 2, 3, 3)
BSK_PLAN(40, 0, 8, 5) BSK_PLAN(40, 1, 4, 2, 5)
BSK_PLAN(48, 0, 8, 2, 3) BSK_PLAN(48, 1, 4, 4, 3)
BSK_PLAN(60, 0, 4, 3, 5) BSK_PLAN(60, 1, 2, 2, 3, 5)
BSK_PLAN(64, 0, 8, 8) BSK_PLAN(64, 1, 4, 4, 4)
BSK_PLAN(72, 0, 8, 3, 3) BSK_PLAN(72, 1, 4, 2, 3, 3)
BSK_PLAN(80, 0, 8, 2, 5) BSK_PLAN(80, 1, 4, 4, 5)
BSK_PLAN(96, 0, 8, 4, 3) BSK_PLAN(96, 1, 4, 4, 2, 3)
BSK_PLAN(100, 0, 4, 5, 5) BSK_PLAN(100, 1, 10, 10)
BSK_PLAN(108, 0, 4, 3, 3, 3) BSK_PLAN(108, 1, 2, 2, 3, 3, 3)
BSK_PLAN(120, 0, 8, 3, 5) BSK_PLAN(120, 1, 4, 2, 3, 5)
BSK_PLAN(128, 0, 8, 4, 4) BSK_PLAN(128, 1, 4, 4, 4, 2)
BSK_PLAN(144, 0, 8, 2, 3, 3) BSK_PLAN(144, 1, 4, 4, 3, 3)
BSK_PLAN(160, 0, 8, 4, 5) BSK_PLAN(160, 1, 4, 4, 2, 5)
BSK_PLAN(192, 0, 8, 8, 3) BSK_PLAN(192, 1, 4, 4, 4, 3)
BSK_PLAN(200, 0, 8, 5, 5) BSK_PLAN(200, 1, 20, 10)
BSK_PLAN(216, 0, 8, 3, 3, 3) BSK_PLAN(216, 1, 4, 2, 3, 3, 3)
BSK_PLAN(240, 0, 8, 2, 3, 5) BSK_PLAN(240, 1, 4, 4, 3, 5)
BSK_PLAN(256, 0, 8, 8, 4) BSK_PLAN(256, 1, 4, 4, 4, 4)
BSK_PLAN(288, 0, 8, 4, 3, 3) BSK_PLAN(288, 1, 4, 4, 2, 3, 3)
BSK_PLAN(320, 0, 8, 8, 5) BSK_PLAN(320, 1, 4, 4, 4, 5)
BSK_PLAN(384, 0, 8, 8, 2, 3) BSK_PLAN(384, 1, 8, 4, 4, 3)
BSK_PLAN(400, 0, 8, 2, 5, 5) BSK_PLAN(400, 1, 4, 4, 5, 5)
BSK_PLAN(432, 0, 8, 2, 3, 3, 3) BSK_PLAN(432, 1, 4, 4, 3, 3, 3)
BSK_PLAN(480, 0, 8, 4, 3, 5) BSK_PLAN(480, 1, 4, 4, 2, 3, 5)
BSK_PLAN(512, 0, 8, 8, 8) BSK_PLAN(512, 1, 4, 4, 4, 8)
BSK_PLAN(576, 0, 8, 8, 3, 3) BSK_PLAN(576, 1, 4, 4, 4, 3, 3)
BSK_PLAN(640, 0, 8, 8, 2, 5) BSK_PLAN(640, 1, 8, 4, 4, 5)
BSK_PLAN(768, 0, 8, 8, 4, 3) BSK_PLAN(768, 1, 4, 4, 4, 4, 3)
BSK_PLAN(1024, 0, 8, 8, 4, 4) BSK_PLAN(1024, 1, 4, 4, 4, 4, 4)
#undef BSK_PLAN
