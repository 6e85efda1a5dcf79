// C ABI of libbskhip: plans, forward transform, shell fields, triple sums.
// Every data buffer is caller-owned device memory; plans own only their small
// constant tables (twiddles, shell thresholds, window compensation).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <mutex>
#include <queue>
#include <vector>

#include "../../include/bskhip.h"
#include "launch.h"

static thread_local char g_err[512];

void bsk_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

struct bsk_plan {
  int N;  // input mesh
  int S;  // shell mesh: the inverse transforms and triple sums run on S^3 cells
  double box;
  int nbin;
  std::vector<int> n2lo, n2hi;
  int m;
  Geom g;
  int variant[6];
  int* d_lo = nullptr;
  int* d_hi = nullptr;
  // block entries of the inverse x and y passes (the blocks that hold modes
  // of each shell; [0]: 16-pencil blocks, [1]: 8-pencil blocks), shells in
  // order, prefix[b] = first entry of shell b
  struct BlockList {
    int4* d = nullptr;
    std::vector<int> prefix;
  };
  BlockList xblk[2], yblk[2];
  cplx<float>* tw_f = nullptr;
  cplx<double>* tw_d = nullptr;
  cplx<float>* tws_f = nullptr;  // exp(-2 pi i t / S) when S != N
  cplx<double>* tws_d = nullptr;
  // CIC window compensation tables, one per window mesh size used on this
  // plan (kept until the plan is destroyed, so a launch in flight never
  // reads a freed table); comp_f / comp_d point at the selected one
  struct Comp {
    int window;
    float* f;
    double* d;
  };
  std::vector<Comp> comps;
  double kpower = 1.0;  // exponent p of the BSK_KIND_KPOW shells
  float* comp_f = nullptr;
  double* comp_d = nullptr;
};

namespace {

const double kTieRtol = 1e-9;

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// The shell fields are band limited: every retained mode has integer
// components |n_a| <= m.  A product of three of them only holds frequencies
// with components in [-3m, 3m], so on any periodic mesh of S > 3m points per
// axis the sum over cells of d_i d_j d_k picks up exactly the zero frequency:
//   sum_{x in S^3} d_i d_j d_k = S^3 sum_{q1+q2+q3=0} F_i(q1) F_j(q2) F_k(q3),
// the same sum the N^3 mesh gives up to the factor (N/S)^3.  The shells are
// therefore evaluated on the smallest supported S > 3m when that is below N.
int auto_shell_mesh(int N, int m) {
  if (2 * m + 1 >= N) return N;
  const int s = bsk_fft_size_above(3 * m);
  return (s > 0 && s < N) ? s : N;
}

// Sublattice level layout of a shell (the parts of ShellLevels in
// fft_passes.h: P0 = C0 minus E, P1 = E, P2 = O, P3 = C1 minus O, each padded
// with zeros to a multiple of 128 cells, the MFMA kernel's chunk): level L's
// cells are the contiguous range [off[L], off[L] + ncell[L]) of the shell
// row.  Level 0: the whole mesh P0..P3, 1: C0 = P0 P1 (x+y+z even, 1/2 of
// the cells), 2: E and O = P1 P2 (x, y, z of equal parity, 1/4), 3: E = P1
// (all even, 1/8).
long long round128(long long v) { return (v + 127) / 128 * 128; }
struct LevelLayout {
  long long off[BSK_LEVELS], ncell[BSK_LEVELS], stride;
};
LevelLayout level_layout(int S) {
  const long long nh = S / 2, ne = round128(nh * nh * nh), nc = round128(3 * nh * nh * nh);
  return LevelLayout{{0, 0, nc, nc}, {2 * nc + 2 * ne, nc + ne, 2 * ne, ne}, 2 * nc + 2 * ne};
}

// Whether the sum over level L's sublattice is exact (|H|/S^3 of the full sum)
// for a product whose frequencies q satisfy |q_a| <= rinf (every component)
// and |q| <= r2: no nonzero vector of the dual lattice is such a q.  The dual
// vectors are (S/2)(a, b, c) with a, b, c all even (level 0), all of one
// parity (1), of even sum (2) or any (3).
bool level_exact(int S, int L, long long rinf, double r2) {
  const long long h = S / 2;
  const long long A = rinf / h;  // |a| * h <= rinf
  for (long long a = -A; a <= A; ++a)
    for (long long b = -A; b <= A; ++b)
      for (long long c = -A; c <= A; ++c) {
        if (!a && !b && !c) continue;
        const bool ok = L == 0   ? !(a & 1) && !(b & 1) && !(c & 1)
                        : L == 1 ? ((a ^ b) & 1) == 0 && ((b ^ c) & 1) == 0
                        : L == 2 ? ((a + b + c) & 1) == 0
                                 : true;
        if (!ok) continue;
        const double n2 = (double)(h * h) * (double)(a * a + b * b + c * c);
        if (n2 <= r2 * r2 * (1.0 + 1e-12)) return false;  // a reachable alias
      }
  return true;
}

template <typename T>
int make_twiddles(int n, cplx<T>** out) {
  std::vector<cplx<T>> h(n);
  for (int t = 0; t < n; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)n;
    h[t] = {(T)cos(a), (T)sin(a)};
  }
  BSK_HIP_CHECK(hipMalloc(out, sizeof(cplx<T>) * n));
  BSK_HIP_CHECK(hipMemcpy(*out, h.data(), sizeof(cplx<T>) * n, hipMemcpyHostToDevice));
  return BSK_OK;
}

int set_shell_mesh(bsk_plan* p, int S) {
  (void)hipFree(p->tws_f);
  (void)hipFree(p->tws_d);
  p->tws_f = nullptr;
  p->tws_d = nullptr;
  p->S = S;
  if (S == p->N) return BSK_OK;
  int rc = make_twiddles<float>(S, &p->tws_f);
  if (!rc) rc = make_twiddles<double>(S, &p->tws_d);
  return rc;
}

template <typename T>
int forward_t(bsk_plan* p, const T* field, cplx<T>* dk, void* work, hipStream_t s,
              const cplx<T>* tw, const T* comp) {
  const int N = p->N;
  const Geom g = p->g;
  cplx<T>* e1 = reinterpret_cast<cplx<T>*>(work);
  cplx<T>* e2 = e1 + (size_t)N * N * g.MZP;
  int rc = bsk_fwd_z<T>(N, p->variant[0], field, e1, tw, (long long)N * N, g.MZ, g.MZP, s);
  if (rc) return rc;
  rc = bsk_fwd_y<T>(N, p->variant[1], e1, e2, tw, g, s);
  if (rc) return rc;
  const T scale = (T)(1.0 / ((double)N * N * N));
  return bsk_fwd_x<T>(N, p->variant[2], e2, dk, tw, g, scale, comp, s);
}

// Host twin of shell_mz: kz from this value on hold no mode of a shell with
// |n|^2 <= hi.
int host_shell_mz(int MZ, int hi) {
  if (hi < 0) return 0;
  long long r = (long long)std::sqrt((double)hi);
  while (r * r > hi) --r;
  while ((r + 1) * (r + 1) <= hi) ++r;
  return (int)std::min<long long>(r + 1, MZ);
}

int upload_blocks(bsk_plan::BlockList& bl, const std::vector<int4>& v) {
  if (v.empty()) return BSK_OK;
  hipError_t e = hipMalloc(&bl.d, sizeof(int4) * v.size());
  if (e == hipSuccess) e = hipMemcpy(bl.d, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    bsk_set_error("plan block lists: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? BSK_ENOMEM : BSK_EHIP;
  }
  return BSK_OK;
}

// The workgroups of the inverse x and y passes that hold modes: x blocks
// (ry, kz0) with ny^2 + kz0^2 <= n2hi, y blocks kz0 < mz_b, each entry with
// the shell's thresholds.  Independent of the shell mesh.
int build_block_lists(bsk_plan* p) {
  const Geom& g = p->g;
  for (int k = 0; k < 2; ++k) {
    const int P = k ? 8 : 16;
    std::vector<int4> xl, yl;
    p->xblk[k].prefix.assign(p->nbin + 1, 0);
    p->yblk[k].prefix.assign(p->nbin + 1, 0);
    for (int b = 0; b < p->nbin; ++b) {
      p->xblk[k].prefix[b] = (int)xl.size();
      p->yblk[k].prefix[b] = (int)yl.size();
      const int lo = p->n2lo[b], hi = p->n2hi[b];
      for (int kz0 = 0; kz0 < host_shell_mz(g.MZ, hi); kz0 += P)
        yl.push_back(make_int4(b, kz0, hi, 0));
      for (int ry = 0; ry < g.MY; ++ry) {
        const long long ny = g.nlo + ry;
        for (int kz0 = 0; kz0 < g.MZ; kz0 += P)
          if (ny * ny + (long long)kz0 * kz0 <= hi)
            xl.push_back(make_int4(b, (ry << 16) | kz0, lo, hi));
      }
    }
    p->xblk[k].prefix[p->nbin] = (int)xl.size();
    p->yblk[k].prefix[p->nbin] = (int)yl.size();
    if (int rc = upload_blocks(p->xblk[k], xl)) return rc;
    if (int rc = upload_blocks(p->yblk[k], yl)) return rc;
  }
  return BSK_OK;
}

template <typename T>
int shells_t(bsk_plan* p, int kind, const cplx<T>* dk, int shell0, int nsh, int x0, int xs,
             int stages, int fmt, void* out, void* out_lo, long long stride, long long cell0,
             void* work, hipStream_t s, const cplx<T>* tw,
             const unsigned char* parts = nullptr) {
  const int N = p->S;
  const Geom g = p->g;
  cplx<T>* d1 = reinterpret_cast<cplx<T>*>(work);
  cplx<T>* d2 = d1 + (size_t)nsh * g.MY * N * g.MZP;
  const T kf = (T)(2.0 * M_PI / p->box);
  // the pass kernels address one shell's inverse-x output, and one shell
  // row of the level layout, with 32-bit byte offsets from a 64-bit base
  const size_t esz0 = (fmt == BSK_FMT_F64) ? 8 : (fmt == BSK_FMT_F32 ? 4 : 2);
  if ((unsigned long long)N * g.MY * g.MZP * sizeof(cplx<T>) >= (1ull << 32) ||
      ((stages & BSK_STAGE_LEVELS) && (unsigned long long)stride * esz0 >= (1ull << 32))) {
    bsk_set_error("bsk_shells: a shell of this plan spans 4 GiB or more in one pass (shell mesh "
                  "%d, retained cube %d x %d x %d); use FP32/BF16X2 or a smaller shell reach",
                  N, g.MX, g.MY, g.MZ);
    return BSK_EUNSUPPORTED;
  }
  int rc = BSK_OK;
  // the block entries of shells [shell0, shell0 + nsh) for this pencil count
  const int kb = bsk_pxy(sizeof(T) == 8, N) == 8 ? 1 : 0;
  const bsk_plan::BlockList &XB = p->xblk[kb], &YB = p->yblk[kb];
  const int4* xb = XB.d + XB.prefix[shell0];
  const int nxb = XB.prefix[shell0 + nsh] - XB.prefix[shell0];
  const int4* yb = YB.d + YB.prefix[shell0];
  const int nyb = YB.prefix[shell0 + nsh] - YB.prefix[shell0];
  if (stages & BSK_STAGE_X) {
    rc = bsk_inv_x<T>(N, p->variant[3], dk, d1, tw, g, shell0, xb, nxb, kind, kf, (T)p->kpower,
                      s);
    if (rc) return rc;
  }
  if (stages & BSK_STAGE_Y)  // y only: D2 stays in the workspace
    return bsk_inv_y<T>(N, p->variant[4], d1, d2, tw, g, x0, xs, yb, nyb, shell0, s);
  if (!(stages & BSK_STAGE_YZ)) return BSK_OK;
  // shells [shell0, shell0+nsh) of the output arrays, cells of the slab
  const size_t esz = (fmt == BSK_FMT_F64) ? 8 : (fmt == BSK_FMT_F32 ? 4 : 2);
  char* o = reinterpret_cast<char*>(out) + (size_t)shell0 * stride * esz;
  char* ol = out_lo ? reinterpret_cast<char*>(out_lo) + (size_t)shell0 * stride * esz : nullptr;
  rc = bsk_inv_y<T>(N, p->variant[4], d1, d2, tw, g, x0, xs, yb, nyb, shell0, s);
  if (rc) return rc;
  return bsk_inv_z<T>(N, p->variant[5], d2, o, ol, tw, g.MZ, g.MZP, (long long)xs * N, cell0,
                      stride, fmt, nsh, p->d_hi, shell0, (stages & BSK_STAGE_LEVELS) ? 1 : 0,
                      parts, s);
}

const cplx<float>* shell_tw(const bsk_plan* p, float) { return p->tws_f ? p->tws_f : p->tw_f; }
const cplx<double>* shell_tw(const bsk_plan* p, double) { return p->tws_d ? p->tws_d : p->tw_d; }

int check_plan(const bsk_plan* p) {
  if (!p) {
    bsk_set_error("null plan");
    return BSK_EINVAL;
  }
  return BSK_OK;
}

}  // namespace

bool bsk_fft_size_supported(int N) {
  switch (N) {
#define X(n) case n:
    BSK_FFT_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

int bsk_fft_size_above(int n) {
  static const int sizes[] = {
#define X(v) v,
      BSK_FFT_SIZES(X)
#undef X
  };
  for (int v : sizes)
    if (v > n) return v;
  return 0;
}

extern "C" {

const char* bsk_last_error(void) { return g_err; }

int bsk_version(void) { return BSK_ABI_VERSION; }

int bsk_shell_thresholds(int nmesh, double box_size, int nbin, const double* k_low,
                         const double* k_high, int* n2lo, int* n2hi) {
  if (nmesh <= 0 || box_size <= 0 || nbin < 0 || (nbin && (!k_low || !k_high || !n2lo || !n2hi))) {
    bsk_set_error("bsk_shell_thresholds: invalid argument");
    return BSK_EINVAL;
  }
  const double kf = 2.0 * M_PI / box_size;
  const double n2max = 3.0 * (nmesh / 2.0) * (nmesh / 2.0);
  for (int b = 0; b < nbin; ++b) {
    const double a = k_low[b] > 0 ? (k_low[b] / kf) * (k_low[b] / kf) : 0.0;
    const double c = k_high[b] >= 0 ? (k_high[b] / kf) * (k_high[b] / kf) : -1.0;
    double lo = a > 0 ? ceil(a * (1.0 - kTieRtol)) : 0.0;
    double hi = c >= 0 ? floor(c * (1.0 + kTieRtol)) : -1.0;
    if (lo > n2max + 1) lo = n2max + 1;
    if (hi > n2max) hi = n2max;
    n2lo[b] = (int)lo;
    n2hi[b] = (int)hi;
  }
  return BSK_OK;
}

int bsk_plan_create(bsk_plan** out, int nmesh, double box_size, int nbin, const double* k_low,
                    const double* k_high) {
  if (!out || nbin <= 0 || !k_low || !k_high || box_size <= 0) {
    bsk_set_error("bsk_plan_create: invalid argument");
    return BSK_EINVAL;
  }
  *out = nullptr;
  if (!bsk_fft_size_supported(nmesh)) {
    bsk_set_error("bsk_plan_create: Nmesh=%d is not a supported FFT size (even, 5-smooth, "
                  "16..1024, half length with a radix plan)", nmesh);
    return BSK_EUNSUPPORTED;
  }
  bsk_plan* p = new bsk_plan();
  p->N = nmesh;
  p->box = box_size;
  p->nbin = nbin;
  p->n2lo.resize(nbin);
  p->n2hi.resize(nbin);
  int rc = bsk_shell_thresholds(nmesh, box_size, nbin, k_low, k_high, p->n2lo.data(), p->n2hi.data());
  if (rc) {
    delete p;
    return rc;
  }
  int n2max = 0;
  for (int b = 0; b < nbin; ++b) n2max = std::max(n2max, p->n2hi[b]);
  int m = (int)floor(sqrt((double)n2max));
  while ((long long)(m + 1) * (m + 1) <= n2max) ++m;
  while ((long long)m * m > n2max) --m;
  p->m = m;
  const int N = nmesh;
  Geom g;
  g.N = N;
  if (2 * m + 1 >= N) {
    g.MX = g.MY = N;
    g.nlo = -N / 2;
  } else {
    g.MX = g.MY = 2 * m + 1;
    g.nlo = -m;
  }
  g.MZ = std::min(m + 1, N / 2 + 1);
  g.MZP = (g.MZ + 15) & ~15;
  p->g = g;
  p->S = N;
  for (int i = 0; i < 6; ++i) p->variant[i] = 0;

  hipError_t e = hipSuccess;
  e = hipMalloc(&p->d_lo, sizeof(int) * nbin);
  if (e == hipSuccess) e = hipMalloc(&p->d_hi, sizeof(int) * nbin);
  if (e == hipSuccess) e = hipMemcpy(p->d_lo, p->n2lo.data(), sizeof(int) * nbin, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_hi, p->n2hi.data(), sizeof(int) * nbin, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    bsk_set_error("bsk_plan_create: %s", hipGetErrorString(e));
    bsk_plan_destroy(p);
    return e == hipErrorOutOfMemory ? BSK_ENOMEM : BSK_EHIP;
  }
  rc = make_twiddles<float>(N, &p->tw_f);
  if (!rc) rc = make_twiddles<double>(N, &p->tw_d);
  if (!rc) rc = build_block_lists(p);
  if (!rc) rc = set_shell_mesh(p, auto_shell_mesh(N, m));
  if (rc) {
    bsk_plan_destroy(p);
    return rc;
  }
  *out = p;
  return BSK_OK;
}

int bsk_plan_destroy(bsk_plan* p) {
  if (!p) return BSK_OK;
  (void)hipFree(p->d_lo);
  (void)hipFree(p->d_hi);
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(p->xblk[k].d);
    (void)hipFree(p->yblk[k].d);
  }
  (void)hipFree(p->tw_f);
  (void)hipFree(p->tw_d);
  (void)hipFree(p->tws_f);
  (void)hipFree(p->tws_d);
  for (auto& c : p->comps) {
    (void)hipFree(c.f);
    (void)hipFree(c.d);
  }
  delete p;
  return BSK_OK;
}

int bsk_plan_geometry(const bsk_plan* p, int* m, int* mx, int* my, int* mz, int* nlo) {
  if (int rc = check_plan(p)) return rc;
  if (m) *m = p->m;
  if (mx) *mx = p->g.MX;
  if (my) *my = p->g.MY;
  if (mz) *mz = p->g.MZ;
  if (nlo) *nlo = p->g.nlo;
  return BSK_OK;
}

int bsk_plan_shell_mesh(const bsk_plan* p, int* shell_mesh) {
  if (int rc = check_plan(p)) return rc;
  if (!shell_mesh) {
    bsk_set_error("bsk_plan_shell_mesh: null output");
    return BSK_EINVAL;
  }
  *shell_mesh = p->S;
  return BSK_OK;
}

int bsk_plan_levels(const bsk_plan* p, long long* offset, long long* ncell, long long* stride) {
  if (int rc = check_plan(p)) return rc;
  if (!offset || !ncell || !stride) {
    bsk_set_error("bsk_plan_levels: null output");
    return BSK_EINVAL;
  }
  if (p->S % 4) {
    bsk_set_error("bsk_plan_levels: shell mesh %d is not a multiple of 4", p->S);
    return BSK_EUNSUPPORTED;
  }
  const LevelLayout l = level_layout(p->S);
  for (int L = 0; L < BSK_LEVELS; ++L) {
    offset[L] = l.off[L];
    ncell[L] = l.ncell[L];
  }
  *stride = l.stride;
  return BSK_OK;
}

int bsk_plan_triangle_levels(const bsk_plan* p, int ntri, const int* tri, int* level) {
  if (int rc = check_plan(p)) return rc;
  if (ntri < 0 || (ntri && (!tri || !level))) {
    bsk_set_error("bsk_plan_triangle_levels: invalid argument");
    return BSK_EINVAL;
  }
  // largest |n| component and |n| of each shell's modes
  std::vector<long long> mb(p->nbin);
  std::vector<double> rb(p->nbin);
  for (int b = 0; b < p->nbin; ++b) {
    long long r = 0;
    while ((r + 1) * (r + 1) <= (long long)p->n2hi[b]) ++r;
    mb[b] = r;
    rb[b] = p->n2hi[b] > 0 ? std::sqrt((double)p->n2hi[b]) : 0.0;
  }
  for (int t = 0; t < ntri; ++t) {
    long long rinf = 0;
    double r2 = 0.0;
    for (int e = 0; e < 3; ++e) {
      const int b = tri[3 * t + e];
      if (b < 0 || b >= p->nbin) {
        bsk_set_error("bsk_plan_triangle_levels: triangle %d has a bin index out of range", t);
        return BSK_EINVAL;
      }
      rinf += mb[b];
      r2 += rb[b];
    }
    int best = 0;
    if (p->S % 4 == 0)
      for (int L = BSK_LEVELS - 1; L > 0; --L)
        if (level_exact(p->S, L, rinf, r2)) {
          best = L;
          break;
        }
    level[t] = best;
  }
  return BSK_OK;
}

int bsk_plan_set_shell_mesh(bsk_plan* p, int shell_mesh) {
  if (int rc = check_plan(p)) return rc;
  const int S = shell_mesh == 0 ? auto_shell_mesh(p->N, p->m) : shell_mesh;
  if (S != p->N && (!bsk_fft_size_supported(S) || S <= 3 * p->m || S > p->N)) {
    bsk_set_error("bsk_plan_set_shell_mesh: %d is neither Nmesh=%d nor a supported size in "
                  "(3m, Nmesh) with m=%d", shell_mesh, p->N, p->m);
    return BSK_EINVAL;
  }
  return set_shell_mesh(p, S);
}

int bsk_plan_thresholds(const bsk_plan* p, int* n2lo, int* n2hi) {
  if (int rc = check_plan(p)) return rc;
  for (int b = 0; b < p->nbin; ++b) {
    if (n2lo) n2lo[b] = p->n2lo[b];
    if (n2hi) n2hi[b] = p->n2hi[b];
  }
  return BSK_OK;
}

int bsk_plan_set_variant(bsk_plan* p, int pass, int variant) {
  if (int rc = check_plan(p)) return rc;
  if (pass < 0 || pass >= 6 || variant < 0 || variant > 1) {
    bsk_set_error("bsk_plan_set_variant: pass in 0..5, variant in 0..1");
    return BSK_EINVAL;
  }
  p->variant[pass] = variant;
  return BSK_OK;
}

int bsk_plan_set_compensation(bsk_plan* p, int nmesh_window) {
  if (int rc = check_plan(p)) return rc;
  p->comp_f = nullptr;
  p->comp_d = nullptr;
  if (nmesh_window <= 0) return BSK_OK;
  for (const auto& c : p->comps)
    if (c.window == nmesh_window) {
      p->comp_f = c.f;
      p->comp_d = c.d;
      return BSK_OK;
    }
  const int N = p->N;
  std::vector<float> cf(N + 1);
  std::vector<double> cd(N + 1);
  for (int n = -N / 2; n <= N / 2; ++n) {
    const double sn = sin(M_PI * (double)n / (double)nmesh_window);
    cd[n + N / 2] = 1.0 / sqrt(1.0 - 2.0 / 3.0 * sn * sn);
    cf[n + N / 2] = (float)cd[n + N / 2];
  }
  bsk_plan::Comp c{nmesh_window, nullptr, nullptr};
  BSK_HIP_CHECK(hipMalloc(&c.f, sizeof(float) * (N + 1)));
  if (hipMalloc(&c.d, sizeof(double) * (N + 1)) != hipSuccess) {
    (void)hipFree(c.f);
    bsk_set_error("bsk_plan_set_compensation: out of device memory");
    return BSK_ENOMEM;
  }
  p->comps.push_back(c);
  BSK_HIP_CHECK(hipMemcpy(c.f, cf.data(), sizeof(float) * (N + 1), hipMemcpyHostToDevice));
  BSK_HIP_CHECK(hipMemcpy(c.d, cd.data(), sizeof(double) * (N + 1), hipMemcpyHostToDevice));
  p->comp_f = c.f;
  p->comp_d = c.d;
  return BSK_OK;
}

int bsk_plan_set_kpower(bsk_plan* p, double power) {
  if (int rc = check_plan(p)) return rc;
  if (!std::isfinite(power)) {
    bsk_set_error("bsk_plan_set_kpower: the power must be finite");
    return BSK_EINVAL;
  }
  p->kpower = power;
  return BSK_OK;
}

int bsk_forward_workspace(const bsk_plan* p, int dtype, size_t* bytes) {
  if (int rc = check_plan(p)) return rc;
  const size_t c = dtype == BSK_DT_F64 ? 16 : 8;
  const size_t N = p->N;
  *bytes = c * (N * N * p->g.MZP + N * p->g.MY * p->g.MZP);
  return BSK_OK;
}

int bsk_forward(bsk_plan* p, int dtype, const void* field, void* dk, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!field || !dk || !work) {
    bsk_set_error("bsk_forward: null buffer");
    return BSK_EINVAL;
  }
  if (dtype == BSK_DT_F32)
    return forward_t<float>(p, (const float*)field, (cplx<float>*)dk, work, as_stream(stream),
                            p->tw_f, p->comp_f);
  if (dtype == BSK_DT_F64)
    return forward_t<double>(p, (const double*)field, (cplx<double>*)dk, work, as_stream(stream),
                             p->tw_d, p->comp_d);
  bsk_set_error("bsk_forward: dtype must be BSK_DT_F32 or BSK_DT_F64");
  return BSK_EINVAL;
}

int bsk_shells_workspace(const bsk_plan* p, int dtype, int nshell, int xslab, size_t* bytes) {
  if (int rc = check_plan(p)) return rc;
  const size_t c = dtype == BSK_DT_F64 ? 16 : 8;
  const size_t S = p->S;
  *bytes = c * (size_t)nshell * p->g.MZP * ((size_t)p->g.MY * S + (size_t)xslab * S);
  return BSK_OK;
}

int bsk_shells(bsk_plan* p, int dtype, int kind, const void* dk, int shell0, int nshell, int x0,
               int xslab, int stages, int fmt, void* out, void* out_lo, long long shell_stride,
               long long out_cell0, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  const long long slab_cells = (long long)xslab * p->S * p->S;
  const bool yz = (stages & BSK_STAGE_YZ) != 0;
  if ((stages & BSK_STAGE_LEVELS) &&
      (!yz || p->S % 4 || x0 != 0 || xslab != p->S || out_cell0 != 0 ||
       shell_stride < level_layout(p->S).stride)) {
    bsk_set_error("bsk_shells: the level layout needs the y and z stages over the whole "
                  "mesh (x0 = 0, xslab = S, out_cell0 = 0), S a multiple of 4 and a shell "
                  "stride of at least the bsk_plan_levels stride");
    return BSK_EINVAL;
  }
  if (nshell <= 0 || shell0 < 0 || shell0 + nshell > p->nbin || x0 < 0 || xslab <= 0 ||
      x0 + xslab > p->S || !work || stages < 1 || stages > 15 ||
      ((stages & BSK_STAGE_Y) && (stages & BSK_STAGE_YZ)) ||
      ((stages & BSK_STAGE_X) && kind == BSK_KIND_DELTA && !dk) ||
      (kind < BSK_KIND_DELTA || kind > BSK_KIND_KPOW) ||
      (yz && (!out || out_cell0 < 0 || shell_stride < out_cell0 + slab_cells))) {
    bsk_set_error("bsk_shells: invalid argument");
    return BSK_EINVAL;
  }
  if (nshell > 65535) {
    bsk_set_error("bsk_shells: at most 65535 shells per call");
    return BSK_EINVAL;
  }
  if (dtype == BSK_DT_F32) {
    if (fmt != BSK_FMT_F32 && fmt != BSK_FMT_BF16X2) {
      bsk_set_error("bsk_shells: FP32 shells are stored as F32 or BF16X2");
      return BSK_EINVAL;
    }
    if (yz && fmt == BSK_FMT_BF16X2 && !out_lo) {
      bsk_set_error("bsk_shells: BF16X2 needs the lo plane");
      return BSK_EINVAL;
    }
    return shells_t<float>(p, kind, (const cplx<float>*)dk, shell0, nshell, x0, xslab, stages, fmt,
                           out, out_lo, shell_stride, out_cell0, work, as_stream(stream),
                           shell_tw(p, 0.f));
  }
  if (dtype == BSK_DT_F64) {
    if (fmt != BSK_FMT_F64) {
      bsk_set_error("bsk_shells: FP64 shells are stored as F64");
      return BSK_EINVAL;
    }
    return shells_t<double>(p, kind, (const cplx<double>*)dk, shell0, nshell, x0, xslab, stages,
                            fmt, out, nullptr, shell_stride, out_cell0, work, as_stream(stream),
                            shell_tw(p, 0.0));
  }
  bsk_set_error("bsk_shells: dtype must be BSK_DT_F32 or BSK_DT_F64");
  return BSK_EINVAL;
}

int bsk_shells_levels(bsk_plan* p, int dtype, int kind, const void* dk, int shell0, int nshell,
                      int fmt, void* out, void* out_lo, long long shell_stride,
                      const unsigned char* parts, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (dtype != BSK_DT_F32 || (fmt != BSK_FMT_F32 && fmt != BSK_FMT_BF16X2)) {
    if (dtype != BSK_DT_F64 || fmt != BSK_FMT_F64) {
      bsk_set_error("bsk_shells_levels: FP32 shells as F32 or BF16X2, FP64 shells as F64");
      return BSK_EINVAL;
    }
  }
  if (nshell <= 0 || shell0 < 0 || shell0 + nshell > p->nbin || nshell > 65535 || !work ||
      !out || (fmt == BSK_FMT_BF16X2 && !out_lo) || (kind == BSK_KIND_DELTA && !dk) ||
      kind < BSK_KIND_DELTA || kind > BSK_KIND_KPOW || p->S % 4 ||
      shell_stride < level_layout(p->S).stride) {
    bsk_set_error("bsk_shells_levels: invalid argument (S a multiple of 4, shell stride of at "
                  "least the bsk_plan_levels stride)");
    return BSK_EINVAL;
  }
  const int stages = BSK_STAGE_X | BSK_STAGE_YZ | BSK_STAGE_LEVELS;
  if (dtype == BSK_DT_F32)
    return shells_t<float>(p, kind, (const cplx<float>*)dk, shell0, nshell, 0, p->S, stages, fmt,
                           out, out_lo, shell_stride, 0, work, as_stream(stream),
                           shell_tw(p, 0.f), parts);
  return shells_t<double>(p, kind, (const cplx<double>*)dk, shell0, nshell, 0, p->S, stages, fmt,
                          out, nullptr, shell_stride, 0, work, as_stream(stream),
                          shell_tw(p, 0.0), parts);
}

// ---------------------------------------------------------------- triples
static int spectrum_chunks(const bsk_plan* p) {
  const long long total = (long long)p->g.MX * p->g.MY * p->g.MZ;
  return (int)std::max(1LL, std::min(64LL, total / 4096));
}

int bsk_power_spectrum_workspace(const bsk_plan* p, size_t* bytes) {
  if (int rc = check_plan(p)) return rc;
  if (!bytes) {
    bsk_set_error("bsk_power_spectrum_workspace: null output");
    return BSK_EINVAL;
  }
  *bytes = sizeof(double) * 3 * (size_t)p->nbin * spectrum_chunks(p);
  return BSK_OK;
}

int bsk_power_spectrum(bsk_plan* p, int dtype, const void* dk, double* out_dev, void* work,
                       void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!dk || !out_dev || !work || (dtype != BSK_DT_F32 && dtype != BSK_DT_F64)) {
    bsk_set_error("bsk_power_spectrum: invalid argument");
    return BSK_EINVAL;
  }
  return bsk_spectrum_launch(dtype, dk, p->g, p->N, p->d_lo, p->d_hi, p->nbin,
                             2.0 * M_PI / p->box, spectrum_chunks(p),
                             reinterpret_cast<double*>(work), out_dev, as_stream(stream));
}

int bsk_triples_plain_workspace(int ntri, int nchunk, size_t* bytes) {
  *bytes = sizeof(double) * (size_t)std::max(ntri, 0) * std::max(nchunk, 1);
  return BSK_OK;
}

int bsk_triples_plain(int dtype, const void* fa, const void* fb, const void* fc,
                      long long shell_stride, long long ncell, int ntri, const int* tri_dev,
                      int nchunk, double scale, double* out_dev, void* work, void* stream) {
  if (!fa || !fb || !fc || !tri_dev || !out_dev || !work || nchunk <= 0 || ncell <= 0) {
    bsk_set_error("bsk_triples_plain: invalid argument");
    return BSK_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  int rc;
  if (dtype == BSK_DT_F32)
    rc = bsk_triple_plain_launch<float>((const float*)fa, (const float*)fb, (const float*)fc,
                                        shell_stride, shell_stride, shell_stride, ncell, ntri,
                                        tri_dev, part, nchunk, s);
  else if (dtype == BSK_DT_F64)
    rc = bsk_triple_plain_launch<double>((const double*)fa, (const double*)fb, (const double*)fc,
                                         shell_stride, shell_stride, shell_stride, ncell, ntri,
                                         tri_dev, part, nchunk, s);
  else {
    bsk_set_error("bsk_triples_plain: bad dtype");
    return BSK_EINVAL;
  }
  if (rc) return rc;
  return bsk_reduce_plain_launch(part, ntri, nchunk, scale, out_dev, s);
}

int bsk_triples_tiled_workspace(int ntri, int nsum, int nchunk, size_t* bytes) {
  *bytes = sizeof(double) * (size_t)std::max(ntri, 0) * nsum * std::max(nchunk, 1);
  return BSK_OK;
}

int bsk_triples_tiled(int dtype, const void* fields, long long shell_stride, int nfield,
                      long long ncell, int ntri, const int* tri_dev, int nsum, int nchunk,
                      double scale, double* out_dev, void* work, void* stream) {
  if (!fields || !tri_dev || !out_dev || !work || nchunk <= 0 || nfield <= 0 ||
      (nsum != 1 && nsum != 4) || (nsum == 4 && (nfield & 1))) {
    bsk_set_error("bsk_triples_tiled: invalid argument");
    return BSK_EINVAL;
  }
  hipStream_t s = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  int rc;
  if (dtype == BSK_DT_F32)
    rc = bsk_triple_tiled_launch<float>((const float*)fields, shell_stride, nfield, ncell, ntri,
                                        tri_dev, nsum, part, nchunk, s);
  else if (dtype == BSK_DT_F64)
    rc = bsk_triple_tiled_launch<double>((const double*)fields, shell_stride, nfield, ncell, ntri,
                                         tri_dev, nsum, part, nchunk, s);
  else {
    bsk_set_error("bsk_triples_tiled: bad dtype");
    return BSK_EINVAL;
  }
  if (rc) return rc;
  return bsk_reduce_tiled_launch(part, ntri, nsum, nchunk, scale, out_dev, s);
}

}  // extern "C"

// ------------------------------------------------------------ pair plans
// Parts of the level layout (fft_passes.h ShellLevels) in row order: P0 =
// C0 minus E, P1 = E, P2 = O, P3 = C1 minus O; the levels whose cells
// include each part (bit L): every level includes E.
constexpr int kParts = 4;
constexpr unsigned kPartLevels[kParts] = {0x3, 0xf, 0x5, 0x1};
static_assert(kParts == kPairSlots, "one partial slot per triangle and segment");

struct bsk_pairs {
  int mode = 0;
  int nshell = 0;
  int sets = 1;
  int ntri = 0;
  int npair = 0, ntile = 0, nitem = 0;
  int nkb = 0;
  int ni = 1;          // largest items per wave of a launch (compiled count)
  int npass = 1;       // launches in all
  // launch segments, in launch order: runs of consecutive parts with the same
  // levels (the first holds E, whose cells every level includes), each with
  // its own pair tiles and items over the triangles of its levels; per
  // segment the parts (bit p), the staged shell window (first, count) and the
  // launches
  int nseg = 1;
  int seg_parts[kParts] = {0xf, 0, 0, 0};
  int seg_win[kParts][2] = {{0, 1}, {0, 1}, {0, 1}, {0, 1}};
  int seg_pass[kParts] = {1, 0, 0, 0};
  int seg_ni[kParts] = {1, 1, 1, 1};  // items per wave of the segment's launches
  int seg_rep[kParts] = {1, 1, 1, 1};  // replicas of its wave lists (each runs 1/rep of the K steps)
  int item_total = 0;  // distinct items: partial positions per workgroup
  std::vector<int> slot;  // [ntri][kParts]: each triangle's slot per segment, -1 for none
  std::vector<int> h_items, h_outpos, h_pairs;
  int* d_items = nullptr;   // per launch [waves * ni of its segment]
  int* d_outpos = nullptr;  // the same shape
  int* d_pairs = nullptr;   // [ntile + 1][16]
  int* d_slot = nullptr;    // [ntri][kParts]
};

namespace {

struct Role {
  int a, b, c;  // pair (a in set 0, b in set 1), third side c in set 2
};

struct PairTiles {
  std::vector<std::pair<long long, unsigned>> pairs;  // (a << 32 | b, kb mask) in tile order
  std::vector<int> tile_of_pair_start;                // first pair index of each tile
  std::vector<unsigned> tile_mask;
  int nitem = 0;
};

long long pkey(int a, int b) { return ((long long)a << 32) | (unsigned)b; }

// Distinct pairs sorted by the set of third-side blocks they need (then by
// (a, b)) and packed into 16-row tiles in that order; a tile's items are the
// union of its rows' blocks.  exact: a new tile at every change of the block
// set (no MFMA on a block a row does not need); otherwise every tile is
// filled (fewer tiles, so fewer pair products and pair-side LDS reads).
PairTiles build_tiles(const std::vector<Role>& roles, bool exact) {
  std::vector<std::pair<long long, unsigned>> pm;
  pm.reserve(roles.size());
  for (const Role& r : roles) pm.push_back({pkey(r.a, r.b), 1u << (r.c / 16)});
  std::sort(pm.begin(), pm.end());
  std::vector<std::pair<long long, unsigned>> uniq;
  for (const auto& e : pm) {
    if (!uniq.empty() && uniq.back().first == e.first)
      uniq.back().second |= e.second;
    else
      uniq.push_back(e);
  }
  std::stable_sort(uniq.begin(), uniq.end(),
                   [](const std::pair<long long, unsigned>& x,
                      const std::pair<long long, unsigned>& y) { return x.second < y.second; });
  PairTiles t;
  t.pairs = uniq;
  for (size_t i = 0; i < uniq.size();) {
    size_t j = i;
    unsigned m = 0;
    while (j < uniq.size() && j - i < 16 && (!exact || uniq[j].second == uniq[i].second))
      m |= uniq[j++].second;
    t.tile_of_pair_start.push_back((int)i);
    t.tile_mask.push_back(m);
    t.nitem += __builtin_popcount(m);
    i = j;
  }
  return t;
}

// Relative kernel cost of a plan: a tile forms 16 pair products per cell
// (VALU) and reads both pair sides from LDS; an item reads one third-side
// block and issues three MFMAs.  Measured on gfx950 a tile costs about
// three items.
int plan_cost(const PairTiles& t) { return 3 * (int)t.tile_mask.size() + t.nitem; }

std::vector<Role> assign_roles(int mode, int option, int ntri, const int* tri) {
  std::vector<Role> r(ntri);
  for (int t = 0; t < ntri; ++t) {
    int i = tri[3 * t], j = tri[3 * t + 1], k = tri[3 * t + 2];
    if (mode == BSK_MODE_AUTO) {
      int v[3] = {i, j, k};
      std::sort(v, v + 3);  // v[0] <= v[1] <= v[2]
      if (option == 0) r[t] = {v[2], v[1], v[0]};       // pair = two largest
      else if (option == 1) r[t] = {v[1], v[0], v[2]};  // pair = two smallest
      else r[t] = {v[2], v[0], v[1]};                   // pair = largest and smallest
    } else if (mode == BSK_MODE_AAB) {
      r[t] = {std::max(i, j), std::min(i, j), k};
    } else {
      r[t] = {i, j, k};
    }
  }
  return r;
}

// Auto spectra, pair chosen per triangle: a greedy cover of the triangles by
// pairs (each step takes the pair that closes the most triangles not yet
// covered, ties to the smaller key; lazy re-counting), then every triangle
// takes the earliest chosen pair among those that close it.  Fewer distinct
// pairs than any single side pairing for typical bin lists (the 512^3 bench
// list: 233 against 334).
std::vector<Role> assign_roles_cover(int ntri, const int* tri) {
  std::vector<std::array<Role, 3>> cand(ntri);
  std::vector<int> ncand(ntri, 0);
  std::vector<std::pair<long long, int>> pt;  // (pair key, triangle)
  pt.reserve((size_t)3 * ntri);
  for (int t = 0; t < ntri; ++t) {
    int v[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
    std::sort(v, v + 3);
    const Role opts[3] = {{v[2], v[1], v[0]}, {v[1], v[0], v[2]}, {v[2], v[0], v[1]}};
    for (const Role& o : opts) {
      bool dup = false;
      for (int q = 0; q < ncand[t]; ++q)
        dup |= cand[t][q].a == o.a && cand[t][q].b == o.b && cand[t][q].c == o.c;
      if (dup) continue;
      cand[t][ncand[t]++] = o;
      pt.push_back({pkey(o.a, o.b), t});
    }
  }
  std::sort(pt.begin(), pt.end());
  std::vector<long long> keys;
  std::vector<size_t> start;  // triangles of key i: pt[start[i] .. start[i+1])
  for (size_t e = 0; e < pt.size(); ++e)
    if (e == 0 || pt[e].first != pt[e - 1].first) {
      keys.push_back(pt[e].first);
      start.push_back(e);
    }
  start.push_back(pt.size());
  std::vector<char> covered(ntri, 0);
  std::vector<int> rank(keys.size(), -1);
  std::priority_queue<std::pair<int, long long>> heap;  // (count, -key index)
  for (size_t i = 0; i < keys.size(); ++i)
    heap.push({(int)(start[i + 1] - start[i]), -(long long)i});
  int nrank = 0;
  while (!heap.empty()) {
    const auto top = heap.top();
    heap.pop();
    const size_t i = (size_t)(-top.second);
    int cnt = 0;
    for (size_t e = start[i]; e < start[i + 1]; ++e) cnt += !covered[pt[e].second];
    if (cnt == 0) continue;
    if (cnt < top.first) {
      heap.push({cnt, top.second});
      continue;
    }
    rank[i] = nrank++;
    for (size_t e = start[i]; e < start[i + 1]; ++e) covered[pt[e].second] = 1;
  }
  std::vector<Role> r(ntri);
  for (int t = 0; t < ntri; ++t) {
    int best = -1, br = INT_MAX;
    for (int q = 0; q < ncand[t]; ++q) {
      const long long k = pkey(cand[t][q].a, cand[t][q].b);
      const size_t i = (size_t)(std::lower_bound(keys.begin(), keys.end(), k) - keys.begin());
      if (rank[i] >= 0 && rank[i] < br) {
        br = rank[i];
        best = q;
      }
    }
    r[t] = cand[t][best];
  }
  return r;
}

int upload(int** dst, const std::vector<int>& v) {
  BSK_HIP_CHECK(hipMalloc(dst, sizeof(int) * std::max<size_t>(v.size(), 1)));
  if (!v.empty())
    BSK_HIP_CHECK(hipMemcpy(*dst, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice));
  return BSK_OK;
}

// One level's share of a pair plan: role choice, tiles and units (a tile
// with at most `cap` of its kb blocks; each unit forms the tile's pair
// products once), in absolute shell indices.
struct LevelPlan {
  std::vector<int> tri;  // global triangle indices of this level
  std::vector<Role> roles;
  PairTiles tiles;
  std::vector<int> unit_tile;
  std::vector<unsigned> unit_mask;
  int lo = 0, hi = 0;  // shells used: [lo, hi)
};

LevelPlan plan_level(int mode, int nkb, int cap, const std::vector<int>& idx, const int* tri) {
  LevelPlan lp;
  lp.tri = idx;
  std::vector<int> sub((size_t)3 * idx.size());
  lp.lo = INT_MAX;
  for (size_t t = 0; t < idx.size(); ++t)
    for (int e = 0; e < 3; ++e) {
      sub[3 * t + e] = tri[3 * (size_t)idx[t] + e];
      lp.lo = std::min(lp.lo, sub[3 * t + e]);
      lp.hi = std::max(lp.hi, sub[3 * t + e] + 1);
    }
  if (idx.empty()) lp.lo = 0;
  const int ntri = (int)idx.size();
  // role assignment: the fixed roles of <AAB>/<ABC>; for auto spectra the
  // choice of the pair sides that needs the fewest items
  const int nopt = mode == BSK_MODE_AUTO ? 4 : 1;
  bool first = true;
  for (int o = 0; o < nopt; ++o) {
    std::vector<Role> r =
        o == 3 ? assign_roles_cover(ntri, sub.data()) : assign_roles(mode, o, ntri, sub.data());
    for (int exact = 1; exact >= 0; --exact) {
      PairTiles t = build_tiles(r, exact != 0);
      if (first || plan_cost(t) < plan_cost(lp.tiles)) {
        lp.roles = r;
        lp.tiles = t;
        first = false;
      }
    }
  }
  const int ntile = (int)lp.tiles.tile_mask.size();
  for (int ti = 0; ti < ntile; ++ti) {
    unsigned m = 0;
    int n = 0;
    for (int kb = 0; kb < nkb; ++kb) {
      if (!(lp.tiles.tile_mask[ti] >> kb & 1u)) continue;
      m |= 1u << kb;
      if (++n == cap) {
        lp.unit_tile.push_back(ti);
        lp.unit_mask.push_back(m);
        m = 0;
        n = 0;
      }
    }
    if (m) {
      lp.unit_tile.push_back(ti);
      lp.unit_mask.push_back(m);
    }
  }
  return lp;
}

// Units (tile, kb mask) to (pass, wave) bins, costliest first (forming a
// tile's products weighs three items), each to the cheapest bin with room for
// its items; the fewest passes for which every unit fits.
std::vector<std::vector<int>> bin_units(const std::vector<unsigned>& mask, int cap, int W,
                                        int* npass, int* maxload) {
  const int nunit = (int)mask.size();
  int nitem = 0;
  for (unsigned m : mask) {
    nitem += __builtin_popcount(m);
    cap = std::max(cap, __builtin_popcount(m));  // every unit fits an empty bin
  }
  auto ucost = [&](int u) { return 3 + __builtin_popcount(mask[u]); };
  std::vector<int> order(nunit);
  for (int i = 0; i < nunit; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ucost(x) > ucost(y); });
  std::vector<std::vector<int>> bins;
  for (*npass = std::max(1, (nitem + W * cap - 1) / (W * cap));; ++*npass) {
    bins.assign((size_t)*npass * W, {});
    std::vector<int> load((size_t)*npass * W, 0), cost((size_t)*npass * W, 0);
    bool ok = true;
    for (int u : order) {
      const int n = __builtin_popcount(mask[u]);
      int b = -1;
      for (int q = 0; q < (int)bins.size(); ++q)
        if (load[q] + n <= cap && (b < 0 || cost[q] < cost[b])) b = q;
      if (b < 0) {
        ok = false;
        break;
      }
      load[b] += n;
      cost[b] += ucost(u);
      bins[b].push_back(u);
    }
    if (ok) break;
  }
  *maxload = 1;
  for (const auto& b : bins) {
    int l = 0;
    for (int u : b) l += __builtin_popcount(mask[u]);
    *maxload = std::max(*maxload, l);
  }
  return bins;
}

int pairs_create(bsk_pairs** out, int mode, int nshell, int ntri, const int* tri,
                 const int* level, const char* fn) {
  if (!out || nshell < 1 || mode < 0 || mode > 2 || ntri < 0 || (ntri && !tri)) {
    bsk_set_error("%s: invalid argument", fn);
    return BSK_EINVAL;
  }
  *out = nullptr;
  for (int t = 0; t < ntri; ++t) {
    for (int e = 0; e < 3; ++e)
      if (tri[3 * t + e] < 0 || tri[3 * t + e] >= nshell) {
        bsk_set_error("%s: triangle %d has a shell index out of range", fn, t);
        return BSK_EINVAL;
      }
    if (level && (level[t] < 0 || level[t] >= BSK_LEVELS)) {
      bsk_set_error("%s: triangle %d has level %d (0 to %d)", fn, t, level[t], BSK_LEVELS - 1);
      return BSK_EINVAL;
    }
  }
  const int sets = mode == BSK_MODE_AUTO ? 1 : (mode == BSK_MODE_AAB ? 2 : 3);
  const int nkb = (nshell + 15) / 16;
  int nsh[3] = {nshell, nshell, nshell};
  if (!bsk_pair_prefetch(nshell) || bsk_pair_lds_bytes(sets, nsh, nkb + 1) > 160 * 1024 ||
      nkb > 16 || nshell >= 0xffff) {
    bsk_set_error("%s: %d shells exceed the MFMA kernel's LDS staging "
                  "(128 shells, 96 for <ABC>)", fn, nshell);
    return BSK_EUNSUPPORTED;
  }
  // at most 4 items per wave per launch keeps two workgroups per CU (more
  // items of a segment take more launches over its cells); units (a tile's
  // items formed from one set of pair products) hold at most that many
  const int W = kPairWaves, cap = std::min(4, bsk_pair_item_cap(sets, nshell));
  std::vector<int> idx[BSK_LEVELS];
  for (int t = 0; t < ntri; ++t) idx[level ? level[t] : 0].push_back(t);
  unsigned present = 0;
  for (int L = 0; L < BSK_LEVELS; ++L)
    if (!idx[L].empty()) present |= 1u << L;
  // segments: runs of consecutive parts with the same present levels, the
  // one holding E (every level) first
  struct Seg {
    int parts;
    unsigned lv;
  };
  std::vector<Seg> segs;
  for (int p = 0; p < kParts; ++p) {
    const unsigned lv = kPartLevels[p] & present;
    if (!lv) continue;
    if (!segs.empty() && segs.back().lv == lv && (segs.back().parts >> (p - 1) & 1))
      segs.back().parts |= 1 << p;
    else
      segs.push_back({1 << p, lv});
  }
  if (segs.empty()) segs.push_back({0xf, 1u});  // empty list: one pass of padding
  std::stable_partition(segs.begin(), segs.end(), [](const Seg& g) { return (g.parts >> 1) & 1; });
  // per segment one plan over the triangles of all its levels: a pair (a, b)
  // that triangles of several levels need is formed once on the segment's
  // cells (on E, whose cells every level includes, the levels share most
  // pairs: at 512^3 / 32 bins 15 tiles instead of 27 planned level by level)
  const int nseg = (int)segs.size();
  std::vector<LevelPlan> lp(nseg);
  std::vector<int> tile0(nseg + 1, 0);
  for (int g = 0; g < nseg; ++g) {
    std::vector<int> sel;
    for (int L = 0; L < BSK_LEVELS; ++L)
      if (segs[g].lv >> L & 1) sel.insert(sel.end(), idx[L].begin(), idx[L].end());
    std::sort(sel.begin(), sel.end());
    lp[g] = plan_level(mode, nkb, cap, sel, tri);
    tile0[g + 1] = tile0[g] + (int)lp[g].tiles.tile_mask.size();
  }
  const int ntile = tile0[nseg];
  if (ntile >= 0xffff) {
    bsk_set_error("%s: %d pair tiles exceed 65534; split the triangle list", fn, ntile);
    return BSK_EUNSUPPORTED;
  }
  // partial positions: one per (segment tile, kb) item, each stored by its
  // segment's launches alone
  std::vector<int> item_of((size_t)ntile * nkb, -1);
  int nitem = 0;
  for (int g = 0; g < nseg; ++g)
    for (int ti = 0; ti < (int)lp[g].tiles.tile_mask.size(); ++ti)
      for (int kb = 0; kb < nkb; ++kb)
        if (lp[g].tiles.tile_mask[ti] >> kb & 1u) item_of[(size_t)(tile0[g] + ti) * nkb + kb] = nitem++;
  if ((long long)std::max(nitem, 1) * 256 >= (1LL << 28)) {  // slot field below the level bits
    bsk_set_error("%s: %d items exceed the slot range; split the triangle list", fn, nitem);
    return BSK_EUNSUPPORTED;
  }
  // per segment: its units binned into passes
  struct SegPlan {
    int lo, nsh, npass, ni, rep = 1;
    std::vector<int> unit_tile;  // global tile
    std::vector<unsigned> unit_mask;
    std::vector<std::vector<int>> bins;
  };
  std::vector<SegPlan> sp(nseg);
  int maxni = 1;
  size_t nentry = 0;
  for (int g = 0; g < nseg; ++g) {
    for (size_t u = 0; u < lp[g].unit_tile.size(); ++u) {
      sp[g].unit_tile.push_back(tile0[g] + lp[g].unit_tile[u]);
      sp[g].unit_mask.push_back(lp[g].unit_mask[u]);
    }
    const bool any = !lp[g].tri.empty();
    sp[g].lo = any ? lp[g].lo : 0;
    sp[g].nsh = any ? lp[g].hi - lp[g].lo : 1;
    int ml = 1;
    if (sp[g].unit_mask.empty()) {
      sp[g].npass = 1;
      sp[g].bins.assign(W, {});
    } else {
      sp[g].bins = bin_units(sp[g].unit_mask, cap, W, &sp[g].npass, &ml);
      // few items (a plan over levels): lists on W / rep waves, replicated
      // rep times, each replica running every rep-th K step
      if (level)
        for (int rep : {4, 2}) {
          int np = 0, m2 = 1;
          auto b2 = bin_units(sp[g].unit_mask, cap, W / rep, &np, &m2);
          if (np == 1 && sp[g].npass == 1) {
            sp[g].bins.clear();
            for (int r = 0; r < rep; ++r) sp[g].bins.insert(sp[g].bins.end(), b2.begin(), b2.end());
            sp[g].rep = rep;
            ml = m2;
            break;
          }
        }
      // a replicated segment with 3 or 4 items per wave: lists of at most 2
      // items on half the waves when they fit, which lets three workgroups
      // share a CU instead of two (its staging latency, not its math, binds)
      if (level && sp[g].rep > 1 && ml > 2) {
        int np = 0, m2 = 1;
        auto b2 = bin_units(sp[g].unit_mask, 2, W / 2, &np, &m2);
        if (np == 1 && m2 <= 2) {
          sp[g].bins.clear();
          for (int r = 0; r < 2; ++r) sp[g].bins.insert(sp[g].bins.end(), b2.begin(), b2.end());
          sp[g].rep = 2;
          ml = m2;
        }
      }
    }
    sp[g].ni = bsk_pair_items_round(ml);
    maxni = std::max(maxni, sp[g].ni);
    nentry += (size_t)sp[g].npass * W * sp[g].ni;
  }
  int npass = 0;
  for (const auto& g : sp) npass += g.npass;
  // launch lists: descriptors tile | kb << 16 | padding << 29 | (first item
  // of a unit: new pair products) << 31, kb counted from the segment's first
  // staged block; partial positions (-1: padding); each launch holds waves x
  // the segment's items per wave
  std::vector<int> items(nentry), outpos(nentry, -1);
  const int pad = ntile | (1 << 29);
  size_t seg0 = 0;
  for (const auto& g : sp) {
    const int kb0 = g.lo / 16, ni = g.ni;
    for (size_t b = 0; b < g.bins.size(); ++b) {
      const size_t e0 = seg0 + b * ni;
      int s = 0;
      for (int u : g.bins[b]) {
        const int ti = g.unit_tile[u];
        bool first = true;
        for (int kb = 0; kb < nkb; ++kb) {
          if (!(g.unit_mask[u] >> kb & 1u)) continue;
          items[e0 + s] = ti | ((kb - kb0) << 16) | (first ? (int)0x80000000u : 0);
          outpos[e0 + s] = item_of[(size_t)ti * nkb + kb];
          first = false;
          ++s;
        }
      }
      for (; s < ni; ++s)  // padding: skipped by the kernel
        items[e0 + s] = pad | (s == 0 ? (int)0x80000000u : 0);
    }
    seg0 += (size_t)g.npass * W * ni;
  }
  // pair rows in absolute shell indices, 0xffff for padding rows (the
  // kernel's zero row); the last tile is all padding
  std::vector<int> prs((size_t)(ntile + 1) * 16, 0xffff | (0xffff << 16));
  bsk_pairs* q = new bsk_pairs();
  // triangle slots, one per segment (-1: the triangle's level does not
  // include the segment's parts; the first segment holds every triangle)
  q->slot.assign((size_t)kParts * ntri, -1);
  int npair = 0;
  for (int g = 0; g < nseg; ++g) {
    const PairTiles& tl = lp[g].tiles;
    const int nt = (int)tl.tile_mask.size();
    std::vector<std::pair<int, int>> where(tl.pairs.size());  // (global tile, row)
    for (int ti = 0; ti < nt; ++ti) {
      const int p0 = tl.tile_of_pair_start[ti];
      const int p1 = ti + 1 < nt ? tl.tile_of_pair_start[ti + 1] : (int)tl.pairs.size();
      for (int p = p0; p < p1; ++p) {
        const long long k = tl.pairs[p].first;
        prs[(size_t)(tile0[g] + ti) * 16 + (p - p0)] =
            (int)(k >> 32) | ((int)(k & 0xffffffff) << 16);
        where[p] = {tile0[g] + ti, p - p0};
      }
    }
    // slot: item * 256 + row * 16 + column, the triangle's level in bits 28-29
    std::vector<long long> keys(tl.pairs.size());
    for (size_t p = 0; p < keys.size(); ++p) keys[p] = tl.pairs[p].first;
    std::vector<size_t> kidx(keys.size());
    for (size_t p = 0; p < kidx.size(); ++p) kidx[p] = p;
    std::sort(kidx.begin(), kidx.end(), [&](size_t x, size_t y) { return keys[x] < keys[y]; });
    for (size_t t = 0; t < lp[g].tri.size(); ++t) {
      const Role& r = lp[g].roles[t];
      const long long k = pkey(r.a, r.b);
      auto it = std::lower_bound(kidx.begin(), kidx.end(), k,
                                 [&](size_t x, long long v) { return keys[x] < v; });
      const int p = (int)*it;
      const int ti = where[p].first, row = where[p].second, c = r.c;
      const int it_ = item_of[(size_t)ti * nkb + c / 16];
      const int gt = lp[g].tri[t];
      const int L = level ? level[gt] : 0;
      q->slot[(size_t)kParts * gt + g] = it_ * 256 + row * 16 + (c % 16) + (L << 28);
    }
    npair += (int)tl.pairs.size();
  }
  q->mode = mode;
  q->nshell = nshell;
  q->sets = sets;
  q->ntri = ntri;
  q->npair = npair;
  q->ntile = ntile;
  q->nitem = nitem;
  q->nkb = nkb;
  q->ni = maxni;
  q->npass = npass;
  q->nseg = (int)sp.size();
  for (int g = 0; g < q->nseg; ++g) {
    q->seg_parts[g] = segs[g].parts;
    q->seg_win[g][0] = sp[g].lo;
    q->seg_win[g][1] = sp[g].nsh;
    q->seg_pass[g] = sp[g].npass;
    q->seg_ni[g] = sp[g].ni;
    q->seg_rep[g] = sp[g].rep;
  }
  q->item_total = std::max(nitem, 1);
  q->h_items.swap(items);
  q->h_outpos.swap(outpos);
  q->h_pairs.swap(prs);
  *out = q;
  return BSK_OK;
}

}  // namespace

extern "C" {

int bsk_pairs_create(bsk_pairs** out, int mode, int nshell, int ntri, const int* tri) {
  return pairs_create(out, mode, nshell, ntri, tri, nullptr, "bsk_pairs_create");
}

int bsk_pairs_create_levels(bsk_pairs** out, int mode, int nshell, int ntri, const int* tri,
                            const int* level) {
  if (ntri > 0 && !level) {
    bsk_set_error("bsk_pairs_create_levels: null level list");
    return BSK_EINVAL;
  }
  return pairs_create(out, mode, nshell, ntri, tri, level, "bsk_pairs_create_levels");
}

int bsk_pairs_segments(const bsk_pairs* q, int* nseg, int* parts, int* first, int* count,
                       int* npass, int* items_per_wave, int* replicas) {
  if (!q || !nseg || !parts || !first || !count || !npass || !items_per_wave || !replicas) {
    bsk_set_error("bsk_pairs_segments: invalid argument");
    return BSK_EINVAL;
  }
  *nseg = q->nseg;
  for (int g = 0; g < kParts; ++g) {
    const bool in = g < q->nseg;
    parts[g] = in ? q->seg_parts[g] : 0;
    first[g] = in ? q->seg_win[g][0] : 0;
    count[g] = in ? q->seg_win[g][1] : 0;
    npass[g] = in ? q->seg_pass[g] : 0;
    items_per_wave[g] = in ? q->seg_ni[g] : 0;
    replicas[g] = in ? q->seg_rep[g] : 0;
  }
  return BSK_OK;
}

int bsk_pairs_shell_parts(const bsk_pairs* q, unsigned char* parts) {
  if (!q || !parts) {
    bsk_set_error("bsk_pairs_shell_parts: invalid argument");
    return BSK_EINVAL;
  }
  std::fill(parts, parts + q->nshell, (unsigned char)0);
  for (int g = 0; g < q->nseg; ++g)
    for (int b = q->seg_win[g][0]; b < q->seg_win[g][0] + q->seg_win[g][1]; ++b)
      parts[b] |= (unsigned char)q->seg_parts[g];
  return BSK_OK;
}

int bsk_pairs_upload(bsk_pairs* q) {
  if (!q) {
    bsk_set_error("bsk_pairs_upload: null plan");
    return BSK_EINVAL;
  }
  if (q->d_items) return BSK_OK;
  int rc = upload(&q->d_items, q->h_items);
  if (!rc) rc = upload(&q->d_outpos, q->h_outpos);
  if (!rc) rc = upload(&q->d_pairs, q->h_pairs);
  if (!rc) rc = upload(&q->d_slot, q->slot);
  if (rc) {
    (void)hipFree(q->d_items);
    (void)hipFree(q->d_outpos);
    (void)hipFree(q->d_pairs);
    (void)hipFree(q->d_slot);
    q->d_items = q->d_outpos = q->d_pairs = q->d_slot = nullptr;
  }
  return rc;
}

int bsk_pairs_destroy(bsk_pairs* q) {
  if (!q) return BSK_OK;
  (void)hipFree(q->d_items);
  (void)hipFree(q->d_outpos);
  (void)hipFree(q->d_pairs);
  (void)hipFree(q->d_slot);
  delete q;
  return BSK_OK;
}

int bsk_pairs_info(const bsk_pairs* q, int* npair, int* ntile, int* nitem, int* items_per_wave,
                   int* npass) {
  if (!q) {
    bsk_set_error("bsk_pairs_info: null plan");
    return BSK_EINVAL;
  }
  if (npair) *npair = q->npair;
  if (ntile) *ntile = q->ntile;
  if (nitem) *nitem = q->nitem;
  if (items_per_wave) *items_per_wave = q->ni;
  if (npass) *npass = q->npass;
  return BSK_OK;
}

int bsk_pairs_slots(const bsk_pairs* q, int* slot_host) {
  if (!q || (q->ntri && !slot_host)) {
    bsk_set_error("bsk_pairs_slots: invalid argument");
    return BSK_EINVAL;
  }
  std::copy(q->slot.begin(), q->slot.end(), slot_host);
  return BSK_OK;
}

int bsk_pairs_tables(const bsk_pairs* q, int* items, int* outpos, int* pairs) {
  if (!q) {
    bsk_set_error("bsk_pairs_tables: null plan");
    return BSK_EINVAL;
  }
  if (items) std::copy(q->h_items.begin(), q->h_items.end(), items);
  if (outpos) std::copy(q->h_outpos.begin(), q->h_outpos.end(), outpos);
  if (pairs) std::copy(q->h_pairs.begin(), q->h_pairs.end(), pairs);
  return BSK_OK;
}

int bsk_triples_mfma_workspace(const bsk_pairs* q, int nwg, size_t* bytes) {
  if (!q || nwg <= 0 || !bytes) {
    bsk_set_error("bsk_triples_mfma_workspace: invalid argument");
    return BSK_EINVAL;
  }
  // per-workgroup FP32 partials, then the reduction's FP64 range sums
  *bytes = sizeof(float) * 256 * (size_t)q->item_total * nwg + bsk_reduce_mfma_bytes(q->item_total);
  return BSK_OK;
}

namespace {
double* reduce_scratch(void* work, const bsk_pairs* q, int nwg) {
  return reinterpret_cast<double*>(static_cast<char*>(work) +
                                   sizeof(float) * 256 * (size_t)q->item_total * nwg);
}
}  // namespace

namespace {
int triples_mfma(const bsk_pairs* q, const uint16_t* hi_i, const uint16_t* lo_i,
                 const uint16_t* hi_j, const uint16_t* lo_j, const uint16_t* hi_k,
                 const uint16_t* lo_k, long long shell_stride, long long ncell,
                 const long long* level_off, const long long* level_ncell, int nwg,
                 double scale, double* out_dev, void* work, void* stream, const char* fn) {
  if (!q || !hi_j || !lo_j || !work || nwg <= 0 || shell_stride < ncell ||
      (q->mode >= 1 && (!hi_k || !lo_k)) || (q->mode == 2 && (!hi_i || !lo_i))) {
    bsk_set_error("%s: invalid argument", fn);
    return BSK_EINVAL;
  }
  // cell range of each layout part: P1 = level 3, P0 = the rest of level 1
  // before it, P2 = the rest of level 2 after it, P3 = the rest of level 0
  long long pbeg[kParts], pend[kParts];
  if (level_off && level_ncell) {
    const long long* o = level_off;
    const long long* n = level_ncell;
    pbeg[0] = o[1], pend[0] = o[3];
    pbeg[1] = o[3], pend[1] = o[3] + n[3];
    pbeg[2] = o[3] + n[3], pend[2] = o[2] + n[2];
    pbeg[3] = o[2] + n[2], pend[3] = o[0] + n[0];
    bool ok = o[0] == o[1] && o[2] == o[3] && o[0] <= o[3] && o[0] + n[0] <= shell_stride;
    for (int p = 0; p < kParts; ++p)
      ok = ok && pbeg[p] <= pend[p] && pbeg[p] % kPairChunk == 0 && pend[p] % kPairChunk == 0;
    if (!ok) {
      bsk_set_error("%s: level ranges do not nest as the level layout's (bsk_plan_levels)", fn);
      return BSK_EINVAL;
    }
  } else {
    if (q->nseg != 1 || q->seg_parts[0] != 0xf) {
      bsk_set_error("%s: the plan has sublattice-level triangles; use bsk_triples_mfma_levels "
                    "with the bsk_plan_levels layout", fn);
      return BSK_EINVAL;
    }
    // one segment of all parts: the whole C-order range
    std::fill(pbeg, pbeg + kParts, 0LL);
    std::fill(pend, pend + kParts, ncell);
  }
  if (q->ntri == 0) return BSK_OK;
  if (!q->d_items) {
    bsk_set_error("%s: pair plan not uploaded (bsk_pairs_upload)", fn);
    return BSK_EINVAL;
  }
  PairArgs a;
  // roles: set 0 / 1 = pair sides, set 2 = third side
  const uint16_t* hi[3] = {q->mode == 2 ? hi_i : hi_j, hi_j, q->mode >= 1 ? hi_k : hi_j};
  const uint16_t* lo[3] = {q->mode == 2 ? lo_i : lo_j, lo_j, q->mode >= 1 ? lo_k : lo_j};
  a.stride = shell_stride;
  a.nwg = nwg;
  a.pairs = q->d_pairs;
  a.partial = reinterpret_cast<float*>(work);
  a.item_total = q->item_total;
  hipStream_t s = as_stream(stream);
  size_t e0 = 0;  // first list entry of the launch
  for (int g = 0; g < q->nseg; ++g) {
    // the segment's cells (its parts are consecutive) of its shell window
    const int fp = __builtin_ctz(q->seg_parts[g]), lp = 31 - __builtin_clz(q->seg_parts[g]);
    const long long c0 = pbeg[fp], c1 = pend[lp];
    const int w0 = q->seg_win[g][0], wn = q->seg_win[g][1];
    const long long base = c0 + (long long)w0 * shell_stride;
    for (int r = 0; r < 3; ++r) {
      a.hi[r] = hi[r] + base;
      a.lo[r] = lo[r] + base;
      a.nsh[r] = wn;
    }
    a.lo_row = w0;
    a.koff = w0 % 16;
    a.nkb = (a.koff + wn + 15) / 16;
    a.ncell = c1 - c0;  // 0 (an empty part) still stores the segment's zero partials
    a.accumulate = 0;   // every item belongs to one segment
    a.rep = q->seg_rep[g];
    const int ni = q->seg_ni[g];
    const int np = q->seg_pass[g];
    // a segment's passes (item lists over the same cells) in one launch,
    // the workgroups of one cell range on one XCD, when the count allows
    const bool fold = np > 1 && a.rep == 1 && nwg % 8 == 0;
    a.npass = fold ? np : 1;
    a.pass_stride = kPairWaves * ni;
    for (int k = 0; k < np; k += a.npass) {
      a.items = q->d_items + e0 + (size_t)k * kPairWaves * ni;
      a.outpos = q->d_outpos + e0 + (size_t)k * kPairWaves * ni;
      int rc = bsk_triple_pairs_launch(a, q->sets, ni, s);
      if (rc) return rc;
    }
    e0 += (size_t)np * kPairWaves * ni;
  }
  if (!out_dev) return BSK_OK;  // partials only; bsk_triples_mfma_reduce finishes
  return bsk_reduce_mfma_launch(a.partial, nwg, q->item_total, q->ntri, q->d_slot, scale,
                                out_dev, reduce_scratch(work, q, nwg), s);
}
}  // namespace

int bsk_triples_mfma(const bsk_pairs* q, const uint16_t* hi_i, const uint16_t* lo_i,
                     const uint16_t* hi_j, const uint16_t* lo_j, const uint16_t* hi_k,
                     const uint16_t* lo_k, long long shell_stride, long long ncell, int nwg,
                     double scale, double* out_dev, void* work, void* stream) {
  return triples_mfma(q, hi_i, lo_i, hi_j, lo_j, hi_k, lo_k, shell_stride, ncell, nullptr,
                      nullptr, nwg, scale, out_dev, work, stream, "bsk_triples_mfma");
}

int bsk_triples_mfma_levels(const bsk_pairs* q, const uint16_t* hi_i, const uint16_t* lo_i,
                            const uint16_t* hi_j, const uint16_t* lo_j, const uint16_t* hi_k,
                            const uint16_t* lo_k, long long shell_stride,
                            const long long* level_off, const long long* level_ncell, int nwg,
                            double scale, double* out_dev, void* work, void* stream) {
  if (!level_off || !level_ncell) {
    bsk_set_error("bsk_triples_mfma_levels: null level layout");
    return BSK_EINVAL;
  }
  return triples_mfma(q, hi_i, lo_i, hi_j, lo_j, hi_k, lo_k, shell_stride, level_ncell[0],
                      level_off, level_ncell, nwg, scale, out_dev, work, stream,
                      "bsk_triples_mfma_levels");
}

// ------------------------------------------------------------------
// The same sums with the library's own scratch memory: a file-scope device
// array holds the per-workgroup partials and the range sums, so the call
// allocates nothing (for callers that cannot allocate: code under graph
// capture, tight serving loops).  Each device holds its own copy; its
// address is looked up once per device with that device current and kept.
namespace {
static __device__ __align__(256) unsigned char scratch[BSK_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

int scratch_on_current_device(void** out) {
  int dev = -1;
  BSK_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    bsk_set_error("bsk_triples_mfma_levels_noalloc: device %d beyond the %d supported", dev,
                  kMaxDevices);
    return BSK_EUNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    BSK_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return BSK_OK;
}
}  // namespace

int bsk_triples_mfma_levels_noalloc(const bsk_pairs* q, const uint16_t* hi_i,
                                    const uint16_t* lo_i, const uint16_t* hi_j,
                                    const uint16_t* lo_j, const uint16_t* hi_k,
                                    const uint16_t* lo_k, long long shell_stride,
                                    const long long* level_off, const long long* level_ncell,
                                    int nwg, double scale, double* out_dev, void* stream) {
  const char* fn = "bsk_triples_mfma_levels_noalloc";
  if (!q || nwg <= 0 || !out_dev || !level_off || !level_ncell) {
    bsk_set_error("%s: invalid argument (out_dev and the level layout are required)", fn);
    return BSK_EINVAL;
  }
  size_t need = 0;
  if (int rc = bsk_triples_mfma_workspace(q, nwg, &need)) return rc;
  if (need > (size_t)BSK_NOALLOC_SCRATCH_BYTES) {
    bsk_set_error("%s: the plan needs %zu scratch bytes at %d workgroups, above the library's "
                  "%zu (fewer workgroups, or bsk_triples_mfma_levels with a workspace)",
                  fn, need, nwg, (size_t)BSK_NOALLOC_SCRATCH_BYTES);
    return BSK_EINVAL;
  }
  void* work = nullptr;
  if (int rc = scratch_on_current_device(&work)) return rc;
  // every partial slot a triangle reads is stored by the first launch before
  // any other reads it, and the range sums are written before they are read,
  // so nothing in the scratch needs clearing between calls
  return triples_mfma(q, hi_i, lo_i, hi_j, lo_j, hi_k, lo_k, shell_stride, level_ncell[0],
                      level_off, level_ncell, nwg, scale, out_dev, work, stream, fn);
}

namespace {
// Why the fused z/triple kernel cannot run this plan (nullptr: it can).
const char* zfused_gap(const bsk_plan* p, const bsk_pairs* q) {
  if (!p || !q) return "null plan";
  if (q->mode != BSK_MODE_AUTO) return "auto spectra only";
  if (q->nseg != 1 || q->seg_parts[0] != 0xf) return "the plan has sublattice-level triangles";
  const int nsh = q->seg_win[0][1];  // shells staged: the plan's window
  if (nsh > kZMaxShells) return "too many shells";
  if (q->seg_ni[0] < 1 || q->seg_ni[0] > 4) return "items per wave above 4";
  if (q->seg_rep[0] != 1) return "replicated wave lists";
  bool size = false;
#define C(n) size |= p->S == n;
  BSK_ZFUSED_SIZES(C)
#undef C
  if (!size) return "no kernel for this shell mesh";
  if (bsk_zrows_lds_bytes(p->S, nsh, (q->seg_win[0][0] % 16 + nsh + 15) / 16) > 160 * 1024)
    return "LDS";
  return nullptr;
}
}  // namespace

int bsk_triples_zfused_supported(const bsk_plan* p, const bsk_pairs* q, int* ok) {
  if (!ok) {
    bsk_set_error("bsk_triples_zfused_supported: null result");
    return BSK_EINVAL;
  }
  *ok = zfused_gap(p, q) == nullptr;
  return BSK_OK;
}

int bsk_triples_zfused(bsk_plan* p, const bsk_pairs* q, const void* shell_work, int shell0,
                       int nwg, double scale, double* out_dev, void* work, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!q || !shell_work || !work || nwg <= 0 || shell0 < 0 || shell0 + q->nshell > p->nbin) {
    bsk_set_error("bsk_triples_zfused: invalid argument");
    return BSK_EINVAL;
  }
  if (const char* why = zfused_gap(p, q)) {
    bsk_set_error("bsk_triples_zfused: unsupported plan (%s)", why);
    return BSK_EUNSUPPORTED;
  }
  if (q->ntri == 0) return BSK_OK;
  if (!q->d_items) {
    bsk_set_error("bsk_triples_zfused: pair plan not uploaded (bsk_pairs_upload)");
    return BSK_EINVAL;
  }
  const Geom& g = p->g;
  const int S = p->S;
  ZArgs a;
  // the plan's shell window [lo, lo + nsh) of the shells of the workspace
  const int lo = q->seg_win[0][0], nsh = q->seg_win[0][1];
  a.d2 = reinterpret_cast<const cplx<float>*>(shell_work) +
         (size_t)q->nshell * g.MY * S * g.MZP +  // after D1 (shells_t layout)
         (size_t)lo * S * S * g.MZP;
  a.nrows = (long long)S * S;
  a.MZ = g.MZ;
  a.MZP = g.MZP;
  a.n2hi = p->d_hi;
  a.shell0 = shell0 + lo;
  a.nsh = nsh;
  a.lo_row = lo;
  a.koff = lo % 16;
  a.nkb = (a.koff + nsh + 15) / 16;
  a.nwg = nwg;
  a.pairs = q->d_pairs;
  a.partial = reinterpret_cast<float*>(work);
  a.item_total = q->item_total;
  a.tw = shell_tw(p, 0.f);
  hipStream_t s = as_stream(stream);
  for (int pass = 0; pass < q->npass; ++pass) {
    const size_t e0 = (size_t)pass * kPairWaves * q->seg_ni[0];
    a.items = q->d_items + e0;
    a.outpos = q->d_outpos + e0;
    const int rc = p->variant[5] ? bsk_triple_zrows_launch_v1(a, S, q->seg_ni[0], s)
                                 : bsk_triple_zrows_launch_v0(a, S, q->seg_ni[0], s);
    if (rc) return rc;
  }
  if (!out_dev) return BSK_OK;
  return bsk_reduce_mfma_launch(a.partial, nwg, q->item_total, q->ntri, q->d_slot, scale,
                                out_dev, reduce_scratch(work, q, nwg), s);
}

int bsk_triples_mfma_reduce(const bsk_pairs* q, int nwg, double scale, double* out_dev,
                            void* work, void* stream) {
  if (!q || nwg <= 0 || !out_dev || !work || !q->d_slot) {
    bsk_set_error("bsk_triples_mfma_reduce: invalid argument (plan uploaded?)");
    return BSK_EINVAL;
  }
  return bsk_reduce_mfma_launch(reinterpret_cast<const float*>(work), nwg, q->item_total,
                                q->ntri, q->d_slot, scale, out_dev, reduce_scratch(work, q, nwg),
                                as_stream(stream));
}

}  // extern "C"
