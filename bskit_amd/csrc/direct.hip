// Direct enumeration of closed triangles from explicit mode lists.
//
// Replaces the O(M^3) Python loop of bk_binned (bskit/main.py:834-1004):
// for every k1 in list A and k2 in list B the third side k3 = -(k1+k2) is
// formed (wrapped into [-N/2, N/2) when periodic, else kept as is and
// rejected unless it is a mode the reference collects: the stored half grid
// x, y in [-N/2, N/2), 0 <= z <= N/2, or a conjugate -n of a mode with
// 0 < n_z < N/2) and accepted when its integer |n|^2 lies in [lo_c, hi_c].  Each accepted triple adds 1 to the count, |k1|,|k2|,|k3|
// (k_f units) to the k sums and dA(k1) dB(k2) dC(k3) to the complex sum.
// delta values come from a compact half cube [ry][rx][kz] (kz >= 0) with
// delta(-k) = conj(delta(k)).  Per-workgroup partials (6 doubles) are written
// in a fixed layout; the caller sums them.
#include "common.h"

namespace {

constexpr int kW = 256;

struct Cube {
  const double2* d;  // may be null (counting only)
  int N, MX, MZ, nlo;
};

BSK_DEV int wrapc(int v, int N) {
  const int h = N / 2;
  int r = (v + h) % N;
  if (r < 0) r += N;
  return r - h;
}

BSK_DEV double2 lookup(const Cube& c, int x, int y, int z) {
  bool conj = false;
  if (z < 0) {
    x = -x;
    y = -y;
    z = -z;
    conj = true;
  }
  x = wrapc(x, c.N);
  y = wrapc(y, c.N);
  const int rx = x - c.nlo, ry = y - c.nlo;
  if (rx < 0 || rx >= c.MX || ry < 0 || ry >= c.MX || z >= c.MZ) return make_double2(0.0, 0.0);
  double2 v = c.d[((long long)ry * c.MX + rx) * c.MZ + z];
  if (conj) v.y = -v.y;
  return v;
}

__global__ __launch_bounds__(kW) void k_direct(const int* __restrict__ ma, int na,
                                               const int* __restrict__ mb, int nb, int lo_c,
                                               int hi_c, int N, int periodic, Cube ca, Cube cb,
                                               Cube cc, double* __restrict__ part) {
  __shared__ double sh[6][kW / 64];
  const long long total = (long long)na * nb;
  const long long stride = (long long)gridDim.x * kW;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long long e = (long long)blockIdx.x * kW + threadIdx.x; e < total; e += stride) {
    const int p = (int)(e / nb), q = (int)(e - (long long)p * nb);
    const int ax = ma[3 * p], ay = ma[3 * p + 1], az = ma[3 * p + 2];
    const int bx = mb[3 * q], by = mb[3 * q + 1], bz = mb[3 * q + 2];
    int x = -(ax + bx), y = -(ay + by), z = -(az + bz);
    if (periodic) {
      x = wrapc(x, N);
      y = wrapc(y, N);
      z = wrapc(z, N);
    } else {
      // k3 must be one of the collected modes: the stored half grid (x, y in
      // [-h, h), 0 <= z <= h) or the conjugate of a mode of Hermitian weight
      // 2 (0 < -z < h, so x, y in (-h, h])
      const int h = N / 2;
      const bool in = z >= 0 ? (z <= h && x >= -h && x < h && y >= -h && y < h)
                             : (z > -h && x > -h && x <= h && y > -h && y <= h);
      if (!in) continue;
    }
    const int n2 = x * x + y * y + z * z;
    if (n2 < lo_c || n2 > hi_c) continue;
    acc[0] += 1.0;
    acc[1] += sqrt((double)(ax * ax + ay * ay + az * az));
    acc[2] += sqrt((double)(bx * bx + by * by + bz * bz));
    acc[3] += sqrt((double)n2);
    if (ca.d) {
      const double2 va = lookup(ca, ax, ay, az), vb = lookup(cb, bx, by, bz);
      const double2 vc = lookup(cc, x, y, z);
      const double re = va.x * vb.x - va.y * vb.y, im = va.x * vb.y + va.y * vb.x;
      acc[4] += re * vc.x - im * vc.y;
      acc[5] += re * vc.y + im * vc.x;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    double v = acc[s];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[s][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = 0.0;
    for (int i = 0; i < kW / 64; ++i) v += sh[threadIdx.x][i];
    part[(long long)blockIdx.x * 6 + threadIdx.x] = v;
  }
}

}  // namespace

extern "C" int bsk_direct_triangles(int nmesh, int periodic, const int* modes_a, int na,
                                    const int* modes_b, int nb, int n2lo_c, int n2hi_c,
                                    const void* dk_a, const void* dk_b, const void* dk_c, int mx,
                                    int mz, int nlo, int nblocks, double* partial, void* stream) {
  if (nmesh <= 0 || (nmesh & 1) || na < 0 || nb < 0 || nblocks <= 0 || !partial ||
      (na && !modes_a) || (nb && !modes_b) || ((dk_a || dk_b || dk_c) && !(dk_a && dk_b && dk_c)) ||
      (dk_a && (mx <= 0 || mz <= 0))) {
    bsk_set_error("bsk_direct_triangles: invalid argument");
    return BSK_EINVAL;
  }
  Cube ca{(const double2*)dk_a, nmesh, mx, mz, nlo};
  Cube cb{(const double2*)dk_b, nmesh, mx, mz, nlo};
  Cube cc{(const double2*)dk_c, nmesh, mx, mz, nlo};
  hipLaunchKernelGGL(k_direct, dim3(nblocks), dim3(kW), 0, reinterpret_cast<hipStream_t>(stream),
                     modes_a, na, modes_b, nb, n2lo_c, n2hi_c, nmesh, periodic, ca, cb, cc,
                     partial);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
