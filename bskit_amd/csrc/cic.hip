// Cloud-in-cell assignment of particles to a periodic N^3 mesh.
//
// Mesh node (i,j,k) sits at position (i,j,k) * L/N; a particle at mesh
// coordinate u = x N / L adds (1 - f) to node floor(u) and f to node
// floor(u)+1 along each axis (f = u - floor(u)), indices wrapped mod N.
// With to_contrast set the mesh is turned into delta = rho / rho_mean - 1.
// Stands in for nbodykit's CatalogSource.to_mesh(window='cic') painting
// used by scripts/measure/measure_bs_fast.py:209-217; the compensation
// half of that call lives in the forward transform (bsk_plan_set_compensation).
#include "common.h"

namespace {

__global__ __launch_bounds__(256) void k_fill(float* g, long long n, float v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) g[i] = v;
}

BSK_DEV int wrapi(int i, int N) {
  i %= N;
  return i < 0 ? i + N : i;
}

__global__ __launch_bounds__(256) void k_cic(const float* __restrict__ pos, long long np, int N,
                                             float to_mesh, float* __restrict__ g) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int i0[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = pos[3 * p + a] * to_mesh;
    const float fl = floorf(u);
    f[a] = u - fl;
    i0[a] = wrapi((int)fl, N);
  }
  const int i1[3] = {i0[0] + 1 == N ? 0 : i0[0] + 1, i0[1] + 1 == N ? 0 : i0[1] + 1,
                     i0[2] + 1 == N ? 0 : i0[2] + 1};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ix = (c & 4) ? i1[0] : i0[0];
    const int iy = (c & 2) ? i1[1] : i0[1];
    const int iz = (c & 1) ? i1[2] : i0[2];
    const float w = ((c & 4) ? f[0] : 1.f - f[0]) * ((c & 2) ? f[1] : 1.f - f[1]) *
                    ((c & 1) ? f[2] : 1.f - f[2]);
    unsafeAtomicAdd(g + ((long long)ix * N + iy) * N + iz, w);
  }
}

__global__ __launch_bounds__(256) void k_contrast(float* g, long long n, float scale) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    g[i] = g[i] * scale - 1.f;
}

}  // namespace

extern "C" int bsk_cic_paint(int nmesh, double box_size, long long npart, const float* pos,
                             float* grid, int to_contrast, void* stream) {
  if (nmesh <= 0 || box_size <= 0 || npart < 0 || (npart && !pos) || !grid) {
    bsk_set_error("bsk_cic_paint: invalid argument");
    return BSK_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long long ncell = (long long)nmesh * nmesh * nmesh;
  hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, s, grid, ncell, 0.f);
  BSK_LAUNCH_CHECK();
  if (npart > 0) {
    hipLaunchKernelGGL(k_cic, dim3((unsigned)((npart + 255) / 256)), dim3(256), 0, s, pos, npart,
                       nmesh, (float)(nmesh / box_size), grid);
    BSK_LAUNCH_CHECK();
  }
  if (to_contrast) {
    if (npart == 0) {
      bsk_set_error("bsk_cic_paint: density contrast of an empty catalogue");
      return BSK_EINVAL;
    }
    hipLaunchKernelGGL(k_contrast, dim3(4096), dim3(256), 0, s, grid, ncell,
                       (float)((double)ncell / (double)npart));
    BSK_LAUNCH_CHECK();
  }
  return BSK_OK;
}
