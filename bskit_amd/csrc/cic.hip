// Cloud-in-cell assignment of particles to a periodic N^3 mesh.
//
// Mesh node (i,j,k) sits at position (i,j,k) * L/N; a particle at mesh
// coordinate u = x N / L adds (1 - f) to node floor(u) and f to node
// floor(u)+1 along each axis (f = u - floor(u)), indices wrapped mod N.
// With to_contrast set the mesh is turned into delta = rho / rho_mean - 1.
// Stands in for nbodykit's CatalogSource.to_mesh(window='cic') painting
// used by scripts/measure/measure_bs_fast.py:209-217; the compensation
// half of that call lives in the forward transform (bsk_plan_set_compensation).
//
// Deterministic: every corner weight is rounded once to a 32.32 fixed-point
// integer and the per-node sums are integer additions, which do not depend
// on the order the particles arrive in (concurrent jobs, scheduling), so a
// painted mesh is the same bits in every run.
//
// Tiled path (N a multiple of 16): the mesh is cut into T^3 tiles (T = 16).
//  1. k_cic_bin: each particle's base node floor(u) (wrapped) and its tile;
//     the particles of a wave that share a tile take consecutive slots inside
//     it from one atomic add on the tile's count.
//  2. k_cic_scan: exclusive scan of the tile counts (one workgroup).
//  3. k_cic_place: the particle's record (base node within the tile, three
//     fractions) goes to offset[tile] + slot: particles grouped by tile.
//  4. k_cic_tile: one workgroup per tile adds its particles' 8 corner weights
//     into a (T+1)^3 fixed-point region in LDS.  Nodes with every local
//     coordinate in [1, T-1] receive weight only from this tile, so they are
//     final and written as float; the nodes with a local coordinate 0 or T
//     (shared with the neighbour tiles) go to a per-tile boundary record.
//  5. k_cic_faces: each node of a tile's low faces (a local coordinate 0)
//     sums its boundary entries from the tile and from the neighbours below
//     (where it is their local coordinate T) and is written as float.
// Other N: fixed-point 64-bit atomics into an N^3 integer workspace, then a
// conversion pass (same integers, so the same bits as the tiled path).
#include <algorithm>

#include "common.h"

namespace {

constexpr int kT = 16;                                   // tile edge (nodes)
constexpr int kR = kT + 1;                               // LDS region edge
constexpr int kRegion = kR * kR * kR;                    // 4913 nodes
constexpr int kBnd = kRegion - (kT - 1) * (kT - 1) * (kT - 1);  // 1538 boundary nodes
constexpr int kFace = kT * kT * kT - (kT - 1) * (kT - 1) * (kT - 1);  // 721 low-face nodes
constexpr double kFix = 4294967296.0;                    // 2^32: one particle

struct CicRec {
  uint32_t local;  // base node within its tile: lx | ly << 8 | lz << 16
  float f[3];
};

BSK_DEV long long wrapl(long long i, long long N) {
  i %= N;
  return i < 0 ? i + N : i;
}

// base node (wrapped) and fractions of particle p
BSK_DEV void particle(const float* __restrict__ pos, long long p, int N, float to_mesh, int (&b)[3],
                      float (&f)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = pos[3 * p + a] * to_mesh;
    const float fl = floorf(u);
    f[a] = u - fl;
    b[a] = (int)wrapl((long long)fl, N);
  }
}

// fixed-point weight of corner c (bit 2: x+1, bit 1: y+1, bit 0: z+1)
BSK_DEV unsigned long long corner_weight(const float (&f)[3], int c) {
  const float w = ((c & 4) ? f[0] : 1.f - f[0]) * ((c & 2) ? f[1] : 1.f - f[1]) *
                  ((c & 1) ? f[2] : 1.f - f[2]);
  return (unsigned long long)llrint((double)w * kFix);
}

// Position of region node (x, y, z) (each in [0, T]) with some coordinate 0
// or T in the tile's boundary record: rows (x, y) with x or y on the region
// edge hold all T+1 nodes, the others only z = 0 and z = T.
BSK_DEV int bnd_index(int x, int y, int z) {
  // rows before row x: x = 0 is edge (all R rows full); inner x: 2 full rows
  // (y = 0, y = T) and T-1 rows of two
  constexpr int kEdgePlane = kR * kR;
  constexpr int kInnerPlane = 2 * kR + (kT - 1) * 2;
  int off;
  if (x == 0)
    off = 0;
  else
    off = kEdgePlane + (x - 1) * kInnerPlane;  // x = T lands after all inner planes
  if (x == 0 || x == kT) return off + y * kR + z;
  // inner plane: row y = 0 full, rows 1..T-1 two entries, row T full
  if (y == 0) return off + z;
  if (y == kT) return off + kR + (kT - 1) * 2 + z;
  return off + kR + (y - 1) * 2 + (z == 0 ? 0 : 1);
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* g, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) g[i] = 0;
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* g, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) g[i] = 0;
}

// Slots are taken per wave: the lanes of a wave holding particles of one tile
// (neighbouring particles of a catalogue usually share tiles) get consecutive
// slots from a single atomic add of their count, so a tile's counter sees
// one atomic per wave instead of one per particle.
__global__ __launch_bounds__(256) void k_cic_bin(const float* __restrict__ pos, long long np, int N,
                                                 float to_mesh, uint32_t* __restrict__ count,
                                                 uint32_t* __restrict__ slot) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = p < np;
  int tile = -1;
  if (valid) {
    int b[3];
    float f[3];
    particle(pos, p, N, to_mesh, b, f);
    const int nt = N / kT;
    tile = ((b[0] / kT) * nt + b[1] / kT) * nt + b[2] / kT;
  }
  const int lane = __lane_id();
  unsigned long long todo = __ballot(valid);  // wave-uniform
  uint32_t mine = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int t = __shfl(tile, leader);
    const unsigned long long same = __ballot(valid && tile == t) & todo;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count + t, (uint32_t)__popcll(same));
    base = __shfl(base, leader);
    if ((same >> lane) & 1ull) mine = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  if (valid) slot[p] = mine;
}

// exclusive scan of n counts into off[0..n] (one workgroup of 1024 threads)
__global__ __launch_bounds__(1024) void k_cic_scan(const uint32_t* __restrict__ count, long long n,
                                                   uint32_t* __restrict__ off) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const long long per = (n + 1023) / 1024;
  const long long a = t * per, e = a + per < n ? a + per : n;
  uint32_t s = 0;
  for (long long i = a; i < e; ++i) s += count[i];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
    const uint32_t v = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (long long i = a; i < e; ++i) {
    off[i] = run;
    run += count[i];
  }
  if (t == 1023) off[n] = part[1023];
}

__global__ __launch_bounds__(256) void k_cic_place(const float* __restrict__ pos, long long np,
                                                   int N, float to_mesh,
                                                   const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ slot,
                                                   CicRec* __restrict__ rec) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int b[3];
  float f[3];
  particle(pos, p, N, to_mesh, b, f);
  const int nt = N / kT;
  const int tile = ((b[0] / kT) * nt + b[1] / kT) * nt + b[2] / kT;
  CicRec r;
  r.local = (uint32_t)(b[0] % kT) | ((uint32_t)(b[1] % kT) << 8) | ((uint32_t)(b[2] % kT) << 16);
  r.f[0] = f[0];
  r.f[1] = f[1];
  r.f[2] = f[2];
  rec[off[tile] + slot[p]] = r;
}

// node value from its fixed-point weight sum
BSK_DEV float node_value(unsigned long long v, double scale, float shift) {
  return (float)((double)v * scale - (double)shift);
}

__global__ __launch_bounds__(256) void k_cic_tile(const CicRec* __restrict__ rec,
                                                  const uint32_t* __restrict__ off, int N,
                                                  double scale, float shift,
                                                  float* __restrict__ grid,
                                                  unsigned long long* __restrict__ bnd) {
  __shared__ unsigned long long acc[kRegion];
  const int tile = blockIdx.x;
  const int nt = N / kT;
  const int tx = tile / (nt * nt), ty = (tile / nt) % nt, tz = tile % nt;
  for (int i = threadIdx.x; i < kRegion; i += blockDim.x) acc[i] = 0ull;
  __syncthreads();
  const uint32_t p0 = off[tile], p1 = off[tile + 1];
  for (uint32_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const CicRec r = rec[p];
    const int lx = r.local & 0xff, ly = (r.local >> 8) & 0xff, lz = (r.local >> 16) & 0xff;
    const float f[3] = {r.f[0], r.f[1], r.f[2]};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = ((lx + ((c >> 2) & 1)) * kR + ly + ((c >> 1) & 1)) * kR + lz + (c & 1);
      atomicAdd(acc + i, corner_weight(f, c));
    }
  }
  __syncthreads();
  unsigned long long* tb = bnd + (long long)tile * kBnd;
  for (int i = threadIdx.x; i < kRegion; i += blockDim.x) {
    const int x = i / (kR * kR), y = (i / kR) % kR, z = i % kR;
    const unsigned long long v = acc[i];
    if (x == 0 || x == kT || y == 0 || y == kT || z == 0 || z == kT) {
      tb[bnd_index(x, y, z)] = v;
    } else {
      const long long g = ((long long)(tx * kT + x) * N + (ty * kT + y)) * N + (tz * kT + z);
      grid[g] = node_value(v, scale, shift);
    }
  }
}

__global__ __launch_bounds__(256) void k_cic_faces(const unsigned long long* __restrict__ bnd,
                                                   int N, double scale, float shift,
                                                   float* __restrict__ grid) {
  const int nt = N / kT;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)nt * nt * nt * kFace) return;
  const int tile = (int)(q / kFace);
  int r = (int)(q - (long long)tile * kFace);
  // low-face nodes: plane x = 0 (T^2), then y = 0 with x >= 1 (T (T-1)), then
  // z = 0 with x, y >= 1 ((T-1)^2)
  int x, y, z;
  if (r < kT * kT) {
    x = 0;
    y = r / kT;
    z = r % kT;
  } else if ((r -= kT * kT) < kT * (kT - 1)) {
    x = 1 + r / kT;
    y = 0;
    z = r % kT;
  } else {
    r -= kT * (kT - 1);
    x = 1 + r / (kT - 1);
    y = 1 + r % (kT - 1);
    z = 0;
  }
  const int tx = tile / (nt * nt), ty = (tile / nt) % nt, tz = tile % nt;
  unsigned long long v = 0;
  // the node is (x, y, z) of this tile and, along each axis where its local
  // coordinate is 0, also coordinate T of the tile below (periodic)
  for (int d = 0; d < 8; ++d) {
    const int dx = (d >> 2) & 1, dy = (d >> 1) & 1, dz = d & 1;
    if ((dx && x) || (dy && y) || (dz && z)) continue;
    const int sx = dx ? (tx + nt - 1) % nt : tx;
    const int sy = dy ? (ty + nt - 1) % nt : ty;
    const int sz = dz ? (tz + nt - 1) % nt : tz;
    const int st = (sx * nt + sy) * nt + sz;
    v += bnd[(long long)st * kBnd + bnd_index(x + dx * kT, y + dy * kT, z + dz * kT)];
  }
  const long long g = ((long long)(tx * kT + x) * N + (ty * kT + y)) * N + (tz * kT + z);
  grid[g] = node_value(v, scale, shift);
}

// fallback (N not a multiple of the tile edge): global fixed-point atomics
__global__ __launch_bounds__(256) void k_cic_atomic(const float* __restrict__ pos, long long np,
                                                    int N, float to_mesh,
                                                    unsigned long long* __restrict__ acc) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int b[3];
  float f[3];
  particle(pos, p, N, to_mesh, b, f);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const long long ix = (c & 4) ? (b[0] + 1 == N ? 0 : b[0] + 1) : b[0];
    const long long iy = (c & 2) ? (b[1] + 1 == N ? 0 : b[1] + 1) : b[1];
    const long long iz = (c & 1) ? (b[2] + 1 == N ? 0 : b[2] + 1) : b[2];
    atomicAdd(acc + (ix * N + iy) * N + iz, corner_weight(f, c));
  }
}

__global__ __launch_bounds__(256) void k_cic_convert(const unsigned long long* __restrict__ acc,
                                                     long long n, double scale, float shift,
                                                     float* __restrict__ grid) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    grid[i] = node_value(acc[i], scale, shift);
}

bool tiled(int nmesh) { return nmesh % kT == 0; }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// workspace layout of the tiled path: count, offsets, slots, records, boundary
struct TiledWs {
  size_t count, off, slot, rec, bnd, total;
};

TiledWs tiled_ws(int nmesh, long long npart) {
  const long long nt = nmesh / kT;
  const long long ntile = nt * nt * nt;
  TiledWs w;
  w.count = 0;
  w.off = align256(ntile * 4);
  w.slot = w.off + align256((ntile + 1) * 4);
  w.rec = w.slot + align256((size_t)npart * 4);
  w.bnd = w.rec + align256((size_t)npart * sizeof(CicRec));
  w.total = w.bnd + align256((size_t)ntile * kBnd * 8);
  return w;
}

}  // namespace

extern "C" int bsk_cic_workspace(int nmesh, long long npart, size_t* bytes) {
  if (nmesh <= 0 || npart < 0 || !bytes) {
    bsk_set_error("bsk_cic_workspace: invalid argument");
    return BSK_EINVAL;
  }
  if (npart >= (1ll << 32)) {
    bsk_set_error("bsk_cic_workspace: at most 2^32 - 1 particles");
    return BSK_EINVAL;
  }
  if (tiled(nmesh))
    *bytes = tiled_ws(nmesh, npart).total;
  else
    *bytes = (size_t)nmesh * nmesh * nmesh * 8;
  return BSK_OK;
}

extern "C" int bsk_cic_paint(int nmesh, double box_size, long long npart, const float* pos,
                             float* grid, int to_contrast, void* work, void* stream) {
  if (nmesh <= 0 || box_size <= 0 || npart < 0 || (npart && !pos) || !grid || !work) {
    bsk_set_error("bsk_cic_paint: invalid argument");
    return BSK_EINVAL;
  }
  if (npart >= (1ll << 32)) {
    bsk_set_error("bsk_cic_paint: at most 2^32 - 1 particles");
    return BSK_EINVAL;
  }
  if (to_contrast && npart == 0) {
    bsk_set_error("bsk_cic_paint: density contrast of an empty catalogue");
    return BSK_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long long ncell = (long long)nmesh * nmesh * nmesh;
  const float to_mesh = (float)(nmesh / box_size);
  // node value = sum / 2^32 (density in particles), times ncell/npart minus 1
  // for the contrast
  const double scale = (to_contrast ? (double)ncell / (double)npart : 1.0) / kFix;
  const float shift = to_contrast ? 1.f : 0.f;
  const unsigned pblocks = (unsigned)((npart + 255) / 256);
  char* wb = static_cast<char*>(work);
  if (!tiled(nmesh)) {
    auto* acc = reinterpret_cast<unsigned long long*>(wb);
    hipLaunchKernelGGL(k_fill_u64, dim3(4096), dim3(256), 0, s, acc, ncell);
    BSK_LAUNCH_CHECK();
    if (npart > 0) {
      hipLaunchKernelGGL(k_cic_atomic, dim3(pblocks), dim3(256), 0, s, pos, npart, nmesh, to_mesh,
                         acc);
      BSK_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_cic_convert, dim3(4096), dim3(256), 0, s, acc, ncell, scale, shift, grid);
    BSK_LAUNCH_CHECK();
    return BSK_OK;
  }
  const TiledWs w = tiled_ws(nmesh, npart);
  const long long nt = nmesh / kT;
  const long long ntile = nt * nt * nt;
  auto* count = reinterpret_cast<uint32_t*>(wb + w.count);
  auto* off = reinterpret_cast<uint32_t*>(wb + w.off);
  auto* slot = reinterpret_cast<uint32_t*>(wb + w.slot);
  auto* rec = reinterpret_cast<CicRec*>(wb + w.rec);
  auto* bnd = reinterpret_cast<unsigned long long*>(wb + w.bnd);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<long long>(4096, (ntile + 255) / 256)),
                     dim3(256), 0, s, count, ntile);
  BSK_LAUNCH_CHECK();
  if (npart > 0) {
    hipLaunchKernelGGL(k_cic_bin, dim3(pblocks), dim3(256), 0, s, pos, npart, nmesh, to_mesh, count,
                       slot);
    BSK_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_cic_scan, dim3(1), dim3(1024), 0, s, count, ntile, off);
  BSK_LAUNCH_CHECK();
  if (npart > 0) {
    hipLaunchKernelGGL(k_cic_place, dim3(pblocks), dim3(256), 0, s, pos, npart, nmesh, to_mesh,
                       off, slot, rec);
    BSK_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_cic_tile, dim3((unsigned)ntile), dim3(256), 0, s, rec, off, nmesh, scale,
                     shift, grid, bnd);
  BSK_LAUNCH_CHECK();
  const long long nface = ntile * kFace;
  hipLaunchKernelGGL(k_cic_faces, dim3((unsigned)((nface + 255) / 256)), dim3(256), 0, s, bnd,
                     nmesh, scale, shift, grid);
  BSK_LAUNCH_CHECK();
  return BSK_OK;
}
