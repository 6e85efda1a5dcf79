// Internal host-side launch API shared by the translation units.
#pragma once
#include "common.h"
#include "fft_passes.h"

// Mesh sizes with compiled FFT kernels: even 5-smooth lengths whose half
// also has a radix plan (the z passes run N/2-point complex transforms).
#define BSK_FFT_SIZES(X)                                                                  \
  X(16) X(24) X(32) X(40) X(48) X(64) X(72) X(80) X(96) X(120) X(128) X(144) X(160)       \
  X(192) X(200) X(216) X(240) X(256) X(288) X(320) X(384) X(400) X(432) X(480) X(512)    \
  X(576) X(640) X(768) X(1024)

bool bsk_fft_size_supported(int N);
// Smallest supported mesh size strictly greater than n (0 if none).
int bsk_fft_size_above(int n);

template <typename T>
int bsk_fwd_z(int N, int V, const T* in, cplx<T>* out, const cplx<T>* tw, long long nrows, int MZ,
              int MZP, hipStream_t s);
template <typename T>
int bsk_fwd_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g,
              hipStream_t s);
template <typename T>
int bsk_fwd_x(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, T scale,
              const T* comp, hipStream_t s);
// Pencils (kz values) per workgroup of the strided x / y passes: 16
// consecutive kz (128 B in FP32) per grid point; FP64 and lengths above 640
// use 8 to stay inside the LDS budget.
constexpr int bsk_pxy(bool f64, int n) { return (f64 || n > 640) ? 8 : 16; }

// The inverse x and y passes launch only their workgroups that hold modes
// of the shell: block lists (shell, ry << 16 | kz0, n2lo, n2hi) for x and
// (shell, kz0, n2hi, 0) for y, P-pencil blocks, shells in order; the launch
// takes the entries of its shells.
template <typename T>
int bsk_inv_x(int N, int V, const cplx<T>* dk, cplx<T>* out, const cplx<T>* tw, Geom g,
              int shell0, const int4* blocks, int nblocks, int kind, T kf, T kp, hipStream_t s);
template <typename T>
int bsk_inv_y(int N, int V, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, Geom g, int x0,
              int xs, const int4* blocks, int nblocks, int shell0, hipStream_t s);
template <typename T>
int bsk_inv_z(int N, int V, const cplx<T>* in, void* out, void* out_lo, const cplx<T>* tw, int MZ,
              int MZP, long long nrows, long long cell0, long long stride, int fmt, int nsh,
              const int* n2hi, int shell0, int lv, const unsigned char* parts,
              hipStream_t s);

// Triple products ------------------------------------------------------------
// Reference-structured FP32/FP64 path: one sweep over all cells per triangle.
template <typename T>
int bsk_triple_plain_launch(const T* a, const T* b, const T* c, long long stride_a,
                            long long stride_b, long long stride_c, long long ncell, int ntri,
                            const int* tri_dev, double* partial, int nchunk, hipStream_t s);
int bsk_reduce_plain_launch(const double* partial, int ntri, int nchunk, double scale, double* out,
                            hipStream_t s);

// Tiled VALU path: every triangle of a batch per pass over a cell chunk.
// nsum = 1 (B) or 4 (N_tri and the three k-mean numerators; fields are the
// unit shells followed by the |k| shells).
template <typename T>
int bsk_triple_tiled_launch(const T* f, long long stride, int nfield, long long ncell, int ntri,
                            const int* tri_dev, int nsum, double* partial, int nchunk,
                            hipStream_t s);
int bsk_reduce_tiled_launch(const double* partial, int ntri, int nsum, int nchunk, double scale,
                            double* out, hipStream_t s);

// Split-BF16 MFMA path over pair lists.
constexpr int kPairWaves = 8;    // waves per workgroup of k_triple_pairs
constexpr int kPairChunk = 128;  // cells per staged chunk
constexpr int kPairMaxItems = 8; // items per wave and launch
constexpr int kPairSlots = 4;    // partial slots per triangle: one per launch segment
struct PairArgs {
  const uint16_t* hi[3];  // role sets: 0 = first pair side, 1 = second, 2 = third side
  const uint16_t* lo[3];
  long long stride;       // elements between consecutive shells of a set
  int nsh[3];             // shells per role set
  int nkb;                // 16-shell blocks of the third side staged in LDS
  int lo_row;             // shell of staged row 0 (pair rows in the table are absolute)
  int koff;               // third-side LDS row of staged shell 0 (rows below are zero)
  long long ncell;        // cells per shell (multiple of kPairChunk)
  int nwg;                // workgroups (cell ranges)
  const int* items;       // [kPairWaves * NI] item descriptors of this launch
  const int* outpos;      // [kPairWaves * NI] partial position of each item (< 0: padding)
  int accumulate;         // add to the partials an earlier launch stored
  int rep;                // replicas of the wave lists (1, 2, 4), splitting the K steps
  const int* pairs;       // [tiles + 1][16]: (row in set 0) | (row in set 1) << 16
  float* partial;         // [nwg][item_total][256]
  int item_total;         // items per workgroup in the partial layout
  // item-list passes run by one launch (1: one list); with npass > 1 the
  // launch has npass * nwg workgroups in groups of 8 * npass, group g's
  // workgroup 8 p + x running pass p (list at items + p * pass_stride) over
  // cell range 8 g + x: the workgroups sharing a cell range are on one XCD
  // (workgroups go round-robin over the 8 XCDs), so the later passes read
  // the staged shells from that XCD's L2
  int npass;
  int pass_stride;
};
int bsk_pair_items_round(int ni);  // compiled items-per-wave count >= ni (0: too many)
int bsk_pair_item_cap(int sets, int nshell);  // items per wave that compile without spills
int bsk_pair_prefetch(int nshell);  // prefetch segments per thread (0: too many shells)
size_t bsk_pair_lds_bytes(int sets, const int nsh[3], int nkb);  // nkb covers koff + nsh[2]
int bsk_triple_pairs_launch(const PairArgs& a, int sets, int ni, hipStream_t s);
size_t bsk_reduce_mfma_bytes(int items_per_wg);  // scratch of the reduction (FP64 range sums)
int bsk_reduce_mfma_launch(const float* partial, int nwg, int items_per_wg, int ntri,
                           const int* slot_dev, double scale, double* out, double* sums,
                           hipStream_t s);

// Fused z c2r + split-BF16 triple sums (zfused.hip): shell meshes with a
// compiled kernel, at most kZMaxShells shells.
#define BSK_ZFUSED_SIZES(X) \
  X(64) X(72) X(80) X(96) X(120) X(128) X(144) X(160) X(192) X(200) X(216) X(240) X(256)
constexpr int kZMaxShells = 40;
struct ZArgs {
  const cplx<float>* d2;  // inverse y output [nsh][nrows][MZP] (whole x range)
  long long nrows;        // S * S (x, y) rows
  int MZ, MZP;
  const int* n2hi;        // shell limits (device), from shell0
  int shell0;
  int nsh, nkb;
  int lo_row;             // shell of row 0 (pair rows in the table are absolute)
  int koff;               // third-side LDS row of row 0
  int nwg;
  const int* items;       // [kPairWaves * NI] item descriptors of this launch
  const int* outpos;      // [kPairWaves * NI] partial position of each item (< 0: padding)
  const int* pairs;       // [tiles + 1][16]
  float* partial;         // [nwg][item_total][256]
  int item_total;
  const cplx<float>* tw;  // S twiddles exp(-2 pi i t / S)
};
size_t bsk_zrows_lds_bytes(int S, int nsh, int nkb);
int bsk_triple_zrows_launch_v0(const ZArgs& a, int S, int ni, hipStream_t s);
int bsk_triple_zrows_launch_v1(const ZArgs& a, int S, int ni, hipStream_t s);

// Power spectrum sums (spectrum.hip).
int bsk_spectrum_launch(int dtype, const void* dk, Geom g, int N, const int* n2lo,
                        const int* n2hi, int nbin, double kf, int nchunk, double* partial,
                        double* out, hipStream_t s);
