/* libbskhip -- C ABI of the gfx950 FFT-bispectrum path.
 *
 * The reference (sjforeman/bskit) is a pure-Python library with no FFI: the
 * hot path is the FFTBispectrum class (bskit/main.py:1392-2161) calling
 * nbodykit/pmesh for its FFTs and numpy for its triple sums.  Each entry
 * point below replaces one of those calls; the citation on each names the
 * reference code it stands in for.  The Python layer (bskit_amd.main)
 * mirrors the reference's class and module functions on top of this ABI.
 *
 * Conventions
 *   - Every data pointer is caller-owned device memory; nothing is freed
 *     across the ABI.  A plan owns only small constant tables.
 *   - All launches are asynchronous on the given stream (hipStream_t passed
 *     as void*; NULL = default stream).  A plan must not be used by two host
 *     threads at once.
 *   - Return value: 0 on success, negative BSK_E* code on error; the message
 *     of the last error on the calling thread is bsk_last_error().
 *   - Grids are cubic: Nmesh^3 cells in a box of side box_size.  Supported
 *     Nmesh: the even 5-smooth sizes 16, 24, 32, 40, 48, 64, 72, 80, 96, 120,
 *     128, 144, 160, 192, 200, 216, 240, 256, 288, 320, 384, 400, 432, 480,
 *     512, 576, 640, 768, 1024.
 *   - Shell membership: bin b holds the modes with n2lo[b] <= |n|^2 <= n2hi[b]
 *     (n = k / k_f integer vector, components in [-N/2, N/2]); the integer
 *     limits are the closed interval k_low <= |k| <= k_high with modes
 *     within 1e-9 (relative) of an edge counted inside.
 */
#ifndef BSKHIP_H
#define BSKHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSK_ABI_VERSION 7

/* status codes */
#define BSK_OK 0
#define BSK_EINVAL (-1)
#define BSK_ENOMEM (-2)
#define BSK_EUNSUPPORTED (-3)
#define BSK_EHIP (-4)

/* element types of real fields / complex spectra */
#define BSK_DT_F32 0
#define BSK_DT_F64 1
/* shell storage formats */
#define BSK_FMT_F32 0
#define BSK_FMT_BF16X2 1 /* two bf16 planes: hi = rne(v), lo = rne(v - hi) */
#define BSK_FMT_F64 2
/* shell kinds */
#define BSK_KIND_DELTA 0 /* mask * delta(k)          (main.py:1846-1861)      */
#define BSK_KIND_UNIT 1  /* mask                     (number_field :280-329)  */
#define BSK_KIND_KMAG 2  /* |k| * mask               (k_field p=1 :227-277)   */
#define BSK_KIND_KPOW 3  /* |k|^p * mask, p from bsk_plan_set_kpower (k_field :264-268) */
/* triple-product modes (field roles of main.py:624-640) */
#define BSK_MODE_AUTO 0 /* <AAA>: triangles taken with i >= j >= k      */
#define BSK_MODE_AAB 1  /* <AAB>: symmetric in the first two sides      */
#define BSK_MODE_ABC 2  /* <ABC>: no symmetry                           */

typedef struct bsk_plan bsk_plan;

/* Message of the last failed call on this thread (replaces the reference's
 * Python exceptions at the ABI). */
const char* bsk_last_error(void);
int bsk_version(void);

/* Integer |n|^2 limits of the closed shells [k_low[b], k_high[b]].
 * Replaces the float mask predicate (kk<=kmax)&(kk>=kmin) of
 * bskit/main.py:264-268, :314-320, :1850-1852. */
int bsk_shell_thresholds(int nmesh, double box_size, int nbin, const double* k_low,
                         const double* k_high, int* n2lo, int* n2hi);

/* Plan for one (Nmesh, BoxSize, k bins) key.  Replaces the ParticleMesh set
 * up per shell in k_field / number_field (bskit/main.py:271, :323) and the
 * bin-edge list of measure_bispectrum_faster (:1828-1833). */
int bsk_plan_create(bsk_plan** plan, int nmesh, double box_size, int nbin, const double* k_low,
                    const double* k_high);
int bsk_plan_destroy(bsk_plan* plan);
/* Compact retained cube: |n_x|,|n_y| <= m (mx = my = min(2m+1, N) values,
 * first signed frequency nlo) and 0 <= n_z < mz. */
int bsk_plan_geometry(const bsk_plan* plan, int* m, int* mx, int* my, int* mz, int* nlo);
int bsk_plan_thresholds(const bsk_plan* plan, int* n2lo, int* n2hi);
/* Shell mesh S: the shell fields and triple sums are evaluated on S^3 cells.
 * Every shell mode has components |n_a| <= m, so a product of three shells
 * has components in [-3m, 3m] and on any mesh with S > 3m the cell sum is
 * S^3 times the sum over closed triangles q1+q2+q3 = 0, the same as on the
 * N^3 mesh of the reference up to (N/S)^3.  Scale sums by 1/S^3 where the
 * reference divides by N^3 (bskit/main.py:1875-1882, :2024-2064).  The
 * default is the smallest supported S > 3m, or N when that is not smaller;
 * set_shell_mesh(0) restores it, set_shell_mesh(N) gives the reference's own
 * mesh. */
int bsk_plan_shell_mesh(const bsk_plan* plan, int* shell_mesh);
int bsk_plan_set_shell_mesh(bsk_plan* plan, int shell_mesh);
/* Sublattice levels (S a multiple of 4).  The sum over a sublattice H of
 * the S^3 mesh is |H|/S^3 times the full sum whenever no nonzero vector of
 * H's dual lattice is a frequency of the product, i.e. for a triangle whose
 * shells' largest |n| components sum to R_inf and whose largest |n| sum to
 * R_2, when every nonzero dual vector v has a component above R_inf or
 * |v| > R_2.  The level layout (BSK_STAGE_LEVELS) orders a shell row as
 * four parts, P0 = C0 minus E, P1 = E, P2 = O, P3 = C1 minus O (C0 / C1:
 * cells of even / odd x+y+z; E / O: x, y, z all even / all odd), each
 * zero-padded to a multiple of 128 cells, a permutation of the mesh in
 * which every level is one contiguous range:
 *   0: the whole mesh P0..P3 (dual S Z^3);
 *   1: C0 = P0 P1, 1/2 of the cells (dual (S/2)(1,1,1) + S Z^3: |v| = 0.87 S);
 *   2: E and O = P1 P2, 1/4 (dual (S/2)(1,1,0) and permutations + S Z^3:
 *      |v| = 0.71 S);
 *   3: E = P1, 1/8 (dual (S/2) Z^3: a component S/2).
 * A cell (x, y, z) of a part sits at ((c S/2 + x/2) S/2 + y/2) S/2 + z/2,
 * c its row class: 0 in E and O, 2 (x%2) + y%2 - 1 in P0, 2 (x%2) + y%2 in P3.
 * bsk_plan_levels: offset[L] and ncell[L] of level L within a shell row
 * (BSK_LEVELS entries each) and the smallest shell stride of the layout.
 * bsk_plan_triangle_levels: level[t] = the cheapest exact level of each
 * triangle (bin indices).  No reference counterpart: the reference sums
 * every triangle over the whole N^3 mesh (bskit/main.py:1871-1882). */
#define BSK_LEVELS 4
int bsk_plan_levels(const bsk_plan* plan, long long* offset, long long* ncell,
                    long long* stride);
int bsk_plan_triangle_levels(const bsk_plan* plan, int ntri, const int* tri_host, int* level);
/* Radix-plan variant (0 or 1) of FFT pass 0..5 (z r2c, y, x, masked x^-1,
 * y^-1, z c2r); chosen by the autotuner. */
int bsk_plan_set_variant(bsk_plan* plan, int pass, int variant);
/* CIC window compensation applied in the forward transform: delta(k) is
 * divided by prod_i (1 - 2/3 sin^2(pi n_i / nmesh_window))^(1/2).  Replaces
 * CompensateCICShotnoiseNgrid (scripts/measure/measure_bs_fast.py:45-57,
 * applied :222) and nbodykit's compensated CIC mesh (:211-217).
 * nmesh_window <= 0 disables it.  The window is read when bsk_forward is
 * called, so fields with different windows (or none) can share one plan:
 * select the window before each forward call.  The table of every window
 * used stays allocated until bsk_plan_destroy, so launches in flight never
 * read a released table. */
int bsk_plan_set_compensation(bsk_plan* plan, int nmesh_window);
/* Exponent p of the BSK_KIND_KPOW shells (|k|^p on the shell, |k| in box
 * units of 2 pi / box_size per k_f), default 1.  Replaces the power
 * argument of k_field (bskit/main.py:227-277, the weight at :264-268).
 * A shell holding k = 0 with p < 0 gets an infinite mode, as numpy's
 * 0.0 ** p gives there. */
int bsk_plan_set_kpower(bsk_plan* plan, double power);

/* Forward transform of a real N^3 field (C order, z fastest) into the
 * compact cube delta(k) [my][mx][mz] (complex, interleaved), normalised by
 * 1/N^3.  Replaces mesh.paint(mode='complex') in
 * FFTBispectrum._paint_meshes (bskit/main.py:1608-1621).
 * work: bsk_forward_workspace() bytes. */
int bsk_forward_workspace(const bsk_plan* plan, int dtype, size_t* bytes);
int bsk_forward(bsk_plan* plan, int dtype, const void* field, void* dk, void* work, void* stream);

/* Shell fields of bins [shell0, shell0+nshell) on the x-slab
 * [x0, x0+xslab) of the shell mesh (S^3 cells, 0 <= x0 < S): unnormalised
 * inverse transform of the masked spectrum.
 * stages: BSK_STAGE_X runs the masked inverse x transform of every
 * retained (ky,kz) line into work; BSK_STAGE_YZ runs the y and z passes of
 * the slab from that work (so one X stage serves many slabs); both = 3.
 * BSK_STAGE_Y instead of BSK_STAGE_YZ runs the y pass only and leaves its
 * output in work for bsk_triples_zfused (FP32, x0 = 0, xslab = S; out
 * unused).  BSK_STAGE_LEVELS with BSK_STAGE_YZ over the whole mesh
 * (x0 = 0, xslab = S, out_cell0 = 0, S a multiple of 4) writes each shell
 * in the sublattice level layout (bsk_plan_levels) instead of C order;
 * shell_stride must hold the layout.
 * Output shell b (relative to shell0) lands at
 * out + (shell0+b)*shell_stride + out_cell0 (elements) with the slab's cells
 * in C order; for BSK_FMT_BF16X2 out is the hi plane and out_lo the lo
 * plane.  kind selects delta / unit / |k| shells (dk unused for the last
 * two).  Replaces the per-bin view + mask action + paint(mode='real') of
 * measure_bispectrum_faster (bskit/main.py:1846-1861) and number_field /
 * k_field (:227-329) as used by measure_gridinfo_faster (:2006-2013).
 * work: bsk_shells_workspace(nshell, xslab) bytes. */
#define BSK_STAGE_X 1
#define BSK_STAGE_YZ 2
#define BSK_STAGE_Y 4
#define BSK_STAGE_LEVELS 8
int bsk_shells_workspace(const bsk_plan* plan, int dtype, int nshell, int xslab, size_t* bytes);
int bsk_shells(bsk_plan* plan, int dtype, int kind, const void* dk, int shell0, int nshell, int x0,
               int xslab, int stages, int fmt, void* out, void* out_lo, long long shell_stride,
               long long out_cell0, void* work, void* stream);

/* Shells [shell0, shell0 + nshell) over the whole shell mesh in the level
 * layout (stages X, YZ and LEVELS of bsk_shells), writing of shell b only
 * the layout parts p with bit p of parts[b - shell0] set (device array of
 * nshell bytes; NULL: every part).  With the masks of bsk_pairs_shell_parts
 * for the triangle list summed next, the parts no launch of its plan reads
 * are not written.  No reference counterpart (see bsk_plan_levels). */
int bsk_shells_levels(bsk_plan* plan, int dtype, int kind, const void* dk, int shell0, int nshell,
                      int fmt, void* out, void* out_lo, long long shell_stride,
                      const unsigned char* parts, void* work, void* stream);

/* Binned power spectrum sums of the plan's bins from the compact delta(k)
 * cube (bsk_forward output): out_dev[3b + 0] = sum w |delta_k|^2,
 * out_dev[3b + 1] = sum w (modes), out_dev[3b + 2] = sum w |k| over the modes
 * of bin b, w counting both members of each conjugate pair.  By Parseval
 * these are the cell sums / N^3 of the squared shell fields that pk_FFT
 * forms (bskit/main.py:333-405): P = BoxSize^3 * out[3b] / out[3b+1],
 * Nbin = out[3b+1], k_mean = out[3b+2] / out[3b+1].  FP64, fixed order.
 * work: bsk_power_spectrum_workspace() bytes. */
int bsk_power_spectrum_workspace(const bsk_plan* plan, size_t* bytes);
int bsk_power_spectrum(bsk_plan* plan, int dtype, const void* dk, double* out_dev, void* work,
                       void* stream);

/* out[t] = scale * sum_x A[i_t](x) B[j_t](x) C[k_t](x) for tri_dev = (i,j,k)
 * per triangle (int32, device), one sweep over the cells per triangle in
 * the shells' own precision.  Replaces the per-triangle
 * np.sum(d_i*d_j*d_k) * L^6/N^3 + comm.gather/bcast of
 * measure_bispectrum_faster (bskit/main.py:1871-1882).
 * work: bsk_triples_plain_workspace(ntri, nchunk) bytes. */
int bsk_triples_plain_workspace(int ntri, int nchunk, size_t* bytes);
int bsk_triples_plain(int dtype, const void* fa, const void* fb, const void* fc,
                      long long shell_stride, long long ncell, int ntri, const int* tri_dev,
                      int nchunk, double scale, double* out_dev, void* work, void* stream);

/* LDS-tiled triple sums over fields[0..nfield) (stride shell_stride).
 * nsum = 1: out[t] = scale * sum_x F[i] F[j] F[k].
 * nsum = 4: fields are nfield/2 unit shells then nfield/2 |k| shells and
 *   out[4t+0] = scale * sum n_i n_j n_k, out[4t+1..3] = scale * sum of the
 *   products with n_i, n_j, n_k replaced by the |k| shell in turn.
 * Replaces the N_tri and k-mean sums of measure_gridinfo_faster
 * (bskit/main.py:2024-2064).  work: bsk_triples_tiled_workspace() bytes. */
int bsk_triples_tiled_workspace(int ntri, int nsum, int nchunk, size_t* bytes);
int bsk_triples_tiled(int dtype, const void* fields, long long shell_stride, int nfield,
                      long long ncell, int ntri, const int* tri_dev, int nsum, int nchunk,
                      double scale, double* out_dev, void* work, void* stream);

/* Split-BF16 MFMA triple sums over pair lists.  A pair plan (host tables
 * uploaded once) lists the distinct pairs (i, j) that close a triangle of
 * the list, packed into 16-row tiles; per tile and 16-shell block kb of the
 * third side the kernel computes the GEMM over the cells of the pair
 * products d_i d_j (A operand) and the shells d_k (B operand), both split
 * into bf16 hi + lo and multiplied as hi*hi + hi*lo + lo*hi on
 * v_mfma_f32_16x16x32_bf16, FP32 accumulation over short chains, FP64
 * second level and FP64 final reduction; out[t] = scale * sum for every
 * triangle in the order given.  Shells are BF16X2 planes (hi, lo).
 * Roles: BSK_MODE_AUTO takes all sides from (hi_j, lo_j) (any side order);
 * BSK_MODE_AAB takes sides i, j from (hi_j, lo_j) and k from (hi_k, lo_k);
 * BSK_MODE_ABC takes i, j, k from the three sets.  Up to 128 shells (96 for
 * <ABC>).  Replaces the per-triangle loop of measure_bispectrum_faster
 * (bskit/main.py:1871-1882) and, on unit shells, the N_tri sum
 * (:2027-2031).  work: bsk_triples_mfma_workspace(pairs, nwg) bytes. */
typedef struct bsk_pairs bsk_pairs;
/* Host-only planning (no device calls); bsk_pairs_upload copies the tables
 * to the current device once, before the first bsk_triples_mfma. */
int bsk_pairs_create(bsk_pairs** pairs, int mode, int nshell, int ntri, const int* tri_host);
int bsk_pairs_upload(bsk_pairs* pairs);
int bsk_pairs_destroy(bsk_pairs* pairs);
int bsk_pairs_info(const bsk_pairs* pairs, int* npair, int* ntile, int* nitem,
                   int* items_per_wave, int* npass);
/* Plan over sublattice levels: triangle t is summed over the cells of
 * level[t] (bsk_plan_triangle_levels) and scaled by 2^level[t]; run it with
 * bsk_triples_mfma_levels.  The launches run over segments of the level
 * layout: runs of consecutive parts (P0..P3) holding the same levels of the
 * plan, the segment holding E (every level) first; each segment has its own
 * pair tiles and items, planned over the triangles of all its levels at
 * once (a pair that triangles of several levels need is formed once), and
 * stores its own partials.  bsk_pairs_segments:
 * *nseg segments in launch order, per segment the parts (bit p), the shell
 * window its launches stage (shells [first, first + count)), its launches,
 * their items per wave and the replicas of their wave lists (a segment with
 * few items puts its lists on 8/replicas waves, repeated, each copy running
 * every replicas-th K step; the first copy adds the others' sums and
 * stores) (arrays of 4 entries).  A plan without levels (all level 0) is
 * one segment of all parts. */
int bsk_pairs_create_levels(bsk_pairs** pairs, int mode, int nshell, int ntri,
                            const int* tri_host, const int* level);
int bsk_pairs_segments(const bsk_pairs* pairs, int* nseg, int* parts, int* first, int* count,
                       int* npass, int* items_per_wave, int* replicas);
/* The layout parts (bit p) in which some launch of the plan stages shell
 * b: parts[b] for the nshell shells of the plan. */
int bsk_pairs_shell_parts(const bsk_pairs* pairs, unsigned char* parts);
/* Partial-buffer slots of each triangle, one per launch segment
 * (slot_host[4 t + g] for segment g: item * 256 + row * 16 + column, plus
 * level << 28; -1 where the triangle's level does not include the segment's
 * parts; the first segment holds every triangle): 4 ntri ints.  The sum of
 * triangle t is the sum of its slots times 2^level. */
int bsk_pairs_slots(const bsk_pairs* pairs, int* slot_host);
/* Host copies of the launch tables, for inspection: items and outpos
 * (per launch 8 waves x its segment's items per wave: descriptor tile | kb << 16 | padding flag
 * << 29 | new-tile flag << 31, kb counted from block first/16 of the
 * launch's segment; partial position of the item, -1 for padding) and pair
 * rows [(ntile + 1) * 16] (shell of set 0 | shell of set 1 << 16, 0xffff for
 * padding rows; the last tile is all padding). */
int bsk_pairs_tables(const bsk_pairs* pairs, int* items, int* outpos, int* pair_rows);
int bsk_triples_mfma_workspace(const bsk_pairs* pairs, int nwg, size_t* bytes);
int bsk_triples_mfma(const bsk_pairs* pairs, const uint16_t* hi_i, const uint16_t* lo_i,
                     const uint16_t* hi_j, const uint16_t* lo_j, const uint16_t* hi_k,
                     const uint16_t* lo_k, long long shell_stride, long long ncell, int nwg,
                     double scale, double* out_dev, void* work, void* stream);
/* bsk_triples_mfma over shells in the level layout: the passes of level L
 * run over cells [level_off[L], level_off[L] + level_ncell[L]) of each
 * shell row (host arrays of BSK_LEVELS entries, from bsk_plan_levels).
 * bsk_triples_mfma refuses a plan with triangles above level 0. */
int bsk_triples_mfma_levels(const bsk_pairs* pairs, const uint16_t* hi_i, const uint16_t* lo_i,
                            const uint16_t* hi_j, const uint16_t* lo_j, const uint16_t* hi_k,
                            const uint16_t* lo_k, long long shell_stride,
                            const long long* level_off, const long long* level_ncell, int nwg,
                            double scale, double* out_dev, void* work, void* stream);
/* bsk_triples_mfma_levels without a workspace: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-workgroup partials and range sums live in the library's own
 * scratch array of BSK_NOALLOC_SCRATCH_BYTES per device (held until the
 * process ends); a plan and workgroup count whose bsk_triples_mfma_workspace
 * exceeds it is refused with BSK_EINVAL.  The call allocates nothing; out_dev
 * is required (the result is reduced before returning to the stream).  All
 * calls on one device share its array: calls that overlap on one device (two
 * streams, or two host threads) race and must be serialised by the caller. */
#define BSK_NOALLOC_SCRATCH_BYTES (64u << 20)
int bsk_triples_mfma_levels_noalloc(const bsk_pairs* pairs, const uint16_t* hi_i,
                                    const uint16_t* lo_i, const uint16_t* hi_j,
                                    const uint16_t* lo_j, const uint16_t* hi_k,
                                    const uint16_t* lo_k, long long shell_stride,
                                    const long long* level_off, const long long* level_ncell,
                                    int nwg, double scale, double* out_dev, void* stream);
/* out_dev == NULL in bsk_triples_mfma leaves the per-workgroup partials in
 * work; bsk_triples_mfma_reduce then sums them in FP64 into out_dev (the
 * same result as one call with out_dev; it uses the tail of work, after the
 * partials, for its range sums). */
int bsk_triples_mfma_reduce(const bsk_pairs* pairs, int nwg, double scale, double* out_dev,
                            void* work, void* stream);
/* The same split-BF16 sums for an auto spectrum with the z c2r of the shells
 * done inside the triple-sum kernel: shell_work is the bsk_shells workspace
 * after stages BSK_STAGE_X | BSK_STAGE_Y of shells [shell0, shell0 +
 * nshell) of an FP32 spectrum (nshell = the pair plan's shell count, x0 = 0,
 * xslab = S), so the shell fields never reach HBM.  Replaces the shell
 * paint and the per-triangle sum of measure_bispectrum_faster
 * (bskit/main.py:1846-1882) like bsk_shells + bsk_triples_mfma.  work:
 * bsk_triples_mfma_workspace(pairs, nwg) bytes; out_dev == NULL leaves the
 * partials for bsk_triples_mfma_reduce.  *ok (supported query): 1 if this
 * plan (shell mesh size, shell count, items per wave, LDS) has a kernel. */
int bsk_triples_zfused_supported(const bsk_plan* plan, const bsk_pairs* pairs, int* ok);
int bsk_triples_zfused(bsk_plan* plan, const bsk_pairs* pairs, const void* shell_work,
                       int shell0, int nwg, double scale, double* out_dev, void* work,
                       void* stream);

/* Cloud-in-cell assignment of npart particles (pos: npart x 3 float, box
 * units, C order) to the periodic mesh grid (nmesh^3 float, overwritten).
 * Mesh node i sits at i*box/nmesh.  With to_contrast the result is
 * delta = rho/rho_mean - 1.  Every corner weight is rounded once to 32.32
 * fixed point and summed as integers, so the mesh is the same bits whatever
 * order the particles are processed in (deterministic).  Replaces nbodykit's
 * CatalogSource.to_mesh(window='cic') painting as called by
 * scripts/measure/measure_bs_fast.py:209-217 (the compensation is
 * bsk_plan_set_compensation).  work: bsk_cic_workspace() bytes; at most
 * 2^32 - 1 particles. */
int bsk_cic_workspace(int nmesh, long long npart, size_t* bytes);
int bsk_cic_paint(int nmesh, double box_size, long long npart, const float* pos, float* grid,
                  int to_contrast, void* work, void* stream);

/* Direct enumeration of closed triangles from explicit mode lists (int32
 * (nx,ny,nz) triples): for every k1 in A and k2 in B, k3 = -(k1+k2) is
 * accepted when n2lo_c <= |n3|^2 <= n2hi_c and, periodic: after wrapping
 * into [-N/2, N/2) (the lists then hold components in [-N/2, N/2)); open:
 * when k3 is a mode bk_binned collects, the stored half grid (n_x, n_y in
 * [-N/2, N/2), 0 <= n_z <= N/2) or the conjugate of a mode with
 * 0 < n_z < N/2 (the lists then hold such modes).  With dk_a/b/c
 * (FP64 compact cubes from bsk_forward, geometry mx, mz, nlo) the complex
 * sum of dA(k1) dB(k2) dC(k3) is formed too.  partial: nblocks x 6 doubles
 * (count, sum|k1|, sum|k2|, sum|k3| in k_f units, Re, Im) to be summed by
 * the caller.  Replaces the brute-force loops of bk_binned
 * (bskit/main.py:834-1004). */
int bsk_direct_triangles(int nmesh, int periodic, const int* modes_a, int na, const int* modes_b,
                         int nb, int n2lo_c, int n2hi_c, const void* dk_a, const void* dk_b,
                         const void* dk_c, int mx, int mz, int nlo, int nblocks, double* partial,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif
