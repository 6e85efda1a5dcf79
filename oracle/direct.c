/* Oracle: exact periodic enumeration of closed triangles on an N^3 Fourier
 * grid.  TEST INFRASTRUCTURE ONLY -- never linked into the product.
 *
 * This is the periodic form of the reference's brute-force estimator
 * bk_binned (bskit/main.py:834-1004): instead of the float closure test
 * |k1+k2+k3|^2 < 1e-4 k_f (main.py:974-979) it uses the integer closure
 * k1+k2+k3 == 0 (mod N), which is what the FFT estimator
 * (main.py:1871-1882, :2024-2064) measures.  Shell membership uses the
 * integer thresholds n2lo <= |n|^2 <= n2hi, |n|^2 formed from frequency
 * components mapped into [-N/2, N/2).
 *
 * For each requested triangle (a,b,c) of shell indices it returns
 *   count[t]      = #{(k1,k2,k3): k1 in S_a, k2 in S_b, k3 in S_c, k1+k2+k3 = 0 mod N}
 *   ksum[3t+s]    = sum over those triples of |n_s| (k_f units)
 *   bsum[2t+{0,1}]= sum over those triples of dA(k1) dB(k2) dC(k3) (re, im)
 * exactly as integers (count) and in FP64 (the sums).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n;          /* member count */
    int *ix, *iy, *iz; /* signed components in [-N/2, N/2) */
    double *kmag;   /* sqrt(|n|^2) */
} shell_t;

static int wrap(int v, int N) {
    int h = N / 2;
    int r = (v + h) % N;
    if (r < 0) r += N;
    return r - h;
}

static int build_shell(int N, long long lo, long long hi, shell_t *s) {
    int h = N / 2;
    int m = (int)floor(sqrt((double)(hi > 0 ? hi : 0))) + 1;
    int a0 = -m, a1 = m;
    if (a0 < -h) a0 = -h;
    if (a1 > h - 1) a1 = h - 1;
    int cap = 0;
    for (int x = a0; x <= a1; x++)
        for (int y = a0; y <= a1; y++)
            for (int z = a0; z <= a1; z++) {
                long long q = (long long)x * x + (long long)y * y + (long long)z * z;
                if (q >= lo && q <= hi) cap++;
            }
    s->n = cap;
    s->ix = (int *)malloc(sizeof(int) * (cap + 1));
    s->iy = (int *)malloc(sizeof(int) * (cap + 1));
    s->iz = (int *)malloc(sizeof(int) * (cap + 1));
    s->kmag = (double *)malloc(sizeof(double) * (cap + 1));
    if (!s->ix || !s->iy || !s->iz || !s->kmag) return -1;
    int c = 0;
    for (int x = a0; x <= a1; x++)
        for (int y = a0; y <= a1; y++)
            for (int z = a0; z <= a1; z++) {
                long long q = (long long)x * x + (long long)y * y + (long long)z * z;
                if (q >= lo && q <= hi) {
                    s->ix[c] = x; s->iy[c] = y; s->iz[c] = z;
                    s->kmag[c] = sqrt((double)q);
                    c++;
                }
            }
    return 0;
}

static void free_shell(shell_t *s) {
    free(s->ix); free(s->iy); free(s->iz); free(s->kmag);
}

static size_t gidx(int x, int y, int z, int N) {
    int ax = x < 0 ? x + N : x, ay = y < 0 ? y + N : y, az = z < 0 ? z + N : z;
    return ((size_t)ax * N + ay) * N + az;
}

/* dA/dB/dC: full complex grids (re, im) in FFT index order [ix][iy][iz],
 * or NULL to skip the bispectrum sums. */
int oracle_direct(int N, int nbin, const long long *n2lo, const long long *n2hi,
                  int ntri, const int *tri,
                  const double *aRe, const double *aIm,
                  const double *bRe, const double *bIm,
                  const double *cRe, const double *cIm,
                  long long *count, double *ksum, double *bsum) {
    if (N <= 0 || (N & 1) || nbin <= 0 || ntri < 0) return -1;
    shell_t *sh = (shell_t *)calloc(nbin, sizeof(shell_t));
    if (!sh) return -2;
    for (int b = 0; b < nbin; b++)
        if (build_shell(N, n2lo[b], n2hi[b], &sh[b])) return -2;
    memset(count, 0, sizeof(long long) * ntri);
    memset(ksum, 0, sizeof(double) * 3 * ntri);
    if (bsum) memset(bsum, 0, sizeof(double) * 2 * ntri);
    for (int t = 0; t < ntri; t++) {
        int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (a < 0 || b < 0 || c < 0 || a >= nbin || b >= nbin || c >= nbin) return -1;
        const shell_t *A = &sh[a], *B = &sh[b];
        long long lo = n2lo[c], hi = n2hi[c];
        long long cnt = 0;
        double k1 = 0, k2 = 0, k3 = 0, br = 0, bi = 0;
#pragma omp parallel for reduction(+ : cnt, k1, k2, k3, br, bi) schedule(dynamic, 16)
        for (int p = 0; p < A->n; p++) {
            for (int q = 0; q < B->n; q++) {
                int x = wrap(-(A->ix[p] + B->ix[q]), N);
                int y = wrap(-(A->iy[p] + B->iy[q]), N);
                int z = wrap(-(A->iz[p] + B->iz[q]), N);
                long long n2 = (long long)x * x + (long long)y * y + (long long)z * z;
                if (n2 < lo || n2 > hi) continue;
                cnt++;
                k1 += A->kmag[p];
                k2 += B->kmag[q];
                k3 += sqrt((double)n2);
                if (aRe) {
                    size_t i1 = gidx(A->ix[p], A->iy[p], A->iz[p], N);
                    size_t i2 = gidx(B->ix[q], B->iy[q], B->iz[q], N);
                    size_t i3 = gidx(x, y, z, N);
                    double r1 = aRe[i1], m1 = aIm[i1], r2 = bRe[i2], m2 = bIm[i2];
                    double r3 = cRe[i3], m3 = cIm[i3];
                    double r12 = r1 * r2 - m1 * m2, m12 = r1 * m2 + m1 * r2;
                    br += r12 * r3 - m12 * m3;
                    bi += r12 * m3 + m12 * r3;
                }
            }
        }
        count[t] = cnt;
        ksum[3 * t] = k1; ksum[3 * t + 1] = k2; ksum[3 * t + 2] = k3;
        if (bsum) { bsum[2 * t] = br; bsum[2 * t + 1] = bi; }
    }
    for (int b = 0; b < nbin; b++) free_shell(&sh[b]);
    free(sh);
    return 0;
}

/* Exact periodic triangle counts for many triangles at once, by direct
 * enumeration of mode pairs (no FFT), for shell sets that never reach the
 * Nyquist plane (hi < (N/2)^2).  For each distinct side pair (a, b) of the
 * list every k1 of shell a is paired with every k2 of shell b and the
 * pairs are histogrammed by |n3|^2, n3 = -(k1 + k2) wrapped into
 * [-N/2, N/2); a triangle (a, b, c) then counts the pairs whose |n3|^2 lies
 * in [n2lo[c], n2hi[c]].  Shells and |wrap(.)|^2 are invariant under the 48
 * signed permutations of the axes, so k1 runs over one representative
 * x >= y >= z >= 0 per orbit, weighted by the orbit size.  Integer counts;
 * equal to oracle_direct's (checked on small grids). */
static int orbit_size(int x, int y, int z) {
    int nz = (x != 0) + (y != 0) + (z != 0);
    int perms = (x == y && y == z) ? 1 : ((x == y || y == z) ? 3 : 6);
    return perms << nz;
}

int oracle_count_sym(int N, int nbin, const long long *n2lo, const long long *n2hi,
                     int ntri, const int *tri, long long *count) {
    if (N <= 0 || (N & 1) || nbin <= 0 || ntri < 0) return -1;
    const long long nyq2 = (long long)(N / 2) * (N / 2);
    for (int b = 0; b < nbin; b++)
        if (n2hi[b] >= nyq2) return -3;
    for (int t = 0; t < ntri; t++)
        for (int s = 0; s < 3; s++)
            if (tri[3 * t + s] < 0 || tri[3 * t + s] >= nbin) return -1;
    shell_t *sh = (shell_t *)calloc(nbin, sizeof(shell_t));
    if (!sh) return -2;
    for (int b = 0; b < nbin; b++)
        if (build_shell(N, n2lo[b], n2hi[b], &sh[b])) return -2;
    /* distinct (a, b) pairs of the list */
    int *pa = (int *)malloc(sizeof(int) * (ntri + 1)), *pb = (int *)malloc(sizeof(int) * (ntri + 1));
    char *seen = (char *)calloc((size_t)nbin * nbin, 1);
    int npair = 0;
    for (int t = 0; t < ntri; t++) {
        int a = tri[3 * t], b = tri[3 * t + 1];
        if (!seen[(size_t)a * nbin + b]) {
            seen[(size_t)a * nbin + b] = 1;
            pa[npair] = a;
            pb[npair] = b;
            npair++;
        }
    }
    const long long hmax = 3 * nyq2 + 1;
    long long **hist = (long long **)calloc(npair ? npair : 1, sizeof(long long *));
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < npair; q++) {
        long long *h = (long long *)calloc(hmax, sizeof(long long));
        if (!h) {
            rc = -2;
            continue;
        }
        const shell_t *A = &sh[pa[q]], *B = &sh[pb[q]];
        for (int p = 0; p < A->n; p++) {
            int x = A->ix[p], y = A->iy[p], z = A->iz[p];
            if (!(x >= y && y >= z && z >= 0)) continue;
            long long w = orbit_size(x, y, z);
            for (int r = 0; r < B->n; r++) {
                int u = wrap(-(x + B->ix[r]), N);
                int v = wrap(-(y + B->iy[r]), N);
                int k = wrap(-(z + B->iz[r]), N);
                h[(long long)u * u + (long long)v * v + (long long)k * k] += w;
            }
        }
        hist[q] = h;
    }
    if (!rc) {
        for (int t = 0; t < ntri; t++) {
            int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
            int q = 0;
            while (pa[q] != a || pb[q] != b) q++;
            long long lo = n2lo[c] > 0 ? n2lo[c] : 0, hi = n2hi[c] < hmax - 1 ? n2hi[c] : hmax - 1;
            long long cnt = 0;
            for (long long n2 = lo; n2 <= hi; n2++) cnt += hist[q][n2];
            count[t] = cnt;
        }
    }
    for (int q = 0; q < npair; q++) free(hist[q]);
    free(hist);
    free(seen);
    free(pa);
    free(pb);
    for (int b = 0; b < nbin; b++) free_shell(&sh[b]);
    free(sh);
    return rc;
}

/* Number of members of each shell (for sanity checks). */
int oracle_shell_sizes(int N, int nbin, const long long *n2lo, const long long *n2hi,
                       long long *sizes) {
    for (int b = 0; b < nbin; b++) {
        shell_t s;
        if (build_shell(N, n2lo[b], n2hi[b], &s)) return -2;
        sizes[b] = s.n;
        free_shell(&s);
    }
    return 0;
}
