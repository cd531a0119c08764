/*
 * CPU oracle for the thread-coarsened stencil path.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are compared against it: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (threadcoarsenedstencilgenerator_amd) never
 * links or calls it.
 *
 * It restates, in plain C with OpenMP, the arithmetic that the reference's
 * generated kernels perform.  Every function names the reference file:line it
 * follows (paths relative to the reference repository root):
 *
 *   offsets / stencil size ..... hip/generate2DStencil_hip.py:63-67,175-186
 *                                hip/generate3DStencil_hip.py:66-70,201-212
 *   mask layout ................ hip/generate2DStencil_hip.py:120-144
 *                                hip/generate3DStencil_hip.py:132-169
 *   summation order / init ..... hip/generate2DStencil_hip.py:146-198,245-298
 *                                hip/generate3DStencil_hip.py:171-218,266-361
 *   Jacobi constant ............ hip/generate2DStencil_hip.py:300-307
 *                                hip/generate3DStencil_hip.py:363-370
 *   geometry / padding ......... hip/benchmark.cpp:140-177, hip/benchmark.h:100-123
 *   thread mapping / predicate . hip/generate{2,3}DStencil_hip.py printFunctionInit/printCheck
 *   reference init ............. hip/generate2DStencil_hip.py:322-351
 *                                hip/generate3DStencil_hip.py:387-431
 *   ping-pong stepping ......... hip/generateLaunchUpdateGrid.py:71-75,129-137
 *   block-shape sweep order .... hip/primes.cpp:4-88, hip/benchmark.cpp:221-316
 *
 * Arithmetic contract: compiled with -ffp-contract=off; the Variable update is
 * acc = v0*w0 followed by acc = fma(vk, wk, acc); Jacobi is acc = v0, acc += vk,
 * out = acc * fl(1/n).  The FP32 compensated variant (no reference
 * counterpart) is restated in oracle_step_f32c below.
 *
 * Pinning: see tests/test_oracle_*.py (delta-mask shifts, closed-form init
 * answers, constant/linear fields, isotropic-vs-Jacobi, traversal-order
 * derivation from the reference's loop nests, brute force on tiny grids, and
 * the reference's own primes.cpp compiled into oracle/_ref).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_BOX 0
#define OR_STAR 1
#define OR_ROW 0
#define OR_COLUMN 1
#define OR_JACOBI 0
#define OR_VARIABLE 1
#define OR_ALIGN_NONE 0
#define OR_ALIGN_READ 1
#define OR_ALIGN_WRITE 2

/* Stencil size n.  hip/benchmark.cpp:62-73 (Box (2r+1)^d, Star 2dr+1). */
int oracle_stencil_size(int dim, int nb, int r)
{
    if (nb == OR_BOX) {
        int w = 2 * r + 1;
        return dim == 2 ? w * w : w * w * w;
    }
    return 2 * dim * r + 1;
}

/* Mask index of offset (dx,dy,dz).  Follows calcMaskAddress,
 * hip/generate2DStencil_hip.py:120-144 and hip/generate3DStencil_hip.py:132-169.
 * Returns -1 if the offset is not in the neighbourhood. */
int oracle_mask_index(int dim, int nb, int r, int dx, int dy, int dz)
{
    int w = 2 * r + 1;
    if (dim == 2) dz = 0;
    if (nb == OR_BOX) {
        if (abs(dx) > r || abs(dy) > r || abs(dz) > r) return -1;
        if (dim == 2) return (dx + r) + (dy + r) * w;
        return (dx + r) + (dy + r) * w + (dz + r) * w * w;
    }
    /* Star: at most one non-zero component */
    int nz = (dx != 0) + (dy != 0) + (dz != 0);
    if (nz > 1 || abs(dx) > r || abs(dy) > r || abs(dz) > r) return -1;
    if (dim == 2) {
        if (dx == 0 && dy < 0) return dy + r;             /* upper arm   :130-134 */
        if (dy == 0) return r + (dx + r);                 /* row         :135-139 */
        return (dy + r) + 2 * r;                          /* lower arm   :140-144 */
    }
    if (dz < 0) return dz + r;                            /* front arm   :144-147 */
    if (dz == 0 && dx == 0 && dy < 0) return (dy + r) + r;/*             :149-153 */
    if (dz == 0 && dy == 0) return (dx + r) + 2 * r;      /*             :154-158 */
    if (dz == 0 && dx == 0 && dy > 0) return (dy + r) + 3 * r; /*        :159-163 */
    return (dz + r) + 4 * r;                              /* back arm    :165-169 */
}

/* Offsets in per-output summation order.  Row: (dz,dy,dx) lexicographic;
 * Column: (dz,dx,dy) lexicographic.  This is the order in which the
 * reference's traversal loops (hip/generate2DStencil_hip.py:245-298,
 * hip/generate3DStencil_hip.py:266-361) visit the neighbourhood of each
 * output point; the first element is the initialising contributor
 * (hip/generate2DStencil_hip.py:179,188-196; hip/generate3DStencil_hip.py:146,155).
 * Writes n triples (dx,dy,dz) to off and the mask index of each to midx. */
int oracle_offsets(int dim, int nb, int orient, int r, int *off, int *midx)
{
    int n = 0;
    int zr = dim == 3 ? r : 0;
    for (int dz = -zr; dz <= zr; ++dz) {
        for (int a = -r; a <= r; ++a) {       /* outer in-plane axis */
            for (int b = -r; b <= r; ++b) {   /* inner in-plane axis */
                int dx, dy;
                if (orient == OR_ROW) { dy = a; dx = b; } else { dx = a; dy = b; }
                int m = oracle_mask_index(dim, nb, r, dx, dy, dz);
                if (m < 0) continue;
                off[3 * n + 0] = dx; off[3 * n + 1] = dy; off[3 * n + 2] = dz;
                midx[n] = m;
                ++n;
            }
        }
    }
    return n;
}

/* Geometry after truncation and alignment padding.
 * hip/benchmark.cpp:140-177 and hip/benchmark.h:100-123.
 * in : user grid size incl. halo (size[0..dim-1]), coarsening c[0..2].
 * out: g[0..2] = padded xdim, ydim, zdim; g[3] = leftOffset, g[4] = rightOffset,
 *      g[5..7] = interior elements per axis, g[8..10] = threads per axis.
 * Returns 0 on success, -1 if the grid is too small for the halo. */
int oracle_geometry(int dim, int r, int align, int cl, const unsigned *size,
                    const unsigned *c, unsigned long *g)
{
    unsigned t[3] = {1, 1, 1}, e[3] = {1, 1, 1};
    for (int a = 0; a < dim; ++a) {
        if (size[a] <= (unsigned)(2 * r) || c[a] == 0) return -1;
        t[a] = size[a] - (size[a] - 2 * r) % c[a];
        e[a] = t[a] - 2 * r;
    }
    unsigned long lo = 0, ro = 0, xmem = t[0];
    if (align == OR_ALIGN_READ) {
        if (xmem % cl) { ro = cl - xmem % cl; xmem += ro; }
    } else if (align == OR_ALIGN_WRITE) {
        if (r % cl) lo = cl - r % cl;
        xmem += lo;
        if (xmem % cl) ro = cl - xmem % cl;
        xmem += ro;
    }
    g[0] = xmem; g[1] = t[1]; g[2] = dim == 3 ? t[2] : 1;
    g[3] = lo; g[4] = ro;
    for (int a = 0; a < 3; ++a) {
        g[5 + a] = a < dim ? e[a] : 1;
        g[8 + a] = a < dim ? e[a] / c[a] : 1;
    }
    return 0;
}

/* Reference init values (both buffers; halos are constant forever).
 * 2D: hip/generate2DStencil_hip.py:322-351; 3D: hip/generate3DStencil_hip.py:387-431.
 * Points outside [lo, xdim-ro) (the padding) are left untouched. */
void oracle_init_reference(int dim, int r, unsigned long xdim, unsigned long ydim,
                           unsigned long zdim, unsigned long lo, unsigned long ro,
                           double *a)
{
    for (unsigned long k = 0; k < zdim; ++k)
        for (unsigned long j = 0; j < ydim; ++j)
            for (unsigned long i = lo; i < xdim - ro; ++i) {
                unsigned long idx = i + j * xdim + k * xdim * ydim;
                int halo = (i - lo < (unsigned long)r) || (j < (unsigned long)r) ||
                           (dim == 3 && k < (unsigned long)r);
                if (dim == 2) a[idx] = halo ? 1.0 : 0.0;
                else a[idx] = halo ? ((idx % 2 == 0) ? 0.75 : 0.25) : 0.0;
            }
}

/* ---- counter-based SplitMix64 inputs (no reference counterpart: the
 * reference only has its fixed init; these are the seeded synthetic inputs
 * of the benchmark workloads) ---- */
static uint64_t sm64(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double oracle_u01(uint64_t seed, uint64_t idx) { return (double)(sm64(seed, idx) >> 11) * 0x1.0p-53; }
double oracle_upm1(uint64_t seed, uint64_t idx) { return (double)(sm64(seed, idx) >> 11) * 0x1.0p-52 - 1.0; }

/* Seeded interior fill; halo and padding set to `halo`.  mode 0: U[0,1),
 * mode 1: U[-1,1), mode 2: N(0,1) by Box-Muller.  The counter is the
 * interior-relative linear index, so values do not depend on padding. */
void oracle_init_seeded(int dim, int r, int mode, uint64_t seed, double halo,
                        unsigned long xdim, unsigned long ydim, unsigned long zdim,
                        unsigned long lo, unsigned long ro, double *a)
{
    unsigned long nx = xdim - lo - ro - 2 * r, ny = ydim - 2 * r;
    unsigned long zr = dim == 3 ? r : 0;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned long k = 0; k < zdim; ++k)
        for (unsigned long j = 0; j < ydim; ++j)
            for (unsigned long i = 0; i < xdim; ++i) {
                unsigned long idx = i + j * xdim + k * xdim * ydim;
                int inside = i >= lo + r && i < xdim - ro - r && j >= (unsigned long)r &&
                             j < ydim - r && k >= zr && k < zdim - zr;
                if (!inside) { a[idx] = halo; continue; }
                uint64_t li = (i - lo - r) + nx * ((j - r) + ny * (k - zr));
                if (mode == 0) a[idx] = oracle_u01(seed, li);
                else if (mode == 1) a[idx] = oracle_upm1(seed, li);
                else {
                    double u1 = 1.0 - oracle_u01(seed, 2 * li);
                    double u2 = oracle_u01(seed, 2 * li + 1);
                    a[idx] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
                }
            }
}

/* Variable-mask weights U[-1,1) from the seed, rescaled so sum |w| = 1. */
void oracle_mask_seeded(uint64_t seed, int n, double *w)
{
    double s = 0.0;
    for (int k = 0; k < n; ++k) { w[k] = oracle_upm1(seed, (uint64_t)k); s += fabs(w[k]); }
    for (int k = 0; k < n; ++k) w[k] = w[k] / s;
}

/* One FP64 time step over the interior: out[p] = stencil(in)[p].
 * Arithmetic per output follows hip/generate2DStencil_hip.py:146-165,206-211
 * (printInit / printSingleAddition / writeBack) with the fma that the
 * reference compiler's contraction of `result += reg*d_mask[i]` produces.
 * The offset loop sits outside the x loop so the x loop vectorises; the
 * per-point order is unchanged. */
void oracle_step_f64(int dim, int nb, int orient, int coeff, int r,
                     unsigned long xdim, unsigned long ydim, unsigned long zdim,
                     unsigned long lo, unsigned long ro,
                     const double *mask, const double *in, double *out)
{
    int nmax = oracle_stencil_size(dim, nb, r);
    int *off = (int *)malloc(sizeof(int) * 3 * nmax), *midx = (int *)malloc(sizeof(int) * nmax);
    long *delta = (long *)malloc(sizeof(long) * nmax);
    double *w = (double *)malloc(sizeof(double) * nmax);
    int n = oracle_offsets(dim, nb, orient, r, off, midx);
    for (int t = 0; t < n; ++t) {
        delta[t] = off[3 * t] + (long)off[3 * t + 1] * (long)xdim +
                   (long)off[3 * t + 2] * (long)(xdim * ydim);
        w[t] = coeff == OR_VARIABLE ? mask[midx[t]] : 1.0;
    }
    const double c = 1.0 / (double)n;
    unsigned long zr = dim == 3 ? r : 0;
    unsigned long x0 = lo + r, x1 = xdim - ro - r, nx = x1 - x0;
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * (nx ? nx : 1));
#pragma omp for collapse(2) schedule(static)
        for (unsigned long k = zr; k < zdim - zr; ++k)
            for (unsigned long j = r; j < ydim - r; ++j) {
                const double *row = in + x0 + j * xdim + k * xdim * ydim;
                double *orow = out + x0 + j * xdim + k * xdim * ydim;
                if (coeff == OR_VARIABLE) {
                    const double *p = row + delta[0];
                    for (unsigned long i = 0; i < nx; ++i) acc[i] = p[i] * w[0];
                    for (int t = 1; t < n; ++t) {
                        const double *q = row + delta[t];
                        const double wt = w[t];
                        for (unsigned long i = 0; i < nx; ++i) acc[i] = fma(q[i], wt, acc[i]);
                    }
                    for (unsigned long i = 0; i < nx; ++i) orow[i] = acc[i];
                } else {
                    const double *p = row + delta[0];
                    for (unsigned long i = 0; i < nx; ++i) acc[i] = p[i];
                    for (int t = 1; t < n; ++t) {
                        const double *q = row + delta[t];
                        for (unsigned long i = 0; i < nx; ++i) acc[i] = acc[i] + q[i];
                    }
                    for (unsigned long i = 0; i < nx; ++i) orow[i] = acc[i] * c;
                }
            }
        free(acc);
    }
    free(off); free(midx); free(delta); free(w);
}

/* `steps` ping-pong steps starting from a (step t reads buffer t%2).
 * hip/generateLaunchUpdateGrid.py:129-137 (launch, then swap).  Returns the
 * index (0 or 1) of the buffer holding the final state. */
int oracle_run_f64(int dim, int nb, int orient, int coeff, int r,
                   unsigned long xdim, unsigned long ydim, unsigned long zdim,
                   unsigned long lo, unsigned long ro, const double *mask,
                   double *a, double *b, int steps)
{
    double *bufs[2] = {a, b};
    for (int s = 0; s < steps; ++s)
        oracle_step_f64(dim, nb, orient, coeff, r, xdim, ydim, zdim, lo, ro, mask,
                        bufs[s & 1], bufs[(s + 1) & 1]);
    return steps & 1;
}

/* ---- FP32 storage with compensated accumulation (no reference counterpart).
 * Per output, the offsets are consumed in the same canonical order as the
 * FP64 path, in consecutive groups of g terms.  A group is summed in FP32
 * with one rounding per term: q = v0*w0, q = fma(vk, wk, q) (Variable) or
 * q = v0, q = q + vk (Jacobi).  Group sums are combined with TwoSum (Knuth):
 * the first group sets s; the second sets (s, c) = TwoSum(s, q); later ones
 * do (s, e) = TwoSum(s, q), c = c + e.  The output is s (one group) or
 * s + c, times c32 = fl32(fl64(1/n)) for Jacobi.  Weights are the FP32
 * rounding of the FP64 mask.  -ffp-contract=off keeps every operation as
 * written. */
static inline void two_sum_f(float a, float b, float *s, float *e)
{
    float x = a + b;
    float bp = x - a;
    float ap = x - bp;
    *e = (a - ap) + (b - bp);
    *s = x;
}

void oracle_step_f32c(int dim, int nb, int orient, int coeff, int r, int group,
                      unsigned long xdim, unsigned long ydim, unsigned long zdim,
                      unsigned long lo, unsigned long ro,
                      const double *mask, const float *in, float *out)
{
    int nmax = oracle_stencil_size(dim, nb, r);
    int *off = (int *)malloc(sizeof(int) * 3 * nmax), *midx = (int *)malloc(sizeof(int) * nmax);
    long *delta = (long *)malloc(sizeof(long) * nmax);
    float *w = (float *)malloc(sizeof(float) * nmax);
    int n = oracle_offsets(dim, nb, orient, r, off, midx);
    for (int t = 0; t < n; ++t) {
        delta[t] = off[3 * t] + (long)off[3 * t + 1] * (long)xdim +
                   (long)off[3 * t + 2] * (long)(xdim * ydim);
        w[t] = coeff == OR_VARIABLE ? (float)mask[midx[t]] : 1.0f;
    }
    const float c32 = (float)(1.0 / (double)n);
    const int g = group;
    const int ngroups = (n + g - 1) / g;
    unsigned long zr = dim == 3 ? r : 0;
    unsigned long x0 = lo + r, x1 = xdim - ro - r;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned long k = zr; k < zdim - zr; ++k)
        for (unsigned long j = r; j < ydim - r; ++j)
            for (unsigned long i = x0; i < x1; ++i) {
                const float *p = in + i + j * xdim + k * xdim * ydim;
                float q = 0.0f, s = 0.0f, c = 0.0f;
                for (int t = 0; t < n; ++t) {
                    const float v = p[delta[t]];
                    if (t % g == 0) q = coeff == OR_VARIABLE ? v * w[t] : v;
                    else q = coeff == OR_VARIABLE ? fmaf(v, w[t], q) : q + v;
                    if (t % g == g - 1 || t == n - 1) {
                        int gi = t / g;
                        if (gi == 0) s = q;
                        else if (gi == 1) two_sum_f(s, q, &s, &c);
                        else { float e; two_sum_f(s, q, &s, &e); c = c + e; }
                    }
                }
                float res = ngroups > 1 ? s + c : s;
                out[i + j * xdim + k * xdim * ydim] = coeff == OR_VARIABLE ? res : res * c32;
            }
    free(off); free(midx); free(delta); free(w);
}

/* ---- block-shape sweep order: hip/primes.cpp:4-88 ----
 * Divisors of T as the mixed-radix product over prime powers in ascending
 * prime order, with the newest prime as the outer loop (earlier primes vary
 * fastest).  Returns the count written to out. */
int oracle_divisors(unsigned T, unsigned *out)
{
    unsigned primes[32], mult[32];
    int np = 0;
    unsigned n = T;
    for (unsigned i = 2; i * i <= n; ++i) {
        unsigned c = 0;
        while (n % i == 0) { n /= i; ++c; }
        if (c) { primes[np] = i; mult[np] = c; ++np; }
    }
    if (n != 1) { primes[np] = n; mult[np] = 1; ++np; }
    int cnt = 1;
    out[0] = 1;
    for (int p = 0; p < np; ++p) {
        unsigned tmp[4096];
        int tc = 0;
        unsigned f = 1;
        for (unsigned e = 0; e <= mult[p]; ++e) {
            for (int i = 0; i < cnt; ++i) tmp[tc++] = out[i] * f;
            f *= primes[p];
        }
        memcpy(out, tmp, sizeof(unsigned) * tc);
        cnt = tc;
    }
    return cnt;
}

/* All swept block shapes, hip/benchmark.cpp:221-249 (2D) and :285-316 (3D),
 * for warps = 1..maxThreads/warpSize, skipping shapes above maxDims.
 * out receives triples (x,y,z) with z = 1 in 2D.  Returns the count. */
int oracle_block_shapes(int dim, unsigned max_threads, unsigned warp,
                        const unsigned *max_dims, unsigned *out, int cap)
{
    int cnt = 0;
    unsigned d1[4096], d2[4096];
    for (unsigned w = 1; w <= max_threads / warp; ++w) {
        unsigned T = w * warp;
        int n1 = oracle_divisors(T, d1);
        for (int a = 0; a < n1; ++a) {
            if (dim == 2) {
                unsigned x = d1[a], y = T / d1[a];
                if (x > max_dims[0] || y > max_dims[1]) continue;
                if (cnt < cap) { out[3 * cnt] = x; out[3 * cnt + 1] = y; out[3 * cnt + 2] = 1; }
                ++cnt;
            } else {
                int n2 = oracle_divisors(T / d1[a], d2);
                for (int b = 0; b < n2; ++b) {
                    unsigned x = d1[a], y = d2[b], z = T / d1[a] / d2[b];
                    if (x > max_dims[0] || y > max_dims[1] || z > max_dims[2]) continue;
                    if (cnt < cap) { out[3 * cnt] = x; out[3 * cnt + 1] = y; out[3 * cnt + 2] = z; }
                    ++cnt;
                }
            }
        }
    }
    return cnt;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
