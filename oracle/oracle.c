/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement of the reference's batched tridiagonal path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing in the
 * product library links, loads or calls this file.
 *
 * Functions and the reference code they restate:
 *   oracle_thomas_f64 / oracle_thomas_f32
 *       Tridiagonal_AVX512/tridiagonal_basic.cpp:6-26 (Thomas<val_t>): forward
 *       sweep c_i <- c_i/(b_i - c_{i-1} a_i), d_i <- (d_i - d_{i-1} a_i)/(...),
 *       backward d_i <- d_i - c_i d_{i+1}; a_0 and c_{n-1} are never read.
 *       Same recurrence as Tridiagonal_CUDA/tridiagonal_basic.cpp:4-30.
 *       Pinned by closed forms (Poisson x==1, n=1, n=2 Cramer, diagonal and
 *       block-diagonal A) and by dense LU (numpy.linalg.solve) in tests/.
 *   oracle_adjoint_f64
 *       No reference counterpart (the reference has no backward pass).  Thomas
 *       on A^T (sub'_i = c_{i-1}, super'_i = a_{i+1}) plus the coefficient
 *       gradients.  Pinned by dense solves of A^T and central differences.
 *   oracle_residual_f64, oracle_errstats_f64
 *       Error statistics of Tridiagonal_CUDA/utils.cu:28-207 and
 *       Tridiagonal_AVX512/check.c:52-163 (mean |err|, RMSE, max |err| with a
 *       NaN-aware argmax), over ALL systems instead of system 0, with the max
 *       taken over |err| (the reference CUDA check takes the signed max).
 *       The relative residual has no reference counterpart; it is pinned by
 *       tests against numpy dense mat-vecs.
 *   oracle_gen
 *       Seeded generators: the reference manufactured-solution recipe of
 *       Tridiagonal_CUDA/test.cpp:19-38 (with a SplitMix64 stream in place of
 *       MKL VSL MT19937, which is not available), and the diagonally dominant
 *       distributions used by the benchmark workloads.  Pinned bit-for-bit by a
 *       pure-Python SplitMix64 in tests/.
 *
 * Layouts (element (s, i) = row i of system s):
 *   0 contiguous  : s*n + i        (Tridiagonal_CUDA/test.cpp:70,76-79)
 *   1 packed      : s*4n + i, b/c/d at +n/+2n/+3n (utils.cu:274-283)
 *   2 interleaved : i*batch + s
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_EXPORT __attribute__((visibility("default")))

static inline int64_t idx_of(int layout, int64_t s, int64_t i, int64_t n, int64_t batch) {
    if (layout == 2) return i * batch + s;
    if (layout == 1) return s * 4 * n + i;
    return s * n + i;
}

/* ---------------------------------------------------------------- Thomas */

/* Thomas<double> of tridiagonal_basic.cpp:6-26, non-destructive: x receives the
 * solution, inputs are untouched.  work must hold n doubles per thread. */
static void thomas_one_f64(const double *a, const double *b, const double *c, const double *d,
                           double *x, double *cw, int64_t n, int64_t st) {
    double cc = c[0] / b[0];
    double dd = d[0] / b[0];
    cw[0] = cc;
    x[0] = dd;
    for (int64_t i = 1; i < n; ++i) {
        double k = b[i * st] - cc * a[i * st];
        cc = c[i * st] / k;
        dd = (d[i * st] - dd * a[i * st]) / k;
        cw[i] = cc;
        x[i * st] = dd;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
        dd = x[i * st] - cw[i] * dd;
        x[i * st] = dd;
    }
}

static void thomas_one_f32(const float *a, const float *b, const float *c, const float *d,
                           float *x, float *cw, int64_t n, int64_t st) {
    float cc = c[0] / b[0];
    float dd = d[0] / b[0];
    cw[0] = cc;
    x[0] = dd;
    for (int64_t i = 1; i < n; ++i) {
        float k = b[i * st] - cc * a[i * st];
        cc = c[i * st] / k;
        dd = (d[i * st] - dd * a[i * st]) / k;
        cw[i] = cc;
        x[i * st] = dd;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
        dd = x[i * st] - cw[i] * dd;
        x[i * st] = dd;
    }
}

static int clamp_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    return nthreads;
}

/* FP64 oracle: x = A^{-1} d for every system (layouts 0, 1, 2). */
OR_EXPORT int oracle_thomas_f64(const double *a, const double *b, const double *c, const double *d,
                                double *x, int64_t n, int64_t batch, int layout, int nthreads) {
    if (n <= 0 || batch < 0) return 1;
    nthreads = clamp_threads(nthreads);
    int64_t st = (layout == 2) ? batch : 1;
    int64_t off = (layout == 1) ? n : 0; /* packed: b,c,d follow a inside the system block */
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        double *cw = (double *)malloc(sizeof(double) * (size_t)n);
        if (!cw) {
#pragma omp atomic write
            err = 2;
        } else {
#pragma omp for schedule(static)
            for (int64_t s = 0; s < batch; ++s) {
                int64_t base = idx_of(layout, s, 0, n, batch);
                thomas_one_f64(a + base, b + base + off, c + base + 2 * off, d + base + 3 * off,
                               x + base + 3 * off, cw, n, st);
            }
            free(cw);
        }
    }
    return err;
}

/* The same recurrence in single precision: the reference CPU path's arithmetic
 * for val_t=float (used as the CPU baseline, not as an oracle). */
OR_EXPORT int oracle_thomas_f32(const float *a, const float *b, const float *c, const float *d,
                                float *x, int64_t n, int64_t batch, int layout, int nthreads) {
    if (n <= 0 || batch < 0) return 1;
    nthreads = clamp_threads(nthreads);
    int64_t st = (layout == 2) ? batch : 1;
    int64_t off = (layout == 1) ? n : 0;
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        float *cw = (float *)malloc(sizeof(float) * (size_t)n);
        if (!cw) {
#pragma omp atomic write
            err = 2;
        } else {
#pragma omp for schedule(static)
            for (int64_t s = 0; s < batch; ++s) {
                int64_t base = idx_of(layout, s, 0, n, batch);
                thomas_one_f32(a + base, b + base + off, c + base + 2 * off, d + base + 3 * off,
                               x + base + 3 * off, cw, n, st);
            }
            free(cw);
        }
    }
    return err;
}

/* Ragged batch: system s occupies rows [offsets[s], offsets[s+1]) of each array. */
OR_EXPORT int oracle_thomas_ragged_f64(const double *a, const double *b, const double *c,
                                       const double *d, double *x, const int64_t *offsets,
                                       int64_t nsys, int nthreads) {
    nthreads = clamp_threads(nthreads);
    int64_t nmax = 0;
    for (int64_t s = 0; s < nsys; ++s) {
        int64_t len = offsets[s + 1] - offsets[s];
        if (len < 0) return 1;   /* empty systems are skipped */
        if (len > nmax) nmax = len;
    }
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        double *cw = (double *)malloc(sizeof(double) * (size_t)(nmax > 0 ? nmax : 1));
        if (!cw) {
#pragma omp atomic write
            err = 2;
        } else {
#pragma omp for schedule(dynamic, 64)
            for (int64_t s = 0; s < nsys; ++s) {
                int64_t o = offsets[s];
                if (offsets[s + 1] > o)
                    thomas_one_f64(a + o, b + o, c + o, d + o, x + o, cw, offsets[s + 1] - o, 1);
            }
            free(cw);
        }
    }
    return err;
}

OR_EXPORT int oracle_thomas_ragged_f32(const float *a, const float *b, const float *c,
                                       const float *d, float *x, const int64_t *offsets,
                                       int64_t nsys, int nthreads) {
    nthreads = clamp_threads(nthreads);
    int64_t nmax = 0;
    for (int64_t s = 0; s < nsys; ++s) {
        int64_t len = offsets[s + 1] - offsets[s];
        if (len < 0) return 1;   /* empty systems are skipped */
        if (len > nmax) nmax = len;
    }
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        float *cw = (float *)malloc(sizeof(float) * (size_t)(nmax > 0 ? nmax : 1));
        if (!cw) {
#pragma omp atomic write
            err = 2;
        } else {
#pragma omp for schedule(dynamic, 64)
            for (int64_t s = 0; s < nsys; ++s) {
                int64_t o = offsets[s];
                if (offsets[s + 1] > o)
                    thomas_one_f32(a + o, b + o, c + o, d + o, x + o, cw, offsets[s + 1] - o, 1);
            }
            free(cw);
        }
    }
    return err;
}

/* ---------------------------------------------------------------- adjoint */

/* Given x = A^{-1} d and the upstream gradient g = dL/dx, computes
 *   lambda = A^{-T} g   (Thomas on A^T: sub'_i = c_{i-1}, diag b_i, super'_i = a_{i+1})
 *   ga_i = -lambda_i x_{i-1} (0 for i = 0), gb_i = -lambda_i x_i,
 *   gc_i = -lambda_i x_{i+1} (0 for i = n-1), gd_i = lambda_i.
 * a_0 and c_{n-1} are not part of A (the reference ignores them), so their
 * gradients are 0. */
OR_EXPORT int oracle_adjoint_f64(const double *a, const double *b, const double *c,
                                 const double *x, const double *g, double *ga, double *gb,
                                 double *gc, double *gd, int64_t n, int64_t batch, int layout,
                                 int nthreads) {
    if (n <= 0 || batch < 0 || layout == 1) return 1;
    nthreads = clamp_threads(nthreads);
    int64_t st = (layout == 2) ? batch : 1;
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        double *w = (double *)malloc(sizeof(double) * (size_t)n * 5);
        if (!w) {
#pragma omp atomic write
            err = 2;
        } else {
            double *sa = w, *sc = w + n, *cw = w + 2 * n, *lam = w + 3 * n, *sb = w + 4 * n;
#pragma omp for schedule(static)
            for (int64_t s = 0; s < batch; ++s) {
                int64_t base = idx_of(layout, s, 0, n, batch);
                for (int64_t i = 0; i < n; ++i) {
                    sa[i] = (i > 0) ? c[base + (i - 1) * st] : 0.0;       /* sub of A^T  */
                    sc[i] = (i < n - 1) ? a[base + (i + 1) * st] : 0.0;   /* super of A^T */
                    sb[i] = b[base + i * st];
                }
                /* lam = A^{-T} g via the same Thomas recurrence (unit stride) */
                for (int64_t i = 0; i < n; ++i) lam[i] = g[base + i * st];
                thomas_one_f64(sa, sb, sc, lam, lam, cw, n, 1);
                for (int64_t i = 0; i < n; ++i) {
                    double li = lam[i];
                    double xm = (i > 0) ? x[base + (i - 1) * st] : 0.0;
                    double xp = (i < n - 1) ? x[base + (i + 1) * st] : 0.0;
                    ga[base + i * st] = (i > 0) ? -li * xm : 0.0;
                    gb[base + i * st] = -li * x[base + i * st];
                    gc[base + i * st] = (i < n - 1) ? -li * xp : 0.0;
                    gd[base + i * st] = li;
                }
            }
            free(w);
        }
    }
    return err;
}

/* ---------------------------------------------------------------- checks */

/* Relative residual per system, max over systems:
 *   ||A x - d||_inf / (||A||_inf ||x||_inf + ||d||_inf),  a_0 = c_{n-1} = 0. */
OR_EXPORT double oracle_residual_f64(const double *a, const double *b, const double *c,
                                     const double *d, const double *x, int64_t n, int64_t batch,
                                     int layout) {
    int64_t st = (layout == 2) ? batch : 1;
    int64_t off = (layout == 1) ? n : 0;
    double worst = 0.0;
    int saw_nan = 0;
    for (int64_t s = 0; s < batch; ++s) {
        int64_t base = idx_of(layout, s, 0, n, batch);
        const double *A = a + base, *B = b + base + off, *C = c + base + 2 * off,
                     *D = d + base + 3 * off;
        const double *X = x + base + 3 * off;
        double rmax = 0.0, anorm = 0.0, xnorm = 0.0, dnorm = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double ai = (i > 0) ? A[i * st] : 0.0;
            double ci = (i < n - 1) ? C[i * st] : 0.0;
            double xm = (i > 0) ? X[(i - 1) * st] : 0.0;
            double xp = (i < n - 1) ? X[(i + 1) * st] : 0.0;
            double r = ai * xm + B[i * st] * X[i * st] + ci * xp - D[i * st];
            if (isnan(r)) saw_nan = 1;
            if (fabs(r) > rmax) rmax = fabs(r);
            double row = fabs(ai) + fabs(B[i * st]) + fabs(ci);
            if (row > anorm) anorm = row;
            if (fabs(X[i * st]) > xnorm) xnorm = fabs(X[i * st]);
            if (fabs(D[i * st]) > dnorm) dnorm = fabs(D[i * st]);
        }
        double den = anorm * xnorm + dnorm;
        double rel = (den > 0.0) ? rmax / den : rmax;
        if (isnan(rel)) saw_nan = 1;
        if (rel > worst) worst = rel;
    }
    return saw_nan ? NAN : worst;
}

/* Error statistics of x against x_ref over all n*batch entries.
 * out[0] mean |err|, out[1] RMSE, out[2] max |err| (NaN if any NaN),
 * out[3] max over systems of ||x - x_ref||_inf / ||x_ref||_inf,
 * out[4] number of non-finite entries in x.
 * Returns the flat index of the max |err| entry (first NaN if any).  Pairwise
 * summation follows the intent of Tridiagonal_AVX512/check.c:1-50. */
static double pairwise_sum(const double *v, int64_t len) {
    if (len <= 16) {
        double s = 0.0;
        for (int64_t i = 0; i < len; ++i) s += v[i];
        return s;
    }
    int64_t h = len / 2;
    return pairwise_sum(v, h) + pairwise_sum(v + h, len - h);
}

OR_EXPORT int64_t oracle_errstats_f64(const double *x, const double *xref, int64_t n, int64_t batch,
                                      int layout, double *out) {
    int64_t total = n * batch;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1) * 2);
    if (!tmp) return -1;
    double *abse = tmp, *sq = tmp + total;
    int64_t st = (layout == 2) ? batch : 1;
    int64_t off = (layout == 1) ? 3 * n : 0;
    double emax = -1.0, relmax = 0.0, nonfinite = 0.0;
    int64_t argmax = 0, k = 0;
    int nan_found = 0, rel_nan = 0;
    for (int64_t s = 0; s < batch; ++s) {
        int64_t base = idx_of(layout, s, 0, n, batch) + off;
        double emax_s = 0.0, rmax_s = 0.0;
        int nan_s = 0;
        for (int64_t i = 0; i < n; ++i, ++k) {
            int64_t id = base + i * st;
            double e = x[id] - xref[id];
            if (!isfinite(x[id])) nonfinite += 1.0;
            double ae = fabs(e);
            abse[k] = ae;
            sq[k] = e * e;
            if (isnan(e)) {
                nan_s = 1;
                if (!nan_found) { nan_found = 1; argmax = id; }
            } else if (!nan_found && ae > emax) {
                emax = ae;
                argmax = id;
            }
            if (!isnan(e) && ae > emax_s) emax_s = ae;
            if (fabs(xref[id]) > rmax_s) rmax_s = fabs(xref[id]);
        }
        double rel = nan_s ? NAN : ((rmax_s > 0.0) ? emax_s / rmax_s : emax_s);
        if (isnan(rel)) rel_nan = 1;
        else if (rel > relmax) relmax = rel;
    }
    out[0] = pairwise_sum(abse, total) / (double)(total > 0 ? total : 1);
    out[1] = sqrt(pairwise_sum(sq, total) / (double)(total > 0 ? total : 1));
    out[2] = nan_found ? NAN : (emax < 0.0 ? 0.0 : emax);
    out[3] = rel_nan ? NAN : relmax;
    out[4] = nonfinite;
    free(tmp);
    return argmax;
}

/* ---------------------------------------------------------------- generators */

/* SplitMix64 (Steele, Lea, Flood 2014).  One independent stream per system:
 * state_s = seed ^ (s * 0xD1B54A32D192ED03). */
static inline uint64_t sm64_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double sm64_unif(uint64_t *state, double lo, double hi) {
    double u = (double)(sm64_next(state) >> 11) * 0x1.0p-53;
    return lo + (hi - lo) * u;
}

OR_EXPORT uint64_t oracle_sm64_stream_seed(uint64_t seed, int64_t s) {
    return seed ^ ((uint64_t)s * 0xD1B54A32D192ED03ULL);
}

OR_EXPORT void oracle_sm64_fill(uint64_t state, double *out, int64_t count, double lo, double hi) {
    for (int64_t i = 0; i < count; ++i) out[i] = sm64_unif(&state, lo, hi);
}

/* Diagonally dominant random systems (benchmark distribution).
 *   a_i, c_i ~ U[-1,1] with a_0 = c_{n-1} = 0,
 *   b_i = |a_i| + |c_i| + U[blo, bhi],  d_i ~ U[-1,1].
 * Per system, draws happen in the order a[0..n), c[0..n), b[0..n), d[0..n).
 * Values are written in double; the caller rounds to float for FP32 runs. */
OR_EXPORT void oracle_gen_dd(double *a, double *b, double *c, double *d, int64_t n, int64_t batch,
                             int layout, uint64_t seed, double blo, double bhi, int nthreads) {
    nthreads = clamp_threads(nthreads);
    int64_t st = (layout == 2) ? batch : 1;
    int64_t off = (layout == 1) ? n : 0;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t s = 0; s < batch; ++s) {
        uint64_t state = oracle_sm64_stream_seed(seed, s);
        int64_t base = idx_of(layout, s, 0, n, batch);
        double *A = a + base, *B = b + base + off, *C = c + base + 2 * off, *D = d + base + 3 * off;
        for (int64_t i = 0; i < n; ++i) A[i * st] = sm64_unif(&state, -1.0, 1.0);
        for (int64_t i = 0; i < n; ++i) C[i * st] = sm64_unif(&state, -1.0, 1.0);
        A[0] = 0.0;
        C[(n - 1) * st] = 0.0;
        for (int64_t i = 0; i < n; ++i)
            B[i * st] = fabs(A[i * st]) + fabs(C[i * st]) + sm64_unif(&state, blo, bhi);
        for (int64_t i = 0; i < n; ++i) D[i * st] = sm64_unif(&state, -1.0, 1.0);
    }
}

/* Ragged variant: system s occupies [offsets[s], offsets[s+1]). */
OR_EXPORT void oracle_gen_dd_ragged(double *a, double *b, double *c, double *d,
                                    const int64_t *offsets, int64_t nsys, uint64_t seed,
                                    double blo, double bhi, int nthreads) {
    nthreads = clamp_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t s = 0; s < nsys; ++s) {
        uint64_t state = oracle_sm64_stream_seed(seed, s);
        int64_t o = offsets[s], n = offsets[s + 1] - offsets[s];
        if (n <= 0) continue;            /* empty system: no rows to write */
        double *A = a + o, *B = b + o, *C = c + o, *D = d + o;
        for (int64_t i = 0; i < n; ++i) A[i] = sm64_unif(&state, -1.0, 1.0);
        for (int64_t i = 0; i < n; ++i) C[i] = sm64_unif(&state, -1.0, 1.0);
        A[0] = 0.0;
        C[n - 1] = 0.0;
        for (int64_t i = 0; i < n; ++i) B[i] = fabs(A[i]) + fabs(C[i]) + sm64_unif(&state, blo, bhi);
        for (int64_t i = 0; i < n; ++i) D[i] = sm64_unif(&state, -1.0, 1.0);
    }
}

/* Manufactured-solution generator: the recipe of Tridiagonal_CUDA/test.cpp:19-38
 * (one system, later replicated across the batch by the caller):
 *   a[1..n) ~ U[-1,1], c[0..n-1) ~ U[-1,1], x ~ U[-1,1], b ~ U[0,1] (draw order
 *   a, c, x, b as in test.cpp:23-26), a_0 = c_{n-1} = 0,
 *   b_0 += max(|c_0|, |a_1|), b_i += max(|a_i|+|c_i|, |a_{i+1}|+|c_{i-1}|),
 *   b_{n-1} += max(|a_{n-1}|, |c_{n-2}|), d = A x.
 * single != 0 rounds a, b, c, x to float and forms d in float arithmetic, as the
 * reference does for val_t = float. */
OR_EXPORT void oracle_gen_manufactured(double *a, double *b, double *c, double *d, double *x,
                                       int64_t n, uint64_t seed, int single) {
    uint64_t state = oracle_sm64_stream_seed(seed, 0);
    for (int64_t i = 1; i < n; ++i) a[i] = sm64_unif(&state, -1.0, 1.0);
    for (int64_t i = 0; i < n - 1; ++i) c[i] = sm64_unif(&state, -1.0, 1.0);
    for (int64_t i = 0; i < n; ++i) x[i] = sm64_unif(&state, -1.0, 1.0);
    for (int64_t i = 0; i < n; ++i) b[i] = sm64_unif(&state, 0.0, 1.0);
    a[0] = 0.0;
    c[n - 1] = 0.0;
    if (single) {
        for (int64_t i = 0; i < n; ++i) {
            a[i] = (float)a[i]; c[i] = (float)c[i]; x[i] = (float)x[i]; b[i] = (float)b[i];
        }
    }
#define BUMP(i, v) do { if (single) b[i] = (float)((float)b[i] + (float)(v)); else b[i] += (v); } while (0)
    if (n == 1) {
        d[0] = single ? (double)((float)b[0] * (float)x[0]) : b[0] * x[0];
        return;
    }
    BUMP(0, fmax(fabs(c[0]), fabs(a[1])));
    for (int64_t i = 1; i < n - 1; ++i)
        BUMP(i, fmax(fabs(a[i]) + fabs(c[i]), fabs(a[i + 1]) + fabs(c[i - 1])));
    BUMP(n - 1, fmax(fabs(a[n - 1]), fabs(c[n - 2])));
#undef BUMP
    if (single) {
        float fa, fb, fc;
        d[0] = (float)((float)b[0] * (float)x[0] + (float)c[0] * (float)x[1]);
        for (int64_t i = 1; i < n - 1; ++i) {
            fa = (float)a[i] * (float)x[i - 1];
            fb = (float)b[i] * (float)x[i];
            fc = (float)c[i] * (float)x[i + 1];
            d[i] = (float)(fa + fb + fc);
        }
        d[n - 1] = (float)((float)a[n - 1] * (float)x[n - 2] + (float)b[n - 1] * (float)x[n - 1]);
    } else {
        d[0] = b[0] * x[0] + c[0] * x[1];
        for (int64_t i = 1; i < n - 1; ++i) d[i] = a[i] * x[i - 1] + b[i] * x[i] + c[i] * x[i + 1];
        d[n - 1] = a[n - 1] * x[n - 2] + b[n - 1] * x[n - 1];
    }
}
