/* CPU oracle for the holomap kNN-graph path -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiled with -O2 -ffp-contract=off and no fast-math: every float
 * operation below rounds once, in source order, exactly as the Rust
 * reference's (Rust never contracts a*b+c and folds iterators left to
 * right). See hm_oracle.h for the reference line each function follows.
 */
#include "hm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* src/metric.rs:42-53: f64 difference, squared, summed left to right from
 * the first feature; sqrt in f64 (correctly rounded); cast to f32 (RNE). */
static double euclidean64(const float* a, const float* b, size_t d) {
    double sq = -0.0; /* Rust's f64 Sum starts at -0.0; terms are >= +0 */
    for (size_t f = 0; f < d; ++f) {
        double t = (double)a[f] - (double)b[f];
        sq = sq + t * t;
    }
    return sqrt(sq);
}
float hmo_euclidean(const float* a, const float* b, size_t d) {
    return (float)euclidean64(a, b, d);
}

/* src/metric.rs:58-70: three independent sequential f64 folds, then
 * (1 - dot/(sqrt(na)*sqrt(nb))).max(0.0) -- f64::max returns the non-NaN
 * operand, so a zero vector (0/0 = NaN) yields distance 0. */
static double cosine64(const float* a, const float* b, size_t d) {
    double dot = 0.0, na = 0.0, nb = 0.0;
    for (size_t f = 0; f < d; ++f) {
        double x = (double)a[f], y = (double)b[f];
        dot = dot + x * y;
        na = na + x * x;
        nb = nb + y * y;
    }
    double v = 1.0 - dot / (sqrt(na) * sqrt(nb));
    return (v != v) ? 0.0 : (v > 0.0 ? v : 0.0); /* f64::max(v, 0.0) */
}
float hmo_cosine(const float* a, const float* b, size_t d) {
    return (float)cosine64(a, b, d);
}

/* The f64 values the kNN distances round (src/metric.rs before the `as
 * f32`), for given pairs: out[i*k+t] for row i and column idx[i*k+t]; the
 * row itself at 0, as knn.rs:38-40 sets it. */
void hmo_pair_dist64(const float* data, size_t n, size_t d, size_t k, int metric,
                     const uint32_t* idx, double* out) {
    for (size_t i = 0; i < n; ++i)
        for (size_t t = 0; t < k; ++t) {
            const size_t j = idx[i * k + t];
            out[i * k + t] = j == i ? 0.0
                           : metric == 0 ? euclidean64(data + i * d, data + j * d, d)
                                         : cosine64(data + i * d, data + j * d, d);
        }
}

/* f32::total_cmp key: order-preserving map of the bit pattern to int32. */
static inline int32_t total_key(float x) {
    int32_t i = (int32_t)f2u(x);
    return i ^ (int32_t)((uint32_t)(i >> 31) >> 1);
}

typedef struct { float d; uint32_t j; } pair_t;

/* src/knn.rs:48: sort_unstable_by(total_cmp(dist).then(idx)). The key
 * (dist, idx) is a strict total order over distinct j, so any correct sort
 * gives the same sequence. */
static int pair_cmp(const void* pa, const void* pb) {
    const pair_t* a = (const pair_t*)pa;
    const pair_t* b = (const pair_t*)pb;
    int32_t ka = total_key(a->d), kb = total_key(b->d);
    if (ka != kb) return ka < kb ? -1 : 1;
    return a->j < b->j ? -1 : (a->j > b->j ? 1 : 0);
}

static void knn_row(const float* data, size_t n, size_t d, size_t k, int metric,
                    size_t i, pair_t* row, uint32_t* idx_out, float* dist_out) {
    const float* a = data + i * d;
    for (size_t j = 0; j < n; ++j) {
        float dj;
        if (j == i) {
            dj = 0.0f; /* knn.rs:38-40: self-distance is axiomatically zero */
        } else {
            const float* b = data + j * d;
            dj = metric == 0 ? hmo_euclidean(a, b, d) : hmo_cosine(a, b, d);
        }
        row[j].d = dj;
        row[j].j = (uint32_t)j;
    }
    qsort(row, n, sizeof(pair_t), pair_cmp);
    for (size_t t = 0; t < k; ++t) {
        idx_out[t] = row[t].j;
        dist_out[t] = row[t].d;
    }
}

/* src/knn.rs:22-55 */
void hmo_exact_knn(const float* data, size_t n, size_t d, size_t k, int metric,
                   uint32_t* indices, float* dists) {
    pair_t* row = (pair_t*)malloc(n * sizeof(pair_t));
    for (size_t i = 0; i < n; ++i)
        knn_row(data, n, d, k, metric, i, row, indices + i * k, dists + i * k);
    free(row);
}

void hmo_exact_knn_rows(const float* data, size_t n, size_t d, size_t k, int metric,
                        const uint32_t* rows, size_t nrows, uint32_t* indices,
                        float* dists, int nthreads) {
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        pair_t* row = (pair_t*)malloc(n * sizeof(pair_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (long r = 0; r < (long)nrows; ++r)
            knn_row(data, n, d, k, metric, rows[r], row, indices + (size_t)r * k,
                    dists + (size_t)r * k);
        free(row);
    }
}

/* libm 0.2.16 expf (src/fmath.rs:119-122), restated from the published
 * FreeBSD msun e_expf.c method the crate follows: reduction x = k*ln2 + r
 * with ln2 split hi/lo, the degree-2 Remez rational approximation
 * c = r - r^2 (P1 + r^2 P2), exp(r) = 1 + (r c/(2-c) - lo + hi), and an exact
 * power-of-two scaling (scalbnf).
 * The method and its constants are those of FreeBSD msun's e_expf.c, which
 * carries this notice:
 *   Conversion to float by Ian Lance Taylor, Cygnus Support, ian@cygnus.com.
 *   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
 *   Developed at SunPro, a Sun Microsystems, Inc. business.
 *   Permission to use, copy, modify, and distribute this software is freely
 *   granted, provided that this notice is preserved. */
static float scalbnf_(float x, int n) {
    float y = x;
    if (n > 127) {
        y *= 0x1p127f;
        n -= 127;
        if (n > 127) {
            y *= 0x1p127f;
            n -= 127;
            if (n > 127) n = 127;
        }
    } else if (n < -126) {
        y *= 0x1p-126f * 0x1p24f;
        n += 126 - 24;
        if (n < -126) {
            y *= 0x1p-126f * 0x1p24f;
            n += 126 - 24;
            if (n < -126) n = -126;
        }
    }
    return y * u2f((uint32_t)(0x7f + n) << 23);
}

float hmo_expf(float x) {
    static const float half[2] = {0.5f, -0.5f};
    const float ln2hi = 6.9314575195e-01f;  /* 0x3f317200 */
    const float ln2lo = 1.4286067653e-06f;  /* 0x35bfbe8e */
    const float invln2 = 1.4426950216e+00f; /* 0x3fb8aa3b */
    const float P1 = 1.6666625440e-1f;      /*  0xaaaa8f.0p-26 */
    const float P2 = -2.7667332906e-3f;     /* -0xb55215.0p-32 */
    uint32_t hx = f2u(x);
    int sign = (int)(hx >> 31);
    hx &= 0x7fffffffu;
    if (hx >= 0x42aeac50u) {            /* |x| >= 87.33655 or NaN */
        if (hx > 0x7f800000u) return x; /* NaN */
        if (hx >= 0x42b17218u && !sign) return x * 0x1p127f; /* overflow */
        if (sign && hx >= 0x42cff1b5u) return 0.0f;         /* x <= -103.972084 */
    }
    int k;
    float hi, lo;
    if (hx > 0x3eb17218u) {      /* |x| > 0.5 ln2 */
        if (hx > 0x3f851592u)    /* |x| > 1.5 ln2 */
            k = (int)(invln2 * x + half[sign]);
        else
            k = 1 - sign - sign;
        float kf = (float)k;
        hi = x - kf * ln2hi; /* exact */
        lo = kf * ln2lo;
        x = hi - lo;
    } else if (hx > 0x39000000u) { /* |x| > 2^-14 */
        k = 0;
        hi = x;
        lo = 0.0f;
    } else {
        return 1.0f + x;
    }
    float xx = x * x;
    float c = x - xx * (P1 + xx * P2);
    float y = 1.0f + (x * c / (2.0f - c) - lo + hi);
    return k == 0 ? y : scalbnf_(y, k);
}

/* src/smooth_knn.rs:82-84: sequential f64 sum, divided by the length, cast. */
static float mean_f32(const float* v, size_t len) {
    double s = -0.0;
    for (size_t t = 0; t < len; ++t) s = s + (double)v[t];
    return (float)(s / (double)len);
}

/* src/smooth_knn.rs:23-80 */
void hmo_smooth_knn(const float* knn_dists, size_t n, size_t k, float* sigmas,
                    float* rhos) {
    const double target = log2((double)k);
    const double tol = 1e-5;             /* SMOOTH_K_TOLERANCE */
    const float min_k_dist_scale = 1e-3f; /* MIN_K_DIST_SCALE */
    const float fmax = 3.40282347e+38f;  /* f32::MAX */
    const float mean_all = mean_f32(knn_dists, n * k);
    for (size_t i = 0; i < n; ++i) {
        const float* row = knn_dists + i * k;
        float rho = 0.0f;
        for (size_t t = 0; t < k; ++t)
            if (row[t] > 0.0f) { rho = row[t]; break; }
        rhos[i] = rho;
        float lo = 0.0f, hi = fmax, mid = 1.0f;
        for (int it = 0; it < 64; ++it) {
            float psum = 0.0f;
            for (size_t t = 1; t < k; ++t) {
                float dd = row[t] - rho;
                if (dd > 0.0f)
                    psum += hmo_expf(-(dd / mid));
                else
                    psum += 1.0f;
            }
            if (fabs((double)psum - target) < tol) break;
            if ((double)psum > target) {
                hi = mid;
                mid = (lo + hi) / 2.0f;
            } else {
                lo = mid;
                if (hi >= fmax)
                    mid *= 2.0f;
                else
                    mid = (lo + hi) / 2.0f;
            }
        }
        float sigma = mid;
        float floor_ = rho > 0.0f ? min_k_dist_scale * mean_f32(row, k)
                                  : min_k_dist_scale * mean_all;
        if (sigma < floor_) sigma = floor_;
        sigmas[i] = sigma;
    }
}

/* FP64 oracle: the same bracket search with double state and double exp,
 * iterated until the bracket stops moving (convergence), so sigma is the
 * root of sum_j exp(-(d_j - rho)/sigma) = log2(k) to double precision.
 * Floors as src/smooth_knn.rs:70-77, evaluated in double. On f64
 * distances (hmo_smooth_knn_f64d) or the f32 ones widened. */
void hmo_smooth_knn_f64d(const double* knn_dists, size_t n, size_t k, double* sigmas,
                         double* rhos) {
    const double target = log2((double)k);
    double mean_all = 0.0;
    for (size_t t = 0; t < n * k; ++t) mean_all += knn_dists[t];
    mean_all /= (double)(n * k);
    for (size_t i = 0; i < n; ++i) {
        const double* row = knn_dists + i * k;
        double rho = 0.0;
        for (size_t t = 0; t < k; ++t)
            if (row[t] > 0.0) { rho = row[t]; break; }
        rhos[i] = rho;
        double lo = 0.0, hi = INFINITY, mid = 1.0;
        for (int it = 0; it < 2200; ++it) {
            double psum = 0.0;
            for (size_t t = 1; t < k; ++t) {
                double dd = row[t] - rho;
                psum += dd > 0.0 ? exp(-(dd / mid)) : 1.0;
            }
            if (psum == target) break;
            double nmid;
            if (psum > target) {
                hi = mid;
                nmid = 0.5 * (lo + hi);
            } else {
                lo = mid;
                nmid = isinf(hi) ? mid * 2.0 : 0.5 * (lo + hi);
            }
            if (nmid == mid || isinf(nmid)) break;
            mid = nmid;
        }
        double rmean = 0.0;
        for (size_t t = 0; t < k; ++t) rmean += row[t];
        rmean /= (double)k;
        double floor_ = rho > 0.0 ? 1e-3 * rmean : 1e-3 * mean_all;
        sigmas[i] = mid < floor_ ? floor_ : mid;
    }
}

void hmo_smooth_knn_f64(const float* knn_dists, size_t n, size_t k, double* sigmas,
                        double* rhos) {
    double* w = (double*)malloc(n * k * sizeof(double) + 1);
    for (size_t t = 0; t < n * k; ++t) w[t] = (double)knn_dists[t];
    hmo_smooth_knn_f64d(w, n, k, sigmas, rhos);
    free(w);
}

typedef struct { uint32_t r, c; float v; } trip_t;
typedef struct { uint32_t r, c; } coord_t;

static int trip_cmp(const void* pa, const void* pb) {
    const trip_t* a = (const trip_t*)pa;
    const trip_t* b = (const trip_t*)pb;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    return a->c < b->c ? -1 : (a->c > b->c ? 1 : 0);
}
static int coord_cmp(const void* pa, const void* pb) {
    const coord_t* a = (const coord_t*)pa;
    const coord_t* b = (const coord_t*)pb;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    return a->c < b->c ? -1 : (a->c > b->c ? 1 : 0);
}

static float lookup(const trip_t* dir, size_t m, uint32_t r, uint32_t c) {
    size_t lo = 0, hi = m;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (dir[mid].r < r || (dir[mid].r == r && dir[mid].c < c))
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < m && dir[lo].r == r && dir[lo].c == c) return dir[lo].v;
    return 0.0f;
}

/* src/fuzzy.rs:50-84: the sorted directed triples' union with their
 * transpose, w = (v + vt) - v * vt in f32. */
static size_t fuzzy_union(trip_t* dir, size_t m, uint32_t* rows, uint32_t* cols, float* vals) {
    qsort(dir, m, sizeof(trip_t), trip_cmp);
    coord_t* co = (coord_t*)malloc(2 * m * sizeof(coord_t) + 1);
    for (size_t t = 0; t < m; ++t) {
        co[2 * t].r = dir[t].r;
        co[2 * t].c = dir[t].c;
        co[2 * t + 1].r = dir[t].c;
        co[2 * t + 1].c = dir[t].r;
    }
    qsort(co, 2 * m, sizeof(coord_t), coord_cmp);
    size_t nnz = 0;
    for (size_t t = 0; t < 2 * m; ++t) {
        if (t > 0 && co[t].r == co[t - 1].r && co[t].c == co[t - 1].c) continue;
        float v = lookup(dir, m, co[t].r, co[t].c);
        float vt = lookup(dir, m, co[t].c, co[t].r);
        float w = (v + vt) - v * vt;
        if (w != 0.0f) {
            rows[nnz] = co[t].r;
            cols[nnz] = co[t].c;
            vals[nnz] = w;
            ++nnz;
        }
    }
    free(co);
    return nnz;
}

/* src/fuzzy.rs:25-85 */
size_t hmo_fuzzy(const uint32_t* indices, const float* dists, const float* sigmas,
                 const float* rhos, size_t n, size_t k, uint32_t* rows,
                 uint32_t* cols, float* vals) {
    trip_t* dir = (trip_t*)malloc(n * k * sizeof(trip_t) + 1);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < k; ++j) {
            uint32_t col = indices[i * k + j];
            if ((size_t)col == i) continue;
            float dd = dists[i * k + j] - rhos[i];
            float val = (dd <= 0.0f || sigmas[i] == 0.0f) ? 1.0f
                                                         : hmo_expf(-(dd / sigmas[i]));
            if (val != 0.0f) {
                dir[m].r = (uint32_t)i;
                dir[m].c = col;
                dir[m].v = val;
                ++m;
            }
        }
    }
    const size_t nnz = fuzzy_union(dir, m, rows, cols, vals);
    free(dir);
    return nnz;
}

/* The FP64 chain's graph: directed strengths exp(-((d - rho) / sigma)) in
 * f64 from f64 distances, sigmas and rhos (1 where d - rho <= 0 or sigma ==
 * 0), rounded to f32 once (dropped when that is 0), then the same f32
 * union as src/fuzzy.rs:50-84. */
size_t hmo_fuzzy_f64(const uint32_t* indices, const double* dists, const double* sigmas,
                     const double* rhos, size_t n, size_t k, uint32_t* rows,
                     uint32_t* cols, float* vals) {
    trip_t* dir = (trip_t*)malloc(n * k * sizeof(trip_t) + 1);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < k; ++j) {
            uint32_t col = indices[i * k + j];
            if ((size_t)col == i) continue;
            double dd = dists[i * k + j] - rhos[i];
            float val = (float)((dd <= 0.0 || sigmas[i] == 0.0) ? 1.0 : exp(-(dd / sigmas[i])));
            if (val != 0.0f) {
                dir[m].r = (uint32_t)i;
                dir[m].c = col;
                dir[m].v = val;
                ++m;
            }
        }
    }
    const size_t nnz = fuzzy_union(dir, m, rows, cols, vals);
    free(dir);
    return nnz;
}

/* src/sparse.rs:21-35: row counts, then an inclusive running sum. */
void hmo_csr_indptr(const uint32_t* rows, size_t nnz, size_t n, uint64_t* indptr) {
    for (size_t i = 0; i <= n; ++i) indptr[i] = 0;
    for (size_t t = 0; t < nnz; ++t) indptr[rows[t] + 1] += 1;
    for (size_t i = 0; i < n; ++i) indptr[i + 1] += indptr[i];
}

/* src/components.rs:13-39. A search from every unlabelled vertex in
 * ascending order; the Vec the reference calls `queue` is used as a stack
 * (push, pop from the end). Labels are assigned when a vertex is pushed. */
size_t hmo_connected_components(const uint64_t* indptr, const uint32_t* indices, size_t n,
                                uint32_t* labels) {
    const uint32_t unlabeled = 0xffffffffu;
    for (size_t v = 0; v < n; ++v) labels[v] = unlabeled;
    size_t* stack = (size_t*)malloc((n + 1) * sizeof(size_t));
    uint32_t ncomp = 0;
    for (size_t start = 0; start < n; ++start) {
        if (labels[start] != unlabeled) continue;
        size_t top = 0;
        labels[start] = ncomp;
        stack[top++] = start;
        while (top > 0) {
            const size_t v = stack[--top];
            for (uint64_t p = indptr[v]; p < indptr[v + 1]; ++p) {
                const uint32_t u = indices[p];
                if (labels[u] == unlabeled) {
                    labels[u] = ncomp;
                    stack[top++] = u;
                }
            }
        }
        ncomp += 1;
    }
    free(stack);
    return ncomp;
}

/* src/sparse.rs:55-78. Rust's f64 Sum folds left to right from -0.0. */
void hmo_laplacian_new(const uint64_t* indptr, const uint32_t* indices, const float* vals,
                       size_t n, double* svals, double* degrees) {
    double* inv = (double*)malloc((n + 1) * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        double d = -0.0;
        for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) d = d + (double)vals[p];
        degrees[i] = d;
        inv[i] = 1.0 / sqrt(d);
    }
    for (size_t i = 0; i < n; ++i)
        for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p)
            svals[p] = (double)vals[p] * (inv[i] * inv[indices[p]]);
    free(inv);
}

/* src/sparse.rs:87-93. */
void hmo_laplacian_matvec(const uint64_t* indptr, const uint32_t* indices, const double* svals,
                          size_t n, const double* x, double* y) {
    for (size_t i = 0; i < n; ++i) {
        double sx = -0.0;
        for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) sx = sx + svals[p] * x[indices[p]];
        y[i] = x[i] - sx;
    }
}

/* src/sgd.rs:23-63: max by fold(0.0, f32::max) (the non-NaN operand wins),
 * cutoff in f32, edges kept unless w < cutoff, eps = f64(max) / f64(w). */
size_t hmo_schedule_edges(const uint32_t* rows, const uint32_t* cols, const float* vals,
                          size_t nnz, size_t n, size_t n_epochs, uint32_t* head, uint32_t* tail,
                          double* eps) {
    float max_w = 0.0f;
    for (size_t e = 0; e < nnz; ++e) max_w = fmaxf(max_w, vals[e]);
    const size_t prune = n_epochs > 10 ? n_epochs : (n <= 10000 ? 500 : 200);
    const float cutoff = max_w / (float)prune;
    size_t c = 0;
    for (size_t e = 0; e < nnz; ++e) {
        const float w = vals[e];
        if (w < cutoff) continue;
        head[c] = rows[e];
        tail[c] = cols[e];
        eps[c] = (double)max_w / (double)w;
        ++c;
    }
    return c;
}

/* crates/holomap-clusterer/src/pipeline.rs:151-162: norm = sqrt(sum of
 * x*x) with Rust's f32 Sum (a left fold from -0.0), zero rows copied, else
 * x / norm, all in f32. */
void hmo_l2_normalise(const float* x, size_t n, size_t d, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const float* r = x + i * d;
        float s = -0.0f;
        for (size_t f = 0; f < d; ++f) s = s + r[f] * r[f];
        const float norm = sqrtf(s);
        for (size_t f = 0; f < d; ++f) out[i * d + f] = norm == 0.0f ? r[f] : r[f] / norm;
    }
}
