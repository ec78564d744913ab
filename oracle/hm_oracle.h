/* CPU oracle for the holomap kNN-graph path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the GPU path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing in holomap_amd/.
 *
 * Every function restates the reference's arithmetic in the reference's
 * order (Rust float semantics: no contraction, sequential folds, IEEE
 * round-to-nearest-even), so its outputs are the reference's outputs:
 *   hmo_euclidean        src/metric.rs:42-53
 *   hmo_cosine           src/metric.rs:58-70
 *   hmo_exact_knn        src/knn.rs:22-55
 *   hmo_expf             src/fmath.rs:119-122 -> libm 0.2.16 expf
 *                        (FreeBSD msun e_expf.c method, restated).
 *                        Pinned to < 1 ulp against the correctly rounded
 *                        exp and through the golden sigmas; bit-level
 *                        agreement with libm 0.2.16 itself (a dependency
 *                        absent here) is parity unpinned.
 *   hmo_smooth_knn       src/smooth_knn.rs:12-84
 *   hmo_fuzzy            src/fuzzy.rs:25-85
 *   hmo_csr_indptr       src/sparse.rs:21-35
 *   hmo_connected_components  src/components.rs:13-39
 *   hmo_laplacian_new    src/sparse.rs:55-78
 *   hmo_laplacian_matvec src/sparse.rs:87-93
 *   hmo_schedule_edges   src/sgd.rs:23-63
 *   hmo_l2_normalise     crates/holomap-clusterer/src/pipeline.rs:151-162
 * hmo_smooth_knn_f64 / _f64d is an FP64 bisection run to convergence, with
 * hmo_pair_dist64 (the f64 values src/metric.rs rounds to f32) and
 * hmo_fuzzy_f64 (strengths in f64, rounded once, then the f32 union): the
 * FP64 chain, the oracle for the GPU's FP64 calibration mode (no reference
 * counterpart).
 */
#ifndef HM_ORACLE_H
#define HM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

float hmo_euclidean(const float* a, const float* b, size_t d);
float hmo_cosine(const float* a, const float* b, size_t d);
/* metric: 0 = euclidean, 1 = cosine. Full (dist, idx) sort per row, as knn.rs. */
void hmo_exact_knn(const float* data, size_t n, size_t d, size_t k, int metric,
                   uint32_t* indices, float* dists);
/* Same semantics for a subset of query rows (row ids in `rows`), output
 * row-major nrows x k. nthreads > 1 splits query rows over OpenMP threads
 * (rows are independent, so results are bit-identical to nthreads = 1). */
void hmo_exact_knn_rows(const float* data, size_t n, size_t d, size_t k, int metric,
                        const uint32_t* rows, size_t nrows, uint32_t* indices,
                        float* dists, int nthreads);
float hmo_expf(float x);
void hmo_smooth_knn(const float* knn_dists, size_t n, size_t k, float* sigmas,
                    float* rhos);
void hmo_smooth_knn_f64(const float* knn_dists, size_t n, size_t k, double* sigmas,
                        double* rhos);
void hmo_smooth_knn_f64d(const double* knn_dists, size_t n, size_t k, double* sigmas,
                         double* rhos);
/* f64 distances of pairs (row i, column idx[i*k+t]); the row itself at 0. */
void hmo_pair_dist64(const float* data, size_t n, size_t d, size_t k, int metric,
                     const uint32_t* idx, double* out);
/* rows/cols/vals capacity must be >= 2*n*k. Returns nnz. */
size_t hmo_fuzzy(const uint32_t* indices, const float* dists, const float* sigmas,
                 const float* rhos, size_t n, size_t k, uint32_t* rows,
                 uint32_t* cols, float* vals);
/* The FP64 chain's graph (capacity >= 2*n*k). Returns nnz. */
size_t hmo_fuzzy_f64(const uint32_t* indices, const double* dists, const double* sigmas,
                     const double* rhos, size_t n, size_t k, uint32_t* rows,
                     uint32_t* cols, float* vals);
void hmo_csr_indptr(const uint32_t* rows, size_t nnz, size_t n, uint64_t* indptr);
/* src/components.rs:13-39: returns the component count. */
size_t hmo_connected_components(const uint64_t* indptr, const uint32_t* indices, size_t n,
                                uint32_t* labels);
/* src/sparse.rs:55-78: S values (nnz) and degrees (n) from f32 weights. */
void hmo_laplacian_new(const uint64_t* indptr, const uint32_t* indices, const float* vals,
                       size_t n, double* svals, double* degrees);
/* src/sgd.rs:36-63: returns the kept count. */
size_t hmo_schedule_edges(const uint32_t* rows, const uint32_t* cols, const float* vals,
                          size_t nnz, size_t n, size_t n_epochs, uint32_t* head, uint32_t* tail,
                          double* eps);
/* crates/holomap-clusterer/src/pipeline.rs:151-162. */
void hmo_l2_normalise(const float* x, size_t n, size_t d, float* out);
/* src/sparse.rs:87-93: y = x - S x. */
void hmo_laplacian_matvec(const uint64_t* indptr, const uint32_t* indices, const double* svals,
                          size_t n, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
