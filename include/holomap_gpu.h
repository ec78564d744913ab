/* holomap_gpu.h -- C ABI of the MI355X (gfx950) kNN-graph library.
 *
 * Drop-in for the graph-construction seam of holomap's
 * Holomap::fit_transform (src/api.rs:147-149), which calls three
 * crate-private stage functions:
 *
 *   exact_knn(&[f32], n_features, k, Metric) -> Knn            src/knn.rs:22
 *   smooth_knn(&[f32] knn_dists, k) -> SmoothKnn                src/smooth_knn.rs:23
 *   fuzzy_simplicial_set(&Knn, &SmoothKnn) -> FuzzyGraph        src/fuzzy.rs:25
 *
 * and whose first consumer builds Csr::from_fuzzy (src/sparse.rs:21).
 * Each stage has a host-buffer entry point below with the reference's
 * argument meaning and output layout (row-major n x k; COO sorted by
 * (row, col)); hm_knn_graph fuses the three; the hm_job_* functions keep
 * inputs and the graph resident in device memory for batching and serving.
 *
 * Plain C types only. The caller owns every host buffer; the library owns
 * device memory, streams and graphs. Every function is re-entrant across
 * contexts; one context must not be used by two threads at once.
 */
#ifndef HOLOMAP_GPU_H
#define HOLOMAP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hm_status;
/* Status codes. HM_EBADSHAPE .. HM_ENONFINITE mirror HolomapError's
 * variants (src/api.rs:17-44): BadShape, TooFewPoints, InvalidParameter,
 * NonFiniteInput. */
#define HM_OK 0
#define HM_EBADSHAPE -1
#define HM_ETOOFEW -2
#define HM_EINVAL -3
#define HM_ENONFINITE -4
#define HM_EDEVICE -5    /* HIP runtime / kernel failure */
#define HM_ENOMEM -6     /* device allocation failed */
#define HM_EBUDGET -7    /* request exceeds the context's HBM budget */
#define HM_EUNSUPPORTED -8

/* Largest n_neighbors the GPU path serves (the reference has no limit
 * below n; larger values return HM_EUNSUPPORTED). */
#define HM_MAX_NEIGHBORS 1024

/* Metric (src/metric.rs:8-14). */
#define HM_EUCLIDEAN 0
#define HM_COSINE 1

/* Candidate-selection precision of the kNN stage. Every mode returns the
 * reference's indices and distances bit for bit: candidates are re-ranked
 * in the reference's f64 arithmetic and rows whose candidate set cannot be
 * certified are recomputed exactly. */
#define HM_PREC_EXACT 0  /* f64 reference arithmetic for every pair */
#define HM_PREC_FP32 1   /* FP32 MFMA Gram candidates */
#define HM_PREC_BF16 2   /* BF16 MFMA Gram candidates (default) */
#define HM_PREC_BF16X3 3 /* BF16 hi/lo split, three MFMAs per step */

/* Calibration (src/smooth_knn.rs). REF_F32 reproduces the reference's f32
 * bisection bit for bit; F64 runs the search in f64 to convergence. */
#define HM_CALIB_REF_F32 0
#define HM_CALIB_F64 1

/* Stages of a job. */
#define HM_STAGE_KNN 1u
#define HM_STAGE_CALIB 2u
#define HM_STAGE_FUZZY 4u
#define HM_STAGE_ALL 7u

typedef struct hm_ctx hm_ctx;
typedef struct hm_job hm_job;

typedef struct {
  int32_t precision;  /* HM_PREC_*, default HM_PREC_BF16 */
  int32_t calib;      /* HM_CALIB_*, default HM_CALIB_REF_F32 */
  int32_t margin;     /* candidates kept per row = k + margin; -1: 2k */
  int32_t nsplit;     /* column splits of the candidate sweep; 0: auto */
  int32_t use_graph;  /* 1: capture the pipeline once, replay after */
  int32_t autotune;   /* 1: time the sweep configurations on first run */
  int32_t timing;     /* 1: record per-stage device times (eager runs only) */
  int32_t reserved;
} hm_options;

/* Per-stage device times of the last hm_job_run with timing enabled (ms). */
#define HM_NTIMES 8
#define HM_T_PREP 0
#define HM_T_SWEEP 1
#define HM_T_RERANK 2
#define HM_T_FALLBACK 3
#define HM_T_CALIB 4
#define HM_T_FUZZY 5
#define HM_T_TOTAL 6
#define HM_T_THRESHOLD 7 /* the sweep's threshold passes, part of HM_T_SWEEP */

/* Counters of the last run (hm_job_counters). */
#define HM_NCOUNTERS 8
#define HM_C_FALLBACK_ROWS 0 /* rows recomputed by the exact scan */
#define HM_C_NNZ 1           /* graph entries, all datasets */
#define HM_C_CANDIDATES 2    /* candidate key slots per row (lists x list capacity) */
#define HM_C_NSPLIT 3        /* column splits used by the sweep (dataset 0) */
#define HM_C_ESCALATED 4     /* rows the primary keys left uncertified, swept again */
#define HM_C_LISTED 5        /* candidate keys listed for dataset 0's rows */
#define HM_C_TARGET 6        /* candidates targeted per row, k + margin (dataset 0) */
#define HM_C_SAMPLE 7        /* block stride of the threshold pass's sample (dataset 0) */

void hm_default_options(hm_options* opt);
const char* hm_version(void);

/* Context: one device, one stream, an HBM budget for hm_job admission
 * (0 = no limit). */
hm_status hm_ctx_create(hm_ctx** out, int device, uint64_t hbm_budget_bytes);
void hm_ctx_destroy(hm_ctx* ctx);
const char* hm_last_error(const hm_ctx* ctx);
int hm_device_count(void);
/* HBM budget of the context (0 = none) and bytes held by its live jobs. */
uint64_t hm_ctx_budget(const hm_ctx* ctx);
uint64_t hm_ctx_in_use(const hm_ctx* ctx);

/* Input validation of src/api.rs:113-144 restricted to what the graph
 * stages see (shape, n_neighbors >= 2, n > n_neighbors, finite data).
 * *bad_index receives the first non-finite index for HM_ENONFINITE. */
hm_status hm_validate(const float* data, uint64_t len, uint32_t n_features, uint32_t k,
                      uint64_t* bad_index);

/* ---- stage-level entry points (host buffers) ---- */

/* src/knn.rs:22-55. indices/dists: n*k, row-major. */
hm_status hm_exact_knn(hm_ctx* ctx, const float* data, uint64_t n, uint32_t n_features,
                       uint32_t k, int32_t metric, const hm_options* opt, uint32_t* indices,
                       float* dists);

/* src/smooth_knn.rs:23-80. knn_dists: n*k sorted rows; sigmas/rhos: n. */
hm_status hm_smooth_knn(hm_ctx* ctx, const float* knn_dists, uint64_t n, uint32_t k,
                        int32_t calib, float* sigmas, float* rhos);

/* src/fuzzy.rs:25-85. rows/cols/vals capacity `cap` (2*n*k always
 * suffices); *nnz receives the entry count (HM_EINVAL if cap is short). */
hm_status hm_fuzzy_simplicial_set(hm_ctx* ctx, const uint32_t* indices, const float* dists,
                                  const float* sigmas, const float* rhos, uint64_t n, uint32_t k,
                                  uint32_t* rows, uint32_t* cols, float* vals, uint64_t cap,
                                  uint64_t* nnz);

/* src/sparse.rs:21-35: indptr (n+1) of the (row, col)-sorted COO. */
hm_status hm_csr_from_fuzzy(hm_ctx* ctx, const uint32_t* rows, uint64_t nnz, uint64_t n,
                            uint64_t* indptr);

/* ---- fused graph build (host buffers) ---- */
typedef struct {
  uint32_t* indices; /* n*k or NULL */
  float* dists;      /* n*k or NULL */
  float* sigmas;     /* n or NULL */
  float* rhos;       /* n or NULL */
  uint32_t* rows;    /* cap or NULL */
  uint32_t* cols;    /* cap or NULL */
  float* vals;       /* cap or NULL */
  uint64_t* indptr;  /* n+1 or NULL */
  double* sigmas64;  /* n or NULL: FP64 calibration values (HM_CALIB_F64) */
  uint64_t cap;      /* capacity of rows/cols/vals */
  uint64_t nnz;      /* out */
  double* dists64;   /* n*k or NULL: the f64 distances the f32 ones round
                        (HM_CALIB_F64 with the kNN on the device) */
  double err64[3];   /* out: with HM_CALIB_F64 and the kNN on the device,
                        the largest relative error of the returned f32
                        distances, sigmas and directed membership strengths
                        against their f64 values (the FP64 chain); -1 each
                        otherwise (and from hm_knn_graph_noalloc) */
} hm_graph_out;

/* src/api.rs:147-149 fused: exact_knn -> smooth_knn -> fuzzy_simplicial_set. */
hm_status hm_knn_graph(hm_ctx* ctx, const float* data, uint64_t n, uint32_t n_features,
                       uint32_t k, int32_t metric, const hm_options* opt, hm_graph_out* out);

/* ---- consumers of the graph's CSR: src/components.rs, src/sparse.rs ----
 * For graphs fuzzy_simplicial_set returns (symmetric, columns ascending in
 * each row), given as CSR: indptr (n+1, from 0), indices (indptr[n]).
 * Every output is the reference's bit for bit (its f64 folds in its order). */

/* src/components.rs:13-39: labels (n) in first-seen order (the component
 * of vertex 0 is 0, ...), *n_components the count. */
hm_status hm_connected_components(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                                  uint64_t n, uint32_t* labels, uint64_t* n_components);
/* src/sparse.rs:55-78 (Laplacian::new): from the graph's f32 weights, the
 * normalised adjacency S = D^-1/2 W D^-1/2 on the same pattern (svals,
 * indptr[n] values) and the degrees (n); either output may be NULL. */
hm_status hm_laplacian_new(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                           const float* vals, uint64_t n, double* svals, double* degrees);
/* src/sparse.rs:87-93 (Laplacian::matvec): y = x - S x (n values each). */
hm_status hm_laplacian_matvec(hm_ctx* ctx, const uint64_t* indptr, const uint32_t* indices,
                              const double* svals, uint64_t n, const double* x, double* y);
/* The same on a job's resident graph of dataset ds (no host copy of the
 * graph): labels to host memory; the Laplacian product with x and y in
 * DEVICE memory (n doubles each), for iterative eigensolvers. */
hm_status hm_job_connected_components(hm_job* job, uint32_t ds, uint32_t* labels,
                                      uint64_t* n_components);
hm_status hm_job_laplacian_matvec(hm_job* job, uint32_t ds, const double* x, double* y);

/* src/sgd.rs:36-63 (schedule_edges): the SGD layout's edge list from the
 * graph's COO (rows, cols, vals: nnz; n vertices): edges with weight below
 * max(vals) / prune_epochs dropped (prune_epochs = n_epochs if > 10, else
 * 500 for n <= 10,000 and 200 above), in COO order; head, tail,
 * epochs_per_sample = max / w (capacity nnz each); *count kept. */
hm_status hm_schedule_edges(hm_ctx* ctx, const uint32_t* rows, const uint32_t* cols,
                            const float* vals, uint64_t nnz, uint64_t n, uint32_t n_epochs,
                            uint32_t* head, uint32_t* tail, double* epochs_per_sample,
                            uint64_t* count);
/* The same on a job's resident graph of dataset ds (host outputs, capacity
 * the dataset's nnz). */
hm_status hm_job_schedule_edges(hm_job* job, uint32_t ds, uint32_t n_epochs, uint32_t* head,
                                uint32_t* tail, double* epochs_per_sample, uint64_t* count);

/* ---- allocation-free graph build ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, serving loops and harnesses that
 * forbid hipMalloc): hm_knn_graph's computation with the input `data`
 * (n * n_features floats) and every non-NULL field of `out` in the
 * caller's DEVICE memory, and every intermediate buffer carved from a fixed
 * scratch array of HM_NOALLOC_SCRATCH_BYTES that each GPU the library runs
 * on holds until the process ends. It calls no hipMalloc or hipFree.
 * Inputs whose graph build needs more than the array (about N = 30,000
 * rows at d = 128, k = 15) fail with HM_EBUDGET. Calls on one GPU share its
 * array: overlapping calls would race on it, so the library runs them one
 * at a time (a lock per device), each returning when its device work and
 * output copies are done; out->nnz is written on the host. Runs on the
 * context's stream; options use_graph, autotune and timing are ignored. */
#define HM_NOALLOC_SCRATCH_BYTES (256ull << 20)
hm_status hm_knn_graph_noalloc(hm_ctx* ctx, const float* data, uint64_t n, uint32_t n_features,
                               uint32_t k, int32_t metric, const hm_options* opt,
                               hm_graph_out* out);

/* ---- device-resident jobs: one or many datasets per launch ---- */

/* Plans a batch of `nds` datasets (shapes n[i] x d[i], k[i], metric[i]).
 * Allocates every device buffer up front (fails with HM_EBUDGET if the
 * context budget would be exceeded). `stages` is a mask of HM_STAGE_*. */
hm_status hm_job_create(hm_ctx* ctx, uint32_t nds, const uint64_t* n, const uint32_t* d,
                        const uint32_t* k, const int32_t* metric, const hm_options* opt,
                        uint32_t stages, hm_job** out);
void hm_job_destroy(hm_job* job);
/* Device bytes held by the job. */
uint64_t hm_job_bytes(const hm_job* job);
/* Estimate without allocating (admission control); metric may be NULL
 * (every dataset counted as cosine, the larger of the two). */
uint64_t hm_job_estimate_bytes(uint32_t nds, const uint64_t* n, const uint32_t* d,
                               const uint32_t* k, const int32_t* metric, const hm_options* opt);
/* Host -> device copy of dataset `ds` (n*d floats). */
hm_status hm_job_upload(hm_job* job, uint32_t ds, const float* data);
/* Device -> device copy from a caller device pointer (n*d floats). */
hm_status hm_job_upload_device(hm_job* job, uint32_t ds, const float* dev_data);
/* Stage-level inputs for jobs without HM_STAGE_KNN / HM_STAGE_CALIB. */
hm_status hm_job_set_knn(hm_job* job, uint32_t ds, const uint32_t* indices, const float* dists);
hm_status hm_job_set_calib(hm_job* job, uint32_t ds, const float* sigmas, const float* rhos);
/* Enqueue the pipeline on the job's stream (asynchronous). */
hm_status hm_job_run(hm_job* job);
hm_status hm_job_sync(hm_job* job);
/* Non-blocking completion test: 1 done, 0 running, <0 error. */
int hm_job_done(hm_job* job);
/* Device -> host copy of dataset `ds` results (any NULL field skipped). */
hm_status hm_job_download(hm_job* job, uint32_t ds, hm_graph_out* out);
hm_status hm_job_times(const hm_job* job, float* ms /* HM_NTIMES */);
hm_status hm_job_counters(hm_job* job, uint64_t* c /* HM_NCOUNTERS */);
/* Device pointers of dataset `ds` results (valid until destroy). */
hm_status hm_job_device_outputs(hm_job* job, uint32_t ds, const uint32_t** indices,
                                const float** dists, const uint32_t** rows, const uint32_t** cols,
                                const float** vals, const uint64_t** indptr);

/* ---- serving loop: fit requests admitted into concurrent jobs ----
 * No counterpart in the reference (it has no server for this path; its
 * JSON-lines clusterer, crates/holomap-clusterer/src/main.rs:10-31, calls
 * fit_transform once per request). Requests queue in arrival order;
 * hm_server_progress retires finished jobs and admits the longest queue
 * prefix that fits max_batch, max_batch_rows and the context's free HBM
 * budget into a new job (continuous batching), up to max_jobs jobs running
 * at once on their own streams. Results stay in device memory until
 * fetched. A request larger than the whole budget ends in HM_EBUDGET.
 * Single-threaded: one thread drives a server. */
typedef struct hm_server hm_server;
typedef struct {
  uint32_t max_jobs;       /* jobs running at once; default 4 */
  uint32_t max_batch;      /* requests per job; default 32 */
  uint64_t max_batch_rows; /* rows per job (sum of n); default 2^20 */
  hm_options opt;          /* per-job options (use_graph/timing forced off) */
} hm_server_config;
void hm_server_default_config(hm_server_config* cfg);
hm_status hm_server_create(hm_ctx* ctx, const hm_server_config* cfg, hm_server** out);
void hm_server_destroy(hm_server* srv);
/* Queue one request (validated as hm_validate; the data is copied). */
hm_status hm_server_submit(hm_server* srv, const float* data, uint64_t n, uint32_t n_features,
                           uint32_t k, int32_t metric, uint64_t* id);
/* Non-blocking step; *completed = requests finished (or failed) in this call. */
hm_status hm_server_progress(hm_server* srv, uint32_t* completed);
/* 0 queued, 1 running, 2 done, < 0: the request's error status. */
int hm_server_state(hm_server* srv, uint64_t id);
/* Requests queued or running. */
uint32_t hm_server_pending(const hm_server* srv);
/* Shape of a done request's result (for sizing hm_graph_out). */
hm_status hm_server_result_info(hm_server* srv, uint64_t id, uint64_t* n, uint32_t* k,
                                uint64_t* nnz);
/* Copy a done request's results to the host (out may be NULL) and release
 * the request; a failed request returns its error status and is released. */
hm_status hm_server_fetch(hm_server* srv, uint64_t id, hm_graph_out* out);

/* crates/holomap-clusterer/src/pipeline.rs:151-162 (l2_normalise): every
 * row of n x d divided by its f32 norm (zero rows copied); host buffers. */
hm_status hm_l2_normalise(hm_ctx* ctx, const float* data, uint64_t n, uint32_t d, float* out);

/* Parity helper: the device expf (the reference's libm method) over an
 * array, for testing the calibration arithmetic. */
hm_status hm_expf_array(hm_ctx* ctx, const float* x, float* y, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
