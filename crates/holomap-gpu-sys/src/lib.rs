//! The C ABI of libholomap_gpu (include/holomap_gpu.h), one to one.
//!
//! Drop-in for holomap's graph-construction seam (src/api.rs:147-149):
//! `hm_knn_graph` returns the reference's kNN indices and distances, sigma,
//! rho and the fuzzy graph bit for bit. Stage functions mirror
//! src/knn.rs:22, src/smooth_knn.rs:23, src/fuzzy.rs:25 and
//! src/sparse.rs:21 (for src/fixture_parity.rs:67,88,102); the graph's
//! consumers src/components.rs:13, src/sparse.rs:55-93 and src/sgd.rs:36.
//! The caller owns every host buffer; the library owns device memory,
//! streams and graphs. One context per thread.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

pub type hm_status = i32;
pub const HM_OK: hm_status = 0;
pub const HM_EBADSHAPE: hm_status = -1;
pub const HM_ETOOFEW: hm_status = -2;
pub const HM_EINVAL: hm_status = -3;
pub const HM_ENONFINITE: hm_status = -4;
pub const HM_EDEVICE: hm_status = -5;
pub const HM_ENOMEM: hm_status = -6;
pub const HM_EBUDGET: hm_status = -7;
pub const HM_EUNSUPPORTED: hm_status = -8;

pub const HM_MAX_NEIGHBORS: u32 = 1024;

pub const HM_EUCLIDEAN: i32 = 0;
pub const HM_COSINE: i32 = 1;

pub const HM_PREC_EXACT: i32 = 0;
pub const HM_PREC_FP32: i32 = 1;
pub const HM_PREC_BF16: i32 = 2;
pub const HM_PREC_BF16X3: i32 = 3;

pub const HM_CALIB_REF_F32: i32 = 0;
pub const HM_CALIB_F64: i32 = 1;

pub const HM_STAGE_KNN: u32 = 1;
pub const HM_STAGE_CALIB: u32 = 2;
pub const HM_STAGE_FUZZY: u32 = 4;
pub const HM_STAGE_ALL: u32 = 7;

pub const HM_NTIMES: usize = 8;
pub const HM_T_PREP: usize = 0;
pub const HM_T_SWEEP: usize = 1;
pub const HM_T_RERANK: usize = 2;
pub const HM_T_FALLBACK: usize = 3;
pub const HM_T_CALIB: usize = 4;
pub const HM_T_FUZZY: usize = 5;
pub const HM_T_TOTAL: usize = 6;
pub const HM_T_THRESHOLD: usize = 7;

pub const HM_NCOUNTERS: usize = 8;
pub const HM_C_FALLBACK_ROWS: usize = 0;
pub const HM_C_NNZ: usize = 1;
pub const HM_C_CANDIDATES: usize = 2;
pub const HM_C_NSPLIT: usize = 3;
pub const HM_C_ESCALATED: usize = 4;
pub const HM_C_LISTED: usize = 5;
pub const HM_C_TARGET: usize = 6;
pub const HM_C_SAMPLE: usize = 7;

pub const HM_NOALLOC_SCRATCH_BYTES: u64 = 256 << 20;

#[repr(C)]
pub struct hm_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct hm_job {
    _private: [u8; 0],
}
#[repr(C)]
pub struct hm_server {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct hm_options {
    pub precision: i32,
    pub calib: i32,
    pub margin: i32,
    pub nsplit: i32,
    pub use_graph: i32,
    pub autotune: i32,
    pub timing: i32,
    pub reserved: i32,
}

#[repr(C)]
#[derive(Debug)]
pub struct hm_graph_out {
    pub indices: *mut u32,
    pub dists: *mut f32,
    pub sigmas: *mut f32,
    pub rhos: *mut f32,
    pub rows: *mut u32,
    pub cols: *mut u32,
    pub vals: *mut f32,
    pub indptr: *mut u64,
    pub sigmas64: *mut f64,
    pub cap: u64,
    pub nnz: u64,
    pub dists64: *mut f64,
    pub err64: [f64; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct hm_server_config {
    pub max_jobs: u32,
    pub max_batch: u32,
    pub max_batch_rows: u64,
    pub opt: hm_options,
}

unsafe extern "C" {
    pub fn hm_default_options(opt: *mut hm_options);
    pub fn hm_version() -> *const c_char;
    pub fn hm_ctx_create(out: *mut *mut hm_ctx, device: c_int, hbm_budget_bytes: u64) -> hm_status;
    pub fn hm_ctx_destroy(ctx: *mut hm_ctx);
    pub fn hm_last_error(ctx: *const hm_ctx) -> *const c_char;
    pub fn hm_device_count() -> c_int;
    pub fn hm_ctx_budget(ctx: *const hm_ctx) -> u64;
    pub fn hm_ctx_in_use(ctx: *const hm_ctx) -> u64;
    pub fn hm_validate(data: *const f32, len: u64, n_features: u32, k: u32,
                       bad_index: *mut u64) -> hm_status;

    // stage-level entry points (host buffers)
    pub fn hm_exact_knn(ctx: *mut hm_ctx, data: *const f32, n: u64, n_features: u32, k: u32,
                        metric: i32, opt: *const hm_options, indices: *mut u32,
                        dists: *mut f32) -> hm_status;
    pub fn hm_smooth_knn(ctx: *mut hm_ctx, knn_dists: *const f32, n: u64, k: u32, calib: i32,
                         sigmas: *mut f32, rhos: *mut f32) -> hm_status;
    pub fn hm_fuzzy_simplicial_set(ctx: *mut hm_ctx, indices: *const u32, dists: *const f32,
                                   sigmas: *const f32, rhos: *const f32, n: u64, k: u32,
                                   rows: *mut u32, cols: *mut u32, vals: *mut f32, cap: u64,
                                   nnz: *mut u64) -> hm_status;
    pub fn hm_csr_from_fuzzy(ctx: *mut hm_ctx, rows: *const u32, nnz: u64, n: u64,
                             indptr: *mut u64) -> hm_status;

    // fused graph build
    pub fn hm_knn_graph(ctx: *mut hm_ctx, data: *const f32, n: u64, n_features: u32, k: u32,
                        metric: i32, opt: *const hm_options, out: *mut hm_graph_out) -> hm_status;

    // consumers of the graph's CSR
    pub fn hm_connected_components(ctx: *mut hm_ctx, indptr: *const u64, indices: *const u32,
                                   n: u64, labels: *mut u32, n_components: *mut u64) -> hm_status;
    pub fn hm_laplacian_new(ctx: *mut hm_ctx, indptr: *const u64, indices: *const u32,
                            vals: *const f32, n: u64, svals: *mut f64,
                            degrees: *mut f64) -> hm_status;
    pub fn hm_laplacian_matvec(ctx: *mut hm_ctx, indptr: *const u64, indices: *const u32,
                               svals: *const f64, n: u64, x: *const f64, y: *mut f64) -> hm_status;
    pub fn hm_job_connected_components(job: *mut hm_job, ds: u32, labels: *mut u32,
                                       n_components: *mut u64) -> hm_status;
    pub fn hm_job_laplacian_matvec(job: *mut hm_job, ds: u32, x: *const f64,
                                   y: *mut f64) -> hm_status;
    pub fn hm_schedule_edges(ctx: *mut hm_ctx, rows: *const u32, cols: *const u32,
                             vals: *const f32, nnz: u64, n: u64, n_epochs: u32, head: *mut u32,
                             tail: *mut u32, epochs_per_sample: *mut f64,
                             count: *mut u64) -> hm_status;
    pub fn hm_job_schedule_edges(job: *mut hm_job, ds: u32, n_epochs: u32, head: *mut u32,
                                 tail: *mut u32, epochs_per_sample: *mut f64,
                                 count: *mut u64) -> hm_status;

    // allocation-free graph build (device buffers)
    pub fn hm_knn_graph_noalloc(ctx: *mut hm_ctx, data: *const f32, n: u64, n_features: u32,
                                k: u32, metric: i32, opt: *const hm_options,
                                out: *mut hm_graph_out) -> hm_status;

    // device-resident jobs
    pub fn hm_job_create(ctx: *mut hm_ctx, nds: u32, n: *const u64, d: *const u32, k: *const u32,
                         metric: *const i32, opt: *const hm_options, stages: u32,
                         out: *mut *mut hm_job) -> hm_status;
    pub fn hm_job_destroy(job: *mut hm_job);
    pub fn hm_job_bytes(job: *const hm_job) -> u64;
    pub fn hm_job_estimate_bytes(nds: u32, n: *const u64, d: *const u32, k: *const u32,
                                 metric: *const i32, opt: *const hm_options) -> u64;
    pub fn hm_job_upload(job: *mut hm_job, ds: u32, data: *const f32) -> hm_status;
    pub fn hm_job_upload_device(job: *mut hm_job, ds: u32, dev_data: *const f32) -> hm_status;
    pub fn hm_job_set_knn(job: *mut hm_job, ds: u32, indices: *const u32,
                          dists: *const f32) -> hm_status;
    pub fn hm_job_set_calib(job: *mut hm_job, ds: u32, sigmas: *const f32,
                            rhos: *const f32) -> hm_status;
    pub fn hm_job_run(job: *mut hm_job) -> hm_status;
    pub fn hm_job_sync(job: *mut hm_job) -> hm_status;
    pub fn hm_job_done(job: *mut hm_job) -> c_int;
    pub fn hm_job_download(job: *mut hm_job, ds: u32, out: *mut hm_graph_out) -> hm_status;
    pub fn hm_job_times(job: *const hm_job, ms: *mut f32) -> hm_status;
    pub fn hm_job_counters(job: *mut hm_job, c: *mut u64) -> hm_status;
    pub fn hm_job_device_outputs(job: *mut hm_job, ds: u32, indices: *mut *const u32,
                                 dists: *mut *const f32, rows: *mut *const u32,
                                 cols: *mut *const u32, vals: *mut *const f32,
                                 indptr: *mut *const u64) -> hm_status;

    // serving loop
    pub fn hm_server_default_config(cfg: *mut hm_server_config);
    pub fn hm_server_create(ctx: *mut hm_ctx, cfg: *const hm_server_config,
                            out: *mut *mut hm_server) -> hm_status;
    pub fn hm_server_destroy(srv: *mut hm_server);
    pub fn hm_server_submit(srv: *mut hm_server, data: *const f32, n: u64, n_features: u32,
                            k: u32, metric: i32, id: *mut u64) -> hm_status;
    pub fn hm_server_progress(srv: *mut hm_server, completed: *mut u32) -> hm_status;
    pub fn hm_server_state(srv: *mut hm_server, id: u64) -> c_int;
    pub fn hm_server_pending(srv: *const hm_server) -> u32;
    pub fn hm_server_result_info(srv: *mut hm_server, id: u64, n: *mut u64, k: *mut u32,
                                 nnz: *mut u64) -> hm_status;
    pub fn hm_server_fetch(srv: *mut hm_server, id: u64, out: *mut hm_graph_out) -> hm_status;

    // the clusterer's f32 normalisation; the device expf (parity helper)
    pub fn hm_l2_normalise(ctx: *mut hm_ctx, data: *const f32, n: u64, d: u32,
                           out: *mut f32) -> hm_status;
    pub fn hm_expf_array(ctx: *mut hm_ctx, x: *const f32, y: *mut f32, n: u64) -> hm_status;
}
