//! The GPU form of holomap's graph-construction seam, a module of the
//! holomap crate itself (src/gpu.rs, behind a `gpu` cargo feature that
//! depends on holomap-gpu-sys). src/api.rs:147-149 then reads
//!
//!     #[cfg(feature = "gpu")]
//!     let (knn, calib, graph) = crate::gpu::knn_graph(data, n_features, self.n_neighbors, self.metric);
//!     #[cfg(not(feature = "gpu"))]
//!     let (knn, calib, graph) = { /* the three CPU calls, unchanged */ };
//!
//! and everything after it (src/api.rs:150-174) consumes the same values:
//! the library returns the reference's kNN indices and distances, sigma,
//! rho and fuzzy graph bit for bit. Validation stays where it is
//! (src/api.rs:113-144); HolomapError has no device variant, so a device
//! failure panics with the library's message rather than falling back.
use crate::{fuzzy::FuzzyGraph, knn::Knn, metric::Metric, smooth_knn::SmoothKnn};
use holomap_gpu_sys as sys;
use std::cell::RefCell;

/// One context per thread: `fit_transform(&self)` may run on several threads
/// at once and a context must not be shared by two.
struct Ctx(*mut sys::hm_ctx);

impl Drop for Ctx {
    fn drop(&mut self) {
        unsafe { sys::hm_ctx_destroy(self.0) }
    }
}

thread_local!(static CTX: RefCell<Option<Ctx>> = const { RefCell::new(None) });

fn fail(ctx: *mut sys::hm_ctx, st: sys::hm_status) -> ! {
    let msg = unsafe { std::ffi::CStr::from_ptr(sys::hm_last_error(ctx)) };
    panic!("holomap GPU graph build failed ({st}): {}", msg.to_string_lossy());
}

fn with_ctx<R>(f: impl FnOnce(*mut sys::hm_ctx) -> R) -> R {
    CTX.with(|c| {
        let mut c = c.borrow_mut();
        if c.is_none() {
            let mut p = std::ptr::null_mut();
            let st = unsafe { sys::hm_ctx_create(&mut p, 0, 0) };
            if st != sys::HM_OK {
                fail(p, st)
            }
            *c = Some(Ctx(p));
        }
        f(c.as_ref().unwrap().0)
    })
}

/// exact_knn -> smooth_knn -> fuzzy_simplicial_set (src/api.rs:147-149) in
/// one call: BF16 MFMA candidates re-ranked in the reference's f64
/// arithmetic, the reference's f32 calibration, the fuzzy union.
pub(crate) fn knn_graph(data: &[f32], n_features: usize, k: usize, metric: Metric)
    -> (Knn, SmoothKnn, FuzzyGraph) {
    let n = data.len() / n_features;
    let mut knn = Knn { k, indices: vec![0; n * k], dists: vec![0.0; n * k] };
    let mut calib = SmoothKnn { sigmas: vec![0.0; n], rhos: vec![0.0; n] };
    let cap = 2 * n * k;
    let (mut rows, mut cols, mut vals) = (vec![0u32; cap], vec![0u32; cap], vec![0f32; cap]);
    with_ctx(|ctx| {
        let mut opt = sys::hm_options::default();
        unsafe { sys::hm_default_options(&mut opt) }; // BF16 candidates, ref_f32 calibration
        let mut out = sys::hm_graph_out {
            indices: knn.indices.as_mut_ptr(),
            dists: knn.dists.as_mut_ptr(),
            sigmas: calib.sigmas.as_mut_ptr(),
            rhos: calib.rhos.as_mut_ptr(),
            rows: rows.as_mut_ptr(),
            cols: cols.as_mut_ptr(),
            vals: vals.as_mut_ptr(),
            indptr: std::ptr::null_mut(),
            sigmas64: std::ptr::null_mut(),
            cap: cap as u64,
            nnz: 0,
            dists64: std::ptr::null_mut(),
            err64: [0.0; 3],
        };
        let m = match metric {
            Metric::Euclidean => sys::HM_EUCLIDEAN,
            Metric::Cosine => sys::HM_COSINE,
        };
        let st = unsafe {
            sys::hm_knn_graph(ctx, data.as_ptr(), n as u64, n_features as u32, k as u32, m, &opt,
                              &mut out)
        };
        if st != sys::HM_OK {
            fail(ctx, st)
        }
        let nnz = out.nnz as usize;
        rows.truncate(nnz);
        cols.truncate(nnz);
        vals.truncate(nnz);
    });
    (knn, calib, FuzzyGraph { n, rows, cols, vals })
}
