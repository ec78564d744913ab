"""CPU oracle for the holomap kNN-graph path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only as the checker. The product (holomap_amd/)
never imports it.

The numerics live in ``hm_oracle.c`` (a line-by-line restatement of the
reference's src/metric.rs, src/knn.rs, src/smooth_knn.rs, src/fuzzy.rs and
src/sparse.rs with Rust float semantics); this module loads the compiled
library with ctypes and wraps it for numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libhm_oracle.so")
_lib = None

EUCLIDEAN = 0
COSINE = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t


def build() -> None:
    """Compile the oracle library (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.hmo_euclidean.argtypes = [_f32p, _f32p, _sz]
        L.hmo_euclidean.restype = ctypes.c_float
        L.hmo_cosine.argtypes = [_f32p, _f32p, _sz]
        L.hmo_cosine.restype = ctypes.c_float
        L.hmo_exact_knn.argtypes = [_f32p, _sz, _sz, _sz, ctypes.c_int, _u32p, _f32p]
        L.hmo_exact_knn.restype = None
        L.hmo_exact_knn_rows.argtypes = [_f32p, _sz, _sz, _sz, ctypes.c_int, _u32p, _sz,
                                         _u32p, _f32p, ctypes.c_int]
        L.hmo_exact_knn_rows.restype = None
        L.hmo_expf.argtypes = [ctypes.c_float]
        L.hmo_expf.restype = ctypes.c_float
        L.hmo_smooth_knn.argtypes = [_f32p, _sz, _sz, _f32p, _f32p]
        L.hmo_smooth_knn.restype = None
        L.hmo_smooth_knn_f64.argtypes = [_f32p, _sz, _sz, _f64p, _f64p]
        L.hmo_smooth_knn_f64.restype = None
        L.hmo_smooth_knn_f64d.argtypes = [_f64p, _sz, _sz, _f64p, _f64p]
        L.hmo_smooth_knn_f64d.restype = None
        L.hmo_pair_dist64.argtypes = [_f32p, _sz, _sz, _sz, ctypes.c_int, _u32p, _f64p]
        L.hmo_pair_dist64.restype = None
        L.hmo_fuzzy_f64.argtypes = [_u32p, _f64p, _f64p, _f64p, _sz, _sz, _u32p, _u32p, _f32p]
        L.hmo_fuzzy_f64.restype = _sz
        L.hmo_fuzzy.argtypes = [_u32p, _f32p, _f32p, _f32p, _sz, _sz, _u32p, _u32p, _f32p]
        L.hmo_fuzzy.restype = _sz
        L.hmo_csr_indptr.argtypes = [_u32p, _sz, _sz, _u64p]
        L.hmo_csr_indptr.restype = None
        L.hmo_connected_components.argtypes = [_u64p, _u32p, _sz, _u32p]
        L.hmo_connected_components.restype = _sz
        L.hmo_laplacian_new.argtypes = [_u64p, _u32p, _f32p, _sz, _f64p, _f64p]
        L.hmo_laplacian_new.restype = None
        L.hmo_laplacian_matvec.argtypes = [_u64p, _u32p, _f64p, _sz, _f64p, _f64p]
        L.hmo_laplacian_matvec.restype = None
        L.hmo_schedule_edges.argtypes = [_u32p, _u32p, _f32p, _sz, _sz, _sz, _u32p, _u32p, _f64p]
        L.hmo_schedule_edges.restype = _sz
        L.hmo_l2_normalise.argtypes = [_f32p, _sz, _sz, _f32p]
        L.hmo_l2_normalise.restype = None
        _lib = L
    return _lib


def _metric_code(metric) -> int:
    if isinstance(metric, str):
        return {"euclidean": EUCLIDEAN, "cosine": COSINE}[metric.lower()]
    return int(metric)


def euclidean(a, b) -> np.float32:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().hmo_euclidean(a, b, a.size))


def cosine(a, b) -> np.float32:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().hmo_cosine(a, b, a.size))


def expf(x: float) -> np.float32:
    return np.float32(lib().hmo_expf(float(np.float32(x))))


@dataclass
class Knn:
    k: int
    indices: np.ndarray  # uint32 [n, k]
    dists: np.ndarray  # float32 [n, k]


def exact_knn(data, n_features: int, k: int, metric="euclidean") -> Knn:
    """src/knn.rs:22-55."""
    x = np.ascontiguousarray(data, np.float32).reshape(-1)
    n = x.size // n_features
    idx = np.empty(n * k, np.uint32)
    dist = np.empty(n * k, np.float32)
    lib().hmo_exact_knn(x, n, n_features, k, _metric_code(metric), idx, dist)
    return Knn(k, idx.reshape(n, k), dist.reshape(n, k))


def exact_knn_rows(data, n_features: int, k: int, metric, rows, nthreads: int = 1) -> Knn:
    """exact_knn restricted to the query rows in ``rows`` (all N columns)."""
    x = np.ascontiguousarray(data, np.float32).reshape(-1)
    n = x.size // n_features
    rows = np.ascontiguousarray(rows, np.uint32)
    idx = np.empty(rows.size * k, np.uint32)
    dist = np.empty(rows.size * k, np.float32)
    lib().hmo_exact_knn_rows(x, n, n_features, k, _metric_code(metric), rows, rows.size,
                             idx, dist, int(nthreads))
    return Knn(k, idx.reshape(-1, k), dist.reshape(-1, k))


@dataclass
class SmoothKnn:
    sigmas: np.ndarray
    rhos: np.ndarray


def smooth_knn(knn_dists, k: int) -> SmoothKnn:
    """src/smooth_knn.rs:23-80 (float32 state, libm-method expf)."""
    d = np.ascontiguousarray(knn_dists, np.float32).reshape(-1)
    n = d.size // k
    s = np.empty(n, np.float32)
    r = np.empty(n, np.float32)
    lib().hmo_smooth_knn(d, n, k, s, r)
    return SmoothKnn(s, r)


def smooth_knn_f64(knn_dists, k: int) -> SmoothKnn:
    """FP64 bisection run to convergence (oracle for the FP64 GPU mode), on
    f64 distances when given float64, else on the f32 distances widened."""
    n = np.asarray(knn_dists).size // k
    s = np.empty(n, np.float64)
    r = np.empty(n, np.float64)
    if np.asarray(knn_dists).dtype == np.float64:
        lib().hmo_smooth_knn_f64d(np.ascontiguousarray(knn_dists, np.float64).reshape(-1), n, k,
                                  s, r)
    else:
        lib().hmo_smooth_knn_f64(np.ascontiguousarray(knn_dists, np.float32).reshape(-1), n, k,
                                 s, r)
    return SmoothKnn(s, r)


def pair_dist64(data, n_features: int, indices, metric="euclidean") -> np.ndarray:
    """The f64 distances (src/metric.rs before the f32 rounding) of row i to
    indices[i, t]; the row itself at 0."""
    x = np.ascontiguousarray(data, np.float32).reshape(-1)
    idx = np.ascontiguousarray(indices, np.uint32)
    n, k = idx.shape
    out = np.empty((n, k), np.float64)
    lib().hmo_pair_dist64(x, n, n_features, k, _metric_code(metric), idx.reshape(-1),
                          out.reshape(-1))
    return out


def fuzzy_simplicial_set_f64(indices, dists64, sigmas64, rhos64) -> FuzzyGraph:
    """The FP64 chain's graph: f64 strengths rounded once, the f32 union."""
    idx = np.ascontiguousarray(indices, np.uint32)
    n, k = idx.shape
    cap = 2 * n * k
    rows = np.empty(cap, np.uint32)
    cols = np.empty(cap, np.uint32)
    vals = np.empty(cap, np.float32)
    nnz = lib().hmo_fuzzy_f64(idx.reshape(-1), np.ascontiguousarray(dists64, np.float64).reshape(-1),
                              np.ascontiguousarray(sigmas64, np.float64),
                              np.ascontiguousarray(rhos64, np.float64), n, k, rows, cols, vals)
    return FuzzyGraph(n, rows[:nnz].copy(), cols[:nnz].copy(), vals[:nnz].copy())


@dataclass
class FuzzyGraph:
    n: int
    rows: np.ndarray
    cols: np.ndarray
    vals: np.ndarray


def fuzzy_simplicial_set(knn: Knn, calib: SmoothKnn) -> FuzzyGraph:
    """src/fuzzy.rs:25-85."""
    idx = np.ascontiguousarray(knn.indices, np.uint32).reshape(-1)
    dist = np.ascontiguousarray(knn.dists, np.float32).reshape(-1)
    k = knn.k
    n = idx.size // k
    cap = 2 * n * k
    rows = np.empty(cap, np.uint32)
    cols = np.empty(cap, np.uint32)
    vals = np.empty(cap, np.float32)
    nnz = lib().hmo_fuzzy(idx, dist, np.ascontiguousarray(calib.sigmas, np.float32),
                          np.ascontiguousarray(calib.rhos, np.float32), n, k, rows, cols, vals)
    return FuzzyGraph(n, rows[:nnz].copy(), cols[:nnz].copy(), vals[:nnz].copy())


def csr_indptr(rows, n: int) -> np.ndarray:
    """src/sparse.rs:21-35 (indptr only; indices = cols, vals widened)."""
    rows = np.ascontiguousarray(rows, np.uint32)
    out = np.empty(n + 1, np.uint64)
    lib().hmo_csr_indptr(rows, rows.size, n, out)
    return out


def knn_graph(data, n_features: int, k: int, metric="euclidean"):
    """The three-call seam of src/api.rs:147-149."""
    knn = exact_knn(data, n_features, k, metric)
    calib = smooth_knn(knn.dists, k)
    graph = fuzzy_simplicial_set(knn, calib)
    return knn, calib, graph


@dataclass
class Components:
    n_components: int
    labels: np.ndarray


def connected_components(indptr, indices, n: int) -> Components:
    """src/components.rs:13-39 on a CSR (indptr n+1, indices)."""
    ip = np.ascontiguousarray(indptr, np.uint64)
    ix = np.ascontiguousarray(indices, np.uint32)
    lab = np.empty(n, np.uint32)
    nc = lib().hmo_connected_components(ip, ix if ix.size else np.zeros(1, np.uint32), n, lab)
    return Components(int(nc), lab)


def laplacian_new(indptr, indices, vals, n: int):
    """src/sparse.rs:55-78: (S values, degrees)."""
    ip = np.ascontiguousarray(indptr, np.uint64)
    ix = np.ascontiguousarray(indices, np.uint32)
    v = np.ascontiguousarray(vals, np.float32)
    s = np.empty(max(v.size, 1), np.float64)
    d = np.empty(n, np.float64)
    lib().hmo_laplacian_new(ip, ix if ix.size else np.zeros(1, np.uint32),
                            v if v.size else np.zeros(1, np.float32), n, s, d)
    return s[:v.size], d


def laplacian_matvec(indptr, indices, svals, n: int, x) -> np.ndarray:
    """src/sparse.rs:87-93: y = x - S x."""
    ip = np.ascontiguousarray(indptr, np.uint64)
    ix = np.ascontiguousarray(indices, np.uint32)
    sv = np.ascontiguousarray(svals, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n, np.float64)
    lib().hmo_laplacian_matvec(ip, ix if ix.size else np.zeros(1, np.uint32),
                               sv if sv.size else np.zeros(1, np.float64), n, x, y)
    return y


@dataclass
class EdgeSchedule:
    head: np.ndarray
    tail: np.ndarray
    epochs_per_sample: np.ndarray


def schedule_edges(graph: FuzzyGraph, n_epochs: int) -> EdgeSchedule:
    """src/sgd.rs:36-63."""
    m = graph.rows.size
    h = np.empty(max(m, 1), np.uint32)
    t = np.empty(max(m, 1), np.uint32)
    e = np.empty(max(m, 1), np.float64)
    z32 = np.zeros(1, np.uint32)
    c = lib().hmo_schedule_edges(np.ascontiguousarray(graph.rows, np.uint32) if m else z32,
                                 np.ascontiguousarray(graph.cols, np.uint32) if m else z32,
                                 np.ascontiguousarray(graph.vals, np.float32) if m
                                 else np.zeros(1, np.float32), m, graph.n, n_epochs, h, t, e)
    return EdgeSchedule(h[:c].copy(), t[:c].copy(), e[:c].copy())


def l2_normalise(x) -> np.ndarray:
    """crates/holomap-clusterer/src/pipeline.rs:151-162."""
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(a)
    lib().hmo_l2_normalise(a.reshape(-1), a.shape[0], a.shape[1], out.reshape(-1))
    return out
