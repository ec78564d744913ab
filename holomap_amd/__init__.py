"""holomap_amd -- MI355X (gfx950) kNN-graph construction for holomap.

Python mirror of the reference's graph-construction seam. The names,
argument meaning, output layout and error behaviour follow the Rust crate:

  Metric                      src/metric.rs:8-39
  exact_knn -> Knn            src/knn.rs:13-55
  smooth_knn -> SmoothKnn     src/smooth_knn.rs:16-80
  fuzzy_simplicial_set        src/fuzzy.rs:17-85
  Csr.from_fuzzy / row        src/sparse.rs:12-41
  connected_components        src/components.rs:7-39
  Laplacian.new / matvec      src/sparse.rs:46-93
  EdgeSchedule, schedule_edges, default_n_epochs   src/sgd.rs:23-63
  HolomapError variants       src/api.rs:17-44
  Holomap.builder(seed)       src/api.rs:94-107, setters :195-235, and
  Holomap.graph(data, n_features) = the validation of src/api.rs:113-144
  followed by the three calls of src/api.rs:147-149.

Every computation runs in libholomap_gpu.so (hand-written HIP kernels for
gfx950) through its C ABI; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import Enum
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (HM_CALIB_F64, HM_CALIB_REF_F32, HM_PREC_BF16, HM_PREC_BF16X3,
                   HM_PREC_EXACT, HM_PREC_FP32, HM_STAGE_ALL, HM_STAGE_CALIB, HM_STAGE_FUZZY,
                   HM_STAGE_KNN, TIME_NAMES)

__all__ = [
    "Metric", "Knn", "SmoothKnn", "FuzzyGraph", "Csr", "HolomapError", "BadShape",
    "TooFewPoints", "InvalidParameter", "NonFiniteInput", "DeviceError", "Context", "Job",
    "exact_knn", "smooth_knn", "fuzzy_simplicial_set", "knn_graph", "Holomap",
    "Components", "connected_components", "Laplacian", "EdgeSchedule", "schedule_edges",
    "default_n_epochs",
    "PRECISIONS", "CALIBRATIONS",
]

PRECISIONS = {"exact": HM_PREC_EXACT, "fp32": HM_PREC_FP32, "bf16": HM_PREC_BF16,
              "bf16x3": HM_PREC_BF16X3}
CALIBRATIONS = {"ref_f32": HM_CALIB_REF_F32, "f64": HM_CALIB_F64}


# ---------------------------------------------------------------- errors
class HolomapError(Exception):
    """src/api.rs:17-44."""


class BadShape(HolomapError):
    def __init__(self, length: int, n_features: int):
        self.len = length
        self.n_features = n_features
        super().__init__(f"data length {length} is not a positive multiple of n_features {n_features}")


class TooFewPoints(HolomapError):
    def __init__(self, n: int, n_neighbors: int):
        self.n = n
        self.n_neighbors = n_neighbors
        super().__init__(f"need more than n_neighbors={n_neighbors} points, got {n}")


class InvalidParameter(HolomapError):
    def __init__(self, what: str):
        self.what = what
        super().__init__(f"invalid parameter: {what}")


class NonFiniteInput(HolomapError):
    def __init__(self, index: int):
        self.index = index
        super().__init__(f"data contains a non-finite value at index {index}")


class DeviceError(RuntimeError):
    """A HIP runtime or allocation failure (no reference counterpart)."""


# ---------------------------------------------------------------- metric
class Metric(Enum):
    EUCLIDEAN = 0
    COSINE = 1

    @classmethod
    def from_str(cls, s: str) -> "Metric":
        """Case-insensitive parse (src/metric.rs:16-30)."""
        t = s.lower()
        if t == "euclidean":
            return cls.EUCLIDEAN
        if t == "cosine":
            return cls.COSINE
        raise InvalidParameter('metric must be "euclidean" or "cosine"')

    def __str__(self) -> str:
        return "euclidean" if self is Metric.EUCLIDEAN else "cosine"


def _metric(m) -> int:
    if isinstance(m, Metric):
        return m.value
    if isinstance(m, str):
        return Metric.from_str(m).value
    return int(m)


# ---------------------------------------------------------------- results
@dataclass
class Knn:
    """Row-major n x k (src/knn.rs:13-19)."""
    k: int
    indices: np.ndarray  # uint32 (n, k)
    dists: np.ndarray  # float32 (n, k)
    dists64: Optional[np.ndarray] = None  # float64 (n, k): FP64 calibration with the kNN here


@dataclass
class SmoothKnn:
    """src/smooth_knn.rs:16-19 (sigmas64: FP64 calibration values, if run)."""
    sigmas: np.ndarray
    rhos: np.ndarray
    sigmas64: Optional[np.ndarray] = None


@dataclass
class FuzzyGraph:
    """COO sorted by (row, col) (src/fuzzy.rs:17-22)."""
    n: int
    rows: np.ndarray
    cols: np.ndarray
    vals: np.ndarray


@dataclass
class Csr:
    """src/sparse.rs:12-17; vals widened to f64."""
    n: int
    indptr: np.ndarray
    indices: np.ndarray
    vals: np.ndarray

    @staticmethod
    def from_fuzzy(g: FuzzyGraph, ctx: "Context" = None) -> "Csr":
        """src/sparse.rs:21-35 (row pointer computed on the device)."""
        c = ctx or Context.default()
        rows = np.ascontiguousarray(g.rows, np.uint32)
        indptr = np.empty(g.n + 1, np.uint64)
        c._check(_lib.load().hm_csr_from_fuzzy(c.handle, rows.ctypes.data, rows.size, g.n,
                                               indptr.ctypes.data))
        return Csr(g.n, indptr, g.cols.copy(), g.vals.astype(np.float64))

    def row(self, i: int):
        lo, hi = int(self.indptr[i]), int(self.indptr[i + 1])
        return self.indices[lo:hi], self.vals[lo:hi]


@dataclass
class Components:
    """src/components.rs:7-11."""
    n_components: int
    labels: np.ndarray  # uint32 (n,)


def _csr_arrays(csr: Csr):
    ip = np.ascontiguousarray(csr.indptr, np.uint64)
    ix = np.ascontiguousarray(csr.indices, np.uint32)
    return ip, ix


def connected_components(csr: Csr, ctx: "Context" = None) -> Components:
    """src/components.rs:13-39 on the device: labels in first-seen order
    (for the symmetric graphs fuzzy_simplicial_set returns)."""
    c = ctx or Context.default()
    ip, ix = _csr_arrays(csr)
    lab = np.empty(csr.n, np.uint32)
    nc = ctypes.c_uint64()
    c._check(_lib.load().hm_connected_components(c.handle, ip.ctypes.data, ix.ctypes.data, csr.n,
                                                 lab.ctypes.data, ctypes.byref(nc)))
    return Components(int(nc.value), lab)


class Laplacian:
    """src/sparse.rs:46-99: the symmetric-normalised Laplacian, stored as
    the normalised adjacency S = D^-1/2 W D^-1/2 (a Csr of f64 values) and
    the degrees; built and applied on the device."""

    def __init__(self, s: Csr, degrees: np.ndarray):
        self.s = s
        self.degrees = degrees

    @staticmethod
    def new(g: FuzzyGraph, ctx: "Context" = None) -> "Laplacian":
        """src/sparse.rs:55-78."""
        c = ctx or Context.default()
        csr = Csr.from_fuzzy(g, ctx=c)
        ip, ix = _csr_arrays(csr)
        w = np.ascontiguousarray(g.vals, np.float32)
        sv = np.empty(max(w.size, 1), np.float64)
        deg = np.empty(g.n, np.float64)
        c._check(_lib.load().hm_laplacian_new(c.handle, ip.ctypes.data, ix.ctypes.data,
                                              w.ctypes.data, g.n, sv.ctypes.data,
                                              deg.ctypes.data))
        return Laplacian(Csr(g.n, csr.indptr, csr.indices, sv[:w.size]), deg)

    def n(self) -> int:
        return self.s.n

    def matvec(self, x, ctx: "Context" = None) -> np.ndarray:
        """src/sparse.rs:87-93: y = L x = x - S x."""
        c = ctx or Context.default()
        ip, ix = _csr_arrays(self.s)
        sv = np.ascontiguousarray(self.s.vals, np.float64)
        xx = np.ascontiguousarray(x, np.float64)
        assert xx.size == self.s.n
        y = np.empty(self.s.n, np.float64)
        c._check(_lib.load().hm_laplacian_matvec(c.handle, ip.ctypes.data, ix.ctypes.data,
                                                 sv.ctypes.data, self.s.n, xx.ctypes.data,
                                                 y.ctypes.data))
        return y


def default_n_epochs(n: int) -> int:
    """src/sgd.rs:23-25."""
    return 500 if n <= 10_000 else 200


@dataclass
class EdgeSchedule:
    """src/sgd.rs:30-34."""
    head: np.ndarray  # uint32
    tail: np.ndarray  # uint32
    epochs_per_sample: np.ndarray  # float64


def schedule_edges(graph: FuzzyGraph, n_epochs: int, ctx: "Context" = None) -> EdgeSchedule:
    """src/sgd.rs:36-63 on the device: edges below max(w) / prune_epochs
    dropped, epochs_per_sample = max(w) / w, in COO order."""
    c = ctx or Context.default()
    r = np.ascontiguousarray(graph.rows, np.uint32)
    cl = np.ascontiguousarray(graph.cols, np.uint32)
    w = np.ascontiguousarray(graph.vals, np.float32)
    m = r.size
    h = np.empty(max(m, 1), np.uint32)
    t = np.empty(max(m, 1), np.uint32)
    e = np.empty(max(m, 1), np.float64)
    cnt = ctypes.c_uint64()
    c._check(_lib.load().hm_schedule_edges(c.handle, r.ctypes.data, cl.ctypes.data, w.ctypes.data,
                                           m, graph.n, int(n_epochs), h.ctypes.data,
                                           t.ctypes.data, e.ctypes.data, ctypes.byref(cnt)))
    k = int(cnt.value)
    return EdgeSchedule(h[:k].copy(), t[:k].copy(), e[:k].copy())


# ---------------------------------------------------------------- context
class Context:
    """One device, one stream, an optional HBM budget (bytes; 0 = none)."""

    _default = None

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        st = L.hm_ctx_create(ctypes.byref(h), int(device), int(hbm_budget))
        if st != 0:
            raise DeviceError(f"hm_ctx_create failed ({st}): no usable GPU")
        self.handle = h

    @property
    def budget(self) -> int:
        """HBM budget in bytes (0 = none)."""
        return int(_lib.load().hm_ctx_budget(self.handle))

    @property
    def in_use(self) -> int:
        """Device bytes held by this context's live jobs."""
        return int(_lib.load().hm_ctx_in_use(self.handle))

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context()
        return cls._default

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().hm_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return (_lib.load().hm_last_error(self.handle) or b"").decode()

    def _check(self, st: int, *, n: int = 0, k: int = 0, length: int = 0, d: int = 0):
        if st == 0:
            return
        msg = self.last_error()
        if st == _lib.HM_EBADSHAPE:
            raise BadShape(length, d)
        if st == _lib.HM_ETOOFEW:
            raise TooFewPoints(n, k)
        if st == _lib.HM_EINVAL:
            raise InvalidParameter(msg)
        if st == _lib.HM_ENONFINITE:
            raise NonFiniteInput(-1)
        if st == _lib.HM_EBUDGET:
            raise MemoryError(msg)
        raise DeviceError(f"status {st}: {msg}")


def _opts(precision="bf16", calib="ref_f32", margin=-1, nsplit=0, use_graph=False,
          timing=False, autotune=False) -> _lib.Options:
    p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    c = CALIBRATIONS[calib] if isinstance(calib, str) else int(calib)
    return _lib.default_options(precision=p, calib=c, margin=margin, nsplit=nsplit,
                                use_graph=int(use_graph), timing=int(timing),
                                autotune=int(autotune))


def _validate(x: np.ndarray, n_features: int, k: int, finite: bool = True):
    """src/api.rs:113-144 (the checks the graph stages depend on). finite=False
    leaves the scan for non-finite values to the C entry, which refuses them
    (HM_ENONFINITE; _raise_nonfinite then names the index)."""
    if n_features == 0 or x.size == 0 or x.size % n_features != 0:
        raise BadShape(x.size, n_features)
    if k < 2:
        raise InvalidParameter("n_neighbors must be >= 2")
    n = x.size // n_features
    if n <= k:
        raise TooFewPoints(n, k)
    if finite:
        _raise_nonfinite(x)


def _raise_nonfinite(x: np.ndarray):
    bad = np.flatnonzero(~np.isfinite(x))
    if bad.size:
        raise NonFiniteInput(int(bad[0]))


# ---------------------------------------------------------------- stages
def exact_knn(data, n_features: int, k: int, metric="euclidean", *, precision="bf16",
              margin: int = -1, ctx: Context = None) -> Knn:
    """src/knn.rs:22-55 -- indices and distances bit-identical to the reference.

    As the reference's stage function it accepts any k <= n; the stricter
    fit_transform rules (n > k >= 2) apply to knn_graph / Holomap.graph."""
    x = np.ascontiguousarray(data, np.float32).reshape(-1)
    if n_features == 0 or x.size == 0 or x.size % n_features != 0:
        raise BadShape(x.size, n_features)
    if k < 1:
        raise InvalidParameter("n_neighbors must be >= 1")
    if x.size // n_features < k:
        raise TooFewPoints(x.size // n_features, k)
    bad = np.flatnonzero(~np.isfinite(x))
    if bad.size:
        raise NonFiniteInput(int(bad[0]))
    c = ctx or Context.default()
    n = x.size // n_features
    idx = np.empty((n, k), np.uint32)
    dist = np.empty((n, k), np.float32)
    o = _opts(precision, margin=margin)
    st = _lib.load().hm_exact_knn(c.handle, x.ctypes.data, n, n_features, k, _metric(metric),
                                  ctypes.byref(o), idx.ctypes.data, dist.ctypes.data)
    c._check(st, n=n, k=k, length=x.size, d=n_features)
    return Knn(k, idx, dist)


def smooth_knn(knn_dists, k: int, *, calib="ref_f32", ctx: Context = None) -> SmoothKnn:
    """src/smooth_knn.rs:23-80."""
    c = ctx or Context.default()
    d = np.ascontiguousarray(knn_dists, np.float32).reshape(-1)
    n = d.size // k
    s = np.empty(n, np.float32)
    r = np.empty(n, np.float32)
    cm = CALIBRATIONS[calib] if isinstance(calib, str) else int(calib)
    st = _lib.load().hm_smooth_knn(c.handle, d.ctypes.data, n, k, cm, s.ctypes.data, r.ctypes.data)
    c._check(st, n=n, k=k)
    return SmoothKnn(s, r)


def fuzzy_simplicial_set(knn: Knn, calib: SmoothKnn, ctx: Context = None) -> FuzzyGraph:
    """src/fuzzy.rs:25-85."""
    c = ctx or Context.default()
    idx = np.ascontiguousarray(knn.indices, np.uint32).reshape(-1)
    dist = np.ascontiguousarray(knn.dists, np.float32).reshape(-1)
    k = knn.k
    n = idx.size // k
    sig = np.ascontiguousarray(calib.sigmas, np.float32)
    rho = np.ascontiguousarray(calib.rhos, np.float32)
    cap = 2 * n * k
    rows = np.empty(cap, np.uint32)
    cols = np.empty(cap, np.uint32)
    vals = np.empty(cap, np.float32)
    nnz = ctypes.c_uint64()
    st = _lib.load().hm_fuzzy_simplicial_set(c.handle, idx.ctypes.data, dist.ctypes.data,
                                             sig.ctypes.data, rho.ctypes.data, n, k,
                                             rows.ctypes.data, cols.ctypes.data, vals.ctypes.data,
                                             cap, ctypes.byref(nnz))
    c._check(st, n=n, k=k)
    m = nnz.value
    return FuzzyGraph(n, rows[:m].copy(), cols[:m].copy(), vals[:m].copy())


@dataclass
class GraphResult:
    knn: Knn
    calib: SmoothKnn
    graph: FuzzyGraph
    indptr: np.ndarray
    # FP64 calibration with the kNN on the device: the largest relative error
    # of the f32 distances, sigmas and directed membership strengths against
    # their f64 values (the FP64 chain); None otherwise
    err64: Optional[tuple] = None


def _err64(o) -> Optional[tuple]:
    e = tuple(float(v) for v in o.err64)
    return None if e[0] < 0 else e


def knn_graph(data, n_features: int, k: int, metric="euclidean", *, precision="bf16",
              calib="ref_f32", margin: int = -1, ctx: Context = None) -> GraphResult:
    """The fused seam of src/api.rs:147-149 plus the CSR row pointer: one
    hm_knn_graph call (host buffers in and out; the C entry plans, runs and
    frees a job on the context's device)."""
    x = np.ascontiguousarray(data, np.float32).reshape(-1)
    _validate(x, n_features, k, finite=False)  # hm_knn_graph scans for non-finite values
    c = ctx or Context.default()
    n = x.size // n_features
    idx = np.empty((n, k), np.uint32)
    dist = np.empty((n, k), np.float32)
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    cm = CALIBRATIONS[calib] if isinstance(calib, str) else int(calib)
    s64 = np.empty(n, np.float64) if cm == HM_CALIB_F64 else None
    d64 = np.empty((n, k), np.float64) if cm == HM_CALIB_F64 else None
    cap = 2 * n * k
    rows = np.empty(cap, np.uint32)
    cols = np.empty(cap, np.uint32)
    vals = np.empty(cap, np.float32)
    indptr = np.empty(n + 1, np.uint64)
    o = _lib.GraphOut()
    o.indices, o.dists = idx.ctypes.data, dist.ctypes.data
    o.sigmas, o.rhos = sig.ctypes.data, rho.ctypes.data
    o.sigmas64 = s64.ctypes.data if s64 is not None else None
    o.dists64 = d64.ctypes.data if d64 is not None else None
    o.rows, o.cols, o.vals = rows.ctypes.data, cols.ctypes.data, vals.ctypes.data
    o.indptr = indptr.ctypes.data
    o.cap = cap
    opt = _opts(precision, calib, margin)
    st = _lib.load().hm_knn_graph(c.handle, x.ctypes.data, n, n_features, k, _metric(metric),
                                  ctypes.byref(opt), ctypes.byref(o))
    if st == _lib.HM_EBUDGET:
        raise MemoryError(c.last_error())
    if st == _lib.HM_ENONFINITE:
        _raise_nonfinite(x)
    c._check(st, n=n, k=k)
    m = int(o.nnz)
    g = FuzzyGraph(n, rows[:m].copy(), cols[:m].copy(), vals[:m].copy())
    return GraphResult(Knn(k, idx, dist, d64), SmoothKnn(sig, rho, s64), g, indptr, _err64(o))


# ---------------------------------------------------------------- jobs
class Job:
    """A planned batch of datasets with device-resident inputs and graphs."""

    def __init__(self, ctx: Context, n: Sequence[int], d: Sequence[int], k: Sequence[int],
                 metric: Sequence = None, *, precision="bf16", calib="ref_f32", margin=-1,
                 nsplit=0, use_graph=False, timing=False, autotune=False, stages=HM_STAGE_ALL):
        self.ctx = ctx
        nds = len(n)
        self.n = np.ascontiguousarray(n, np.uint64)
        self.d = np.ascontiguousarray(d, np.uint32)
        self.k = np.ascontiguousarray(k, np.uint32)
        m = [0] * nds if metric is None else [_metric(v) for v in metric]
        self.metric = np.ascontiguousarray(m, np.int32)
        self.calib_mode = CALIBRATIONS[calib] if isinstance(calib, str) else int(calib)
        o = _opts(precision, calib, margin, nsplit, use_graph, timing, autotune)
        h = ctypes.c_void_p()
        st = _lib.load().hm_job_create(ctx.handle, nds, self.n.ctypes.data, self.d.ctypes.data,
                                       self.k.ctypes.data, self.metric.ctypes.data,
                                       ctypes.byref(o), int(stages), ctypes.byref(h))
        if st == _lib.HM_EBUDGET:
            raise MemoryError(ctx.last_error())
        ctx._check(st, n=int(self.n[0]), k=int(self.k[0]))
        self.handle = h
        self.stages = stages
        self._keep = []

    @property
    def bytes(self) -> int:
        return int(_lib.load().hm_job_bytes(self.handle))

    def upload(self, ds: int, data) -> None:
        x = np.ascontiguousarray(data, np.float32).reshape(-1)
        assert x.size == int(self.n[ds]) * int(self.d[ds])
        self.ctx._check(_lib.load().hm_job_upload(self.handle, ds, x.ctypes.data))
        self._keep.append(x)  # the copy is asynchronous; keep the sources alive until sync

    def upload_device(self, ds: int, ptr: int) -> None:
        self.ctx._check(_lib.load().hm_job_upload_device(self.handle, ds, ctypes.c_void_p(ptr)))

    def run(self) -> None:
        self.ctx._check(_lib.load().hm_job_run(self.handle))

    def sync(self) -> None:
        self.ctx._check(_lib.load().hm_job_sync(self.handle))
        self._keep = []

    def done(self) -> bool:
        r = _lib.load().hm_job_done(self.handle)
        if r < 0:
            self.ctx._check(r)
        return r == 1

    def times(self) -> dict:
        a = np.zeros(_lib.HM_NTIMES, np.float32)
        self.ctx._check(_lib.load().hm_job_times(self.handle, a.ctypes.data))
        return {name: float(a[i]) for i, name in enumerate(TIME_NAMES)}

    def counters(self) -> dict:
        a = np.zeros(_lib.HM_NCOUNTERS, np.uint64)
        self.ctx._check(_lib.load().hm_job_counters(self.handle, a.ctypes.data))
        return {"fallback_rows": int(a[0]), "nnz": int(a[1]), "candidates": int(a[2]),
                "nsplit": int(a[3]), "escalated_rows": int(a[4]), "listed": int(a[5]),
                "target": int(a[6]), "sample_stride": int(a[7])}

    def download(self, ds: int, with_sigma64: bool = False) -> GraphResult:
        n, k = int(self.n[ds]), int(self.k[ds])
        idx = np.empty((n, k), np.uint32)
        dist = np.empty((n, k), np.float32)
        sig = np.empty(n, np.float32)
        rho = np.empty(n, np.float32)
        f64 = self.calib_mode == HM_CALIB_F64
        s64 = np.empty(n, np.float64) if (with_sigma64 or f64) else None
        d64 = np.empty((n, k), np.float64) if (f64 and self.stages & HM_STAGE_KNN) else None
        cap = 2 * n * k
        rows = np.empty(cap, np.uint32)
        cols = np.empty(cap, np.uint32)
        vals = np.empty(cap, np.float32)
        indptr = np.empty(n + 1, np.uint64)
        o = _lib.GraphOut()
        o.indices, o.dists = idx.ctypes.data, dist.ctypes.data
        o.sigmas, o.rhos = sig.ctypes.data, rho.ctypes.data
        o.sigmas64 = s64.ctypes.data if s64 is not None else None
        o.dists64 = d64.ctypes.data if d64 is not None else None
        if self.stages & HM_STAGE_FUZZY:
            o.rows, o.cols, o.vals = rows.ctypes.data, cols.ctypes.data, vals.ctypes.data
            o.indptr = indptr.ctypes.data
        o.cap = cap
        self.ctx._check(_lib.load().hm_job_download(self.handle, ds, ctypes.byref(o)))
        m = int(o.nnz)
        g = FuzzyGraph(n, rows[:m].copy(), cols[:m].copy(), vals[:m].copy())
        return GraphResult(Knn(k, idx, dist, d64), SmoothKnn(sig, rho, s64), g, indptr,
                           _err64(o))

    def connected_components(self, ds: int) -> Components:
        """src/components.rs:13-39 on dataset ds's resident graph."""
        lab = np.empty(int(self.n[ds]), np.uint32)
        nc = ctypes.c_uint64()
        self.ctx._check(_lib.load().hm_job_connected_components(self.handle, ds, lab.ctypes.data,
                                                                ctypes.byref(nc)))
        return Components(int(nc.value), lab)

    def schedule_edges(self, ds: int, n_epochs: int) -> EdgeSchedule:
        """src/sgd.rs:36-63 on dataset ds's resident graph."""
        nnz = self.counters_nnz(ds)
        h = np.empty(max(nnz, 1), np.uint32)
        t = np.empty(max(nnz, 1), np.uint32)
        e = np.empty(max(nnz, 1), np.float64)
        cnt = ctypes.c_uint64()
        self.ctx._check(_lib.load().hm_job_schedule_edges(self.handle, ds, int(n_epochs),
                                                          h.ctypes.data, t.ctypes.data,
                                                          e.ctypes.data, ctypes.byref(cnt)))
        k = int(cnt.value)
        return EdgeSchedule(h[:k].copy(), t[:k].copy(), e[:k].copy())

    def counters_nnz(self, ds: int) -> int:
        """Graph entries of dataset ds."""
        o = _lib.GraphOut()
        self.ctx._check(_lib.load().hm_job_download(self.handle, ds, ctypes.byref(o)))
        return int(o.nnz)

    def laplacian_matvec(self, ds: int, x_dev: int, y_dev: int) -> None:
        """y = L x (src/sparse.rs:87-93) on dataset ds's resident graph; x and
        y are device pointers to n doubles."""
        self.ctx._check(_lib.load().hm_job_laplacian_matvec(self.handle, ds,
                                                            ctypes.c_void_p(x_dev),
                                                            ctypes.c_void_p(y_dev)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().hm_job_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving
class Server:
    """Fit-request serving loop (hm_server_*): requests queue in arrival
    order; progress() retires finished jobs and admits the longest queue
    prefix that fits max_batch / max_batch_rows and the context's free HBM
    budget into a new job, with up to max_jobs jobs running at once on their
    own streams. Results stay on the device until fetch()."""

    QUEUED, RUNNING, DONE = 0, 1, 2

    def __init__(self, ctx: Context, *, max_jobs=4, max_batch=32, max_batch_rows=1 << 20,
                 precision="bf16", calib="ref_f32", margin=-1):
        self.ctx = ctx
        L = _lib.load()
        cfg = _lib.ServerConfig()
        L.hm_server_default_config(ctypes.byref(cfg))
        cfg.max_jobs, cfg.max_batch, cfg.max_batch_rows = max_jobs, max_batch, max_batch_rows
        cfg.opt = _opts(precision, calib, margin, 0, False, False)
        h = ctypes.c_void_p()
        ctx._check(L.hm_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._shape = {}

    def submit(self, data, n_features: int, k: int, metric="euclidean") -> int:
        x = np.ascontiguousarray(data, np.float32).reshape(-1)
        _validate(x, n_features, k, finite=False)  # hm_server_submit scans for them
        rid = ctypes.c_uint64()
        st = _lib.load().hm_server_submit(self.handle, x.ctypes.data, x.size // n_features,
                                          n_features, k, _metric(metric), ctypes.byref(rid))
        if st == _lib.HM_ENONFINITE:
            _raise_nonfinite(x)
        self.ctx._check(st)
        self._shape[rid.value] = (x.size // n_features, k)
        return rid.value

    def progress(self) -> int:
        c = ctypes.c_uint32()
        self.ctx._check(_lib.load().hm_server_progress(self.handle, ctypes.byref(c)))
        return c.value

    def state(self, rid: int) -> int:
        return int(_lib.load().hm_server_state(self.handle, rid))

    def pending(self) -> int:
        return int(_lib.load().hm_server_pending(self.handle))

    def fetch(self, rid: int) -> GraphResult:
        """Results of a done request (raises its error for a failed one)."""
        n, k = self._shape.pop(rid)
        L = _lib.load()
        st = self.state(rid)
        if st < 0:
            self.ctx._check(L.hm_server_fetch(self.handle, rid, None), n=n, k=k)
        nnz = ctypes.c_uint64()
        self.ctx._check(L.hm_server_result_info(self.handle, rid, None, None, ctypes.byref(nnz)))
        m = int(nnz.value)
        idx = np.empty((n, k), np.uint32)
        dist = np.empty((n, k), np.float32)
        sig = np.empty(n, np.float32)
        rho = np.empty(n, np.float32)
        # exactly nnz entries (known before the copy): no truncating copies
        rows = np.empty(max(m, 1), np.uint32)
        cols = np.empty(max(m, 1), np.uint32)
        vals = np.empty(max(m, 1), np.float32)
        indptr = np.empty(n + 1, np.uint64)
        o = _lib.GraphOut()
        o.indices, o.dists = idx.ctypes.data, dist.ctypes.data
        o.sigmas, o.rhos = sig.ctypes.data, rho.ctypes.data
        o.rows, o.cols, o.vals = rows.ctypes.data, cols.ctypes.data, vals.ctypes.data
        o.indptr = indptr.ctypes.data
        o.cap = max(m, 1)
        self.ctx._check(L.hm_server_fetch(self.handle, rid, ctypes.byref(o)))
        g = FuzzyGraph(n, rows[:m], cols[:m], vals[:m])
        return GraphResult(Knn(k, idx, dist), SmoothKnn(sig, rho, None), g, indptr)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().hm_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def l2_normalise(vectors, ctx: Context = None) -> np.ndarray:
    """crates/holomap-clusterer/src/pipeline.rs:151-162 on the device: each
    row divided by its f32 norm, zero rows as they are."""
    c = ctx or Context.default()
    a = np.ascontiguousarray(vectors, np.float32)
    assert a.ndim == 2
    out = np.empty_like(a)
    c._check(_lib.load().hm_l2_normalise(c.handle, a.ctypes.data, a.shape[0], a.shape[1],
                                         out.ctypes.data))
    return out


def expf_device(x, ctx: Context = None) -> np.ndarray:
    """The device's f32 exp (the reference's libm method) over an array."""
    c = ctx or Context.default()
    a = np.ascontiguousarray(x, np.float32).reshape(-1)
    y = np.empty_like(a)
    c._check(_lib.load().hm_expf_array(c.handle, a.ctypes.data, y.ctypes.data, a.size))
    return y


# ---------------------------------------------------------------- builder
class Holomap:
    """src/api.rs:66-107 -- a seeded builder; only the graph stages run here.

    ``graph(data, n_features)`` performs fit_transform's validation
    (src/api.rs:113-144) and the three graph calls (src/api.rs:147-149);
    the spectral initialisation and SGD layout that consume the graph stay
    on the reference's CPU code.
    """

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.n_components = 2
        self.n_neighbors = 15
        self.min_dist = 0.1
        self.spread = 1.0
        self.metric = Metric.EUCLIDEAN
        self.n_epochs = None
        self.precision = "bf16"

    @staticmethod
    def builder(seed: int) -> "HolomapBuilder":
        return HolomapBuilder(Holomap(seed))

    def graph(self, data, n_features: int, ctx: Context = None) -> GraphResult:
        x = np.ascontiguousarray(data, np.float32).reshape(-1)
        if n_features == 0 or x.size == 0 or x.size % n_features != 0:
            raise BadShape(x.size, n_features)
        if self.n_components == 0:
            raise InvalidParameter("n_components must be >= 1")
        if self.n_neighbors < 2:
            raise InvalidParameter("n_neighbors must be >= 2")
        if not (self.spread > 0.0):
            raise InvalidParameter("spread must be > 0")
        if not (self.min_dist >= 0.0):
            raise InvalidParameter("min_dist must be >= 0")
        if self.n_epochs == 0:
            raise InvalidParameter("n_epochs must be >= 1")
        n = x.size // n_features
        if n <= self.n_neighbors:
            raise TooFewPoints(n, self.n_neighbors)
        return knn_graph(x, n_features, self.n_neighbors, self.metric, precision=self.precision,
                         ctx=ctx)


class HolomapBuilder:
    """Setters of src/api.rs:195-235."""

    def __init__(self, inner: Holomap):
        self.inner = inner

    def n_components(self, v: int):
        self.inner.n_components = int(v)
        return self

    def n_neighbors(self, v: int):
        self.inner.n_neighbors = int(v)
        return self

    def min_dist(self, v: float):
        self.inner.min_dist = float(v)
        return self

    def spread(self, v: float):
        self.inner.spread = float(v)
        return self

    def metric(self, m):
        self.inner.metric = m if isinstance(m, Metric) else Metric.from_str(m)
        return self

    def n_epochs(self, v: int):
        self.inner.n_epochs = int(v)
        return self

    def precision(self, p: str):
        if p not in PRECISIONS:
            raise InvalidParameter("precision must be exact, fp32, bf16 or bf16x3")
        self.inner.precision = p
        return self

    def build(self) -> Holomap:
        return self.inner

    def graph(self, data, n_features: int, ctx: Context = None) -> GraphResult:
        return self.inner.graph(data, n_features, ctx)
