"""hm_knn_graph_noalloc: the graph build with device-memory inputs and
outputs and every intermediate buffer in the library's fixed scratch array.
Checked against hm_knn_graph (host buffers, pooled allocations) bit for bit
on two calls in a row, the second with a smaller input (the array keeps the
first call's contents), plus its refusals: an input too large for the array,
a non-finite value, a short COO capacity."""
import ctypes

import numpy as np
import pytest

import holomap_amd as H
from holomap_amd import _lib
from holomap_amd import datasets as DS
from hip_util import Hip

pytestmark = pytest.mark.gpu


def noalloc_graph(ctx, hip, x, d, k, metric, cap=None, watch=None):
    """Device buffers for the call (allocated here, by the test), the call,
    the results back on the host. watch: a list that receives the device's
    free bytes just before and just after the call."""
    L = _lib.load()
    x = np.ascontiguousarray(x, np.float32)
    n = x.size // d
    cap = 2 * n * k if cap is None else cap
    dx = hip.alloc(x.nbytes)
    hip.h2d(dx, x)
    o = _lib.GraphOut()
    o.indices, o.dists = hip.alloc(n * k * 4), hip.alloc(n * k * 4)
    o.sigmas, o.rhos = hip.alloc(n * 4), hip.alloc(n * 4)
    o.rows, o.cols, o.vals = hip.alloc(cap * 4), hip.alloc(cap * 4), hip.alloc(cap * 4)
    o.indptr = hip.alloc((n + 1) * 8)
    o.sigmas64 = None
    o.cap = cap
    opt = _lib.default_options()
    m = _lib.HM_EUCLIDEAN if metric == "euclidean" else _lib.HM_COSINE
    if watch is not None:
        watch.append(hip.free_bytes())
    st = L.hm_knn_graph_noalloc(ctx.handle, dx, n, d, k, m, ctypes.byref(opt), ctypes.byref(o))
    if watch is not None:
        watch.append(hip.free_bytes())
    if st != 0:
        return st, None
    nnz = int(o.nnz)
    res = {"indices": np.empty((n, k), np.uint32), "dists": np.empty((n, k), np.float32),
           "sigmas": np.empty(n, np.float32), "rhos": np.empty(n, np.float32),
           "rows": np.empty(nnz, np.uint32), "cols": np.empty(nnz, np.uint32),
           "vals": np.empty(nnz, np.float32), "indptr": np.empty(n + 1, np.uint64)}
    for name in res:
        hip.d2h(res[name], getattr(o, name))
    return st, res


def check_same(res, ref):
    assert res["indices"].tobytes() == ref.knn.indices.tobytes()
    assert res["dists"].tobytes() == ref.knn.dists.tobytes()
    assert res["sigmas"].tobytes() == ref.calib.sigmas.tobytes()
    assert res["rhos"].tobytes() == ref.calib.rhos.tobytes()
    assert res["rows"].tobytes() == ref.graph.rows.tobytes()
    assert res["cols"].tobytes() == ref.graph.cols.tobytes()
    assert res["vals"].tobytes() == ref.graph.vals.tobytes()
    assert res["indptr"].tobytes() == ref.indptr.tobytes()


def test_noalloc_matches_hm_knn_graph_twice(ctx):
    hip = Hip()
    try:
        cases = [(DS.c2(n=12000, d=128), 128, 15, "euclidean"),
                 (DS.c2(seed=9, n=3000, d=64), 64, 30, "cosine")]
        for i, (x, d, k, metric) in enumerate(cases):
            watch = []
            st, res = noalloc_graph(ctx, hip, x, d, k, metric, watch=watch)
            assert st == 0, ctx.last_error()
            if i > 0:
                # no device memory taken by the call itself
                assert watch[1] == watch[0], watch
            ref = H.knn_graph(x, d, k, metric, ctx=ctx)
            check_same(res, ref)
        st, res = noalloc_graph(ctx, hip, cases[1][0], 64, 30, "cosine")
        assert st == 0
        check_same(res, H.knn_graph(cases[1][0], 64, 30, "cosine", ctx=ctx))
    finally:
        hip.close()


def test_noalloc_refusals(ctx):
    hip = Hip()
    try:
        # larger than the scratch array
        big = np.zeros((60000, 256), np.float32)
        big[:, 0] = np.arange(60000, dtype=np.float32)
        st, _ = noalloc_graph(ctx, hip, big, 256, 15, "euclidean")
        assert st == _lib.HM_EBUDGET
        assert "scratch" in ctx.last_error()
        # a non-finite value, found on the device
        x = DS.c2(n=2000, d=32)
        x[777, 5] = np.nan
        st, _ = noalloc_graph(ctx, hip, x, 32, 15, "euclidean")
        assert st == _lib.HM_ENONFINITE
        assert str(777 * 32 + 5) in ctx.last_error()
        # a COO capacity below the entry count
        x[777, 5] = 0.0
        st, _ = noalloc_graph(ctx, hip, x, 32, 15, "euclidean", cap=100)
        assert st == _lib.HM_EINVAL
        # and a normal call afterwards
        st, res = noalloc_graph(ctx, hip, x, 32, 15, "euclidean")
        assert st == 0
        check_same(res, H.knn_graph(x, 32, 15, "euclidean", ctx=ctx))
    finally:
        hip.close()
