"""GPU parity at the benchmark configurations' full sizes.

  * N = 50,000 x d = 784 image-like rows (k = 15): kNN indices and distance
    bits against the oracle on sampled query rows, all precision modes byte
    for byte, sigma / rho and the graph against the oracle on the GPU's kNN,
    the FP64 calibration against the FP64 oracle;
  * the 256-dataset batch (N log-uniform in [200, 8000]): every dataset equal
    to its single-dataset run, the BF16 batch equal to the exact-arithmetic
    batch on every row (escalated rows included), the exact batch against the
    oracle on sampled rows, and the downstream stages against the oracle;
  * the serving loop with requests up to N = 20,000 at d = 768 under both
    metrics, each served result checked against the oracle.

The kNN oracle runs on sampled rows (all N columns each) with 8 threads;
rows are independent, so its outputs equal the single-threaded ones.
"""
import numpy as np
import pytest

import holomap_amd as H
import oracle as O
from holomap_amd import datasets as DS
from fp64_chain import check_fp64_chain

pytestmark = pytest.mark.gpu


def run_job(ctx, x, d, k, metric, **kw):
    n = x.shape[0]
    job = H.Job(ctx, [n], [d], [k], [metric], **kw)
    job.upload(0, x)
    job.run()
    job.sync()
    r = job.download(0, with_sigma64=True)
    c = job.counters()
    job.close()
    return r, c


def same_bytes(a, b):
    return (a.knn.indices.tobytes() == b.knn.indices.tobytes()
            and a.knn.dists.tobytes() == b.knn.dists.tobytes()
            and a.calib.sigmas.tobytes() == b.calib.sigmas.tobytes()
            and a.calib.rhos.tobytes() == b.calib.rhos.tobytes()
            and a.graph.rows.tobytes() == b.graph.rows.tobytes()
            and a.graph.cols.tobytes() == b.graph.cols.tobytes()
            and a.graph.vals.tobytes() == b.graph.vals.tobytes())


def check_rows_vs_oracle(res, x, d, k, metric, rows):
    sub = O.exact_knn_rows(x, d, k, metric, rows, nthreads=8)
    assert np.array_equal(res.knn.indices[rows], sub.indices)
    assert np.array_equal(res.knn.dists[rows].view(np.uint32), sub.dists.view(np.uint32))


def check_downstream_vs_oracle(res, k):
    """sigma / rho and the fuzzy graph of the oracle, fed the GPU's kNN."""
    cal = O.smooth_knn(res.knn.dists, k)
    assert np.array_equal(res.calib.rhos.view(np.uint32), cal.rhos.view(np.uint32))
    assert np.array_equal(res.calib.sigmas.view(np.uint32), cal.sigmas.view(np.uint32))
    g = O.fuzzy_simplicial_set(O.Knn(k, res.knn.indices, res.knn.dists), cal)
    assert np.array_equal(res.graph.rows, g.rows)
    assert np.array_equal(res.graph.cols, g.cols)
    assert np.array_equal(res.graph.vals.view(np.uint32), g.vals.view(np.uint32))
    assert np.array_equal(res.indptr, O.csr_indptr(g.rows, res.knn.indices.shape[0]))


@pytest.fixture(scope="module")
def c3_data():
    return DS.c3()


def test_c3_bf16_vs_oracle_and_modes(ctx, c3_data):
    x = c3_data
    n, d, k = x.shape[0], 784, 15
    out = {}
    for prec in ["bf16", "fp32", "exact"]:
        # the configuration's margin 2k (45 candidates per row)
        out[prec] = run_job(ctx, x, d, k, "euclidean", precision=prec, margin=2 * k)
    res, cnt = out["bf16"]
    rows = np.random.default_rng(5).choice(n, 256, replace=False).astype(np.uint32)
    check_rows_vs_oracle(res, x, d, k, "euclidean", rows)
    for prec in ["fp32", "exact"]:
        assert same_bytes(res, out[prec][0]), prec
    check_downstream_vs_oracle(res, k)


def test_c3_f64_calibration_vs_fp64_oracle(ctx, c3_data):
    # the configuration: margin 2k, FP64 smooth-kNN calibration
    x = c3_data
    k = 15
    r, _ = run_job(ctx, x, 784, k, "euclidean", calib="f64", margin=2 * k)
    check_fp64_chain(r, x, 784, k, "euclidean")
    # the FP64 root's distance from the reference's f32 search
    ref32 = O.smooth_knn(r.knn.dists, k)
    assert (np.abs(ref32.sigmas - r.calib.sigmas64) / r.calib.sigmas64).max() < 1e-3


def test_c4_full_batch(ctx):
    sets = DS.c4()
    assert len(sets) == 256
    shape = ([s.n for s in sets], [s.d for s in sets], [s.k for s in sets],
             [s.metric for s in sets])
    got = {}
    for prec in ["bf16", "exact"]:
        job = H.Job(ctx, *shape, precision=prec)
        for i, s in enumerate(sets):
            job.upload(i, s.data)
        job.run()
        job.sync()
        got[prec] = [job.download(i) for i in range(len(sets))]
        job.close()
    # every row -- escalated or exact-scanned ones included, where the bf16
    # keys leave some uncertified -- equals the exact arithmetic
    for i in range(len(sets)):
        assert same_bytes(got["bf16"][i], got["exact"][i]), i
    rng = np.random.default_rng(8)
    for i, s in enumerate(sets):
        alone = H.knn_graph(s.data, s.d, s.k, s.metric, ctx=ctx)
        assert same_bytes(got["bf16"][i], alone), i
        if i % 4 == 0:
            rows = rng.choice(s.n, min(s.n, 24), replace=False).astype(np.uint32)
            check_rows_vs_oracle(got["exact"][i], s.data, s.d, s.k, s.metric, rows)
        check_downstream_vs_oracle(got["bf16"][i], s.k)


def c5_shaped(seed, n, d, metric, k):
    rng = np.random.default_rng(seed)
    if metric == "euclidean":
        x = DS.gaussian_mixture(n, d, int(rng.integers(5, 41)), 4.0, rng)
    else:
        x = rng.normal(0.0, 1.0, size=(n, d))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        x = x.astype(np.float32)
    return DS.Dataset(x, n, d, k, metric)


def test_server_results_vs_oracle(ctx):
    # two requests at N = 20,000, d = 768 (one per metric) among smaller
    # ones of the C5 mix; every served result is checked against the oracle
    reqs = [c5_shaped(501, 20000, 768, "euclidean", 15), c5_shaped(502, 20000, 768, "cosine", 30)]
    reqs += [DS.c5_request(600 + i, n_min=500, n_max=6000) for i in range(8)]
    srv = H.Server(ctx, max_jobs=3, max_batch=4)
    ids = [srv.submit(r.data, r.d, r.k, r.metric) for r in reqs]
    out = {}
    while len(out) < len(ids):
        srv.progress()
        for rid in ids:
            if rid not in out and srv.state(rid) == srv.DONE:
                out[rid] = srv.fetch(rid)
        for rid in ids:
            assert srv.state(rid) >= 0 or rid in out, (rid, srv.state(rid))
    srv.close()
    rng = np.random.default_rng(9)
    for rid, r in zip(ids, reqs):
        res = out[rid]
        rows = rng.choice(r.n, min(r.n, 48), replace=False).astype(np.uint32)
        check_rows_vs_oracle(res, r.data, r.d, r.k, r.metric, rows)
        check_downstream_vs_oracle(res, r.k)
