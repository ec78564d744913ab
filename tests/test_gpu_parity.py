"""GPU parity: the HIP path against the oracle and the reference's golden
fixtures. kNN indices/distances, rho, sigma (ref_f32 mode) and the fuzzy
graph must be bit-identical; FP64 calibration within the stated tolerance."""
import ctypes
import json
import os

import numpy as np
import pytest

import holomap_amd as H
import oracle as O
from holomap_amd import datasets as DS
from fp64_chain import check_fp64_chain

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = ["blobs_euclidean", "cosine", "duplicates_euclidean"]
PRECS = ["exact", "fp32", "bf16", "bf16x3"]


def load_fixture(name):
    with open(os.path.join(GOLD, f"stage2_{name}.json")) as f:
        fx = json.load(f)
    return fx, np.array(fx["data"], np.float64).astype(np.float32)


def assert_graph_equal(res, ref_knn, ref_cal, ref_g):
    assert np.array_equal(res.knn.indices, ref_knn.indices)
    assert np.array_equal(res.knn.dists.view(np.uint32), ref_knn.dists.view(np.uint32))
    assert np.array_equal(res.calib.rhos.view(np.uint32), ref_cal.rhos.view(np.uint32))
    assert np.array_equal(res.calib.sigmas.view(np.uint32), ref_cal.sigmas.view(np.uint32))
    assert np.array_equal(res.graph.rows, ref_g.rows)
    assert np.array_equal(res.graph.cols, ref_g.cols)
    assert np.array_equal(res.graph.vals.view(np.uint32), ref_g.vals.view(np.uint32))


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("prec", PRECS)
def test_fixtures_all_precisions(ctx, name, prec):
    fx, data = load_fixture(name)
    d, k = fx["n_features"], fx["n_neighbors"]
    res = H.knn_graph(data, d, k, fx["metric"], precision=prec, ctx=ctx)
    assert np.array_equal(res.knn.indices, np.array(fx["knn_indices"], np.uint32))
    assert np.array_equal(res.knn.dists, np.array(fx["knn_dists"], np.float64).astype(np.float32))
    assert np.array_equal(res.graph.rows, np.array(fx["graph_rows"], np.uint32))
    assert np.array_equal(res.graph.cols, np.array(fx["graph_cols"], np.uint32))
    e = np.array(fx["graph_vals"])
    assert np.all(np.abs(res.graph.vals - e) <= 1e-5 * (1 + np.abs(e)))
    ref = O.knn_graph(data, d, k, fx["metric"])
    assert_graph_equal(res, *ref)
    ip = O.csr_indptr(ref[2].rows, data.shape[0])
    assert np.array_equal(res.indptr, ip)


def test_known_answers(ctx):
    # src/knn.rs:64-104, src/smooth_knn.rs:103-124, src/fuzzy.rs:111-153
    knn = H.exact_knn(np.array([0.0, 1.0, 3.0, 7.0], np.float32), 1, 3, ctx=ctx)
    assert knn.indices[1].tolist() == [1, 0, 2] and knn.dists[3].tolist() == [0.0, 4.0, 6.0]
    dup = H.exact_knn(np.array([5.0, 9.0, 5.0], np.float32), 1, 2, ctx=ctx)
    assert dup.indices[2].tolist() == [0, 2]
    c = H.exact_knn(np.array([0.3, 0.7, 0.31, 0.69], np.float32), 2, 2, "cosine", ctx=ctx)
    assert c.dists[0, 0] == 0.0 and c.dists[0, 1] > 0.0
    a = H.exact_knn(np.array([1.0, 0.0, 100.0, 1.0, 0.1, 0.1], np.float32), 2, 3, "cosine", ctx=ctx)
    assert a.indices[0].tolist() == [0, 1, 2]
    cal = H.smooth_knn(np.array([0, 1, 3, 0, 1, 2, 0, 2, 3], np.float32), 3, ctx=ctx)
    assert cal.sigmas.tolist() == [3.72998046875, 1.864990234375, 1.864990234375]
    assert cal.rhos.tolist() == [1.0, 1.0, 2.0]
    z = H.smooth_knn(np.zeros(4, np.float32), 2, ctx=ctx)
    assert z.sigmas.tolist() == [1.0, 1.0]
    k2 = H.Knn(2, np.array([0, 1, 1, 2, 2, 1], np.uint32).reshape(3, 2),
               np.array([0, 1, 0, 0.1, 0, 0.1], np.float32).reshape(3, 2))
    c2 = H.SmoothKnn(np.ones(3, np.float32), np.array([0.5, 0.1, 0.1], np.float32))
    g = H.fuzzy_simplicial_set(k2, c2, ctx=ctx)
    w = O.expf(-0.5)
    assert g.rows.tolist() == [0, 1, 1, 2] and g.cols.tolist() == [1, 0, 2, 1]
    assert g.vals.tolist() == [w, w, 1.0, 1.0]
    csr = H.Csr.from_fuzzy(H.FuzzyGraph(3, np.array([0, 1, 1, 2], np.uint32),
                                        np.array([1, 0, 2, 1], np.uint32),
                                        np.ones(4, np.float32)), ctx=ctx)
    assert csr.indptr.tolist() == [0, 1, 3, 4]
    idx, vals = csr.row(1)
    assert idx.tolist() == [0, 2] and vals.tolist() == [1.0, 1.0]


def test_device_expf_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    bits = rng.integers(0x80000000, 0xC2D00000, size=200000, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([bits.view(np.float32), np.linspace(-110, 0, 5001, dtype=np.float32),
                        np.array([0.0, -0.0, -1e-30, -87.5, -103.9, -104.0], np.float32)])
    y = H.expf_device(x, ctx=ctx)
    L = O.lib()
    ref = np.array([L.hmo_expf(float(v)) for v in x], np.float32)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_c1_full_graph(ctx, prec, metric):
    x = DS.c1()
    res = H.knn_graph(x, 50, 15, metric, precision=prec, ctx=ctx)
    assert_graph_equal(res, *O.knn_graph(x, 50, 15, metric))


# d <= 192 runs the query-resident bf16 sweep, longer rows the streaming one
# (d = 200: four 128-byte slices with a zero-padded tail; d = 449: eight)
@pytest.mark.parametrize("shape", [(257, 3, 5), (130, 33, 7), (129, 1, 2), (200, 50, 64),
                                   (65, 130, 64), (1000, 17, 30), (300, 200, 15),
                                   (385, 449, 10)])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_ragged_shapes(ctx, shape, metric):
    n, d, k = shape
    rng = np.random.default_rng(n * 7 + d)
    x = rng.normal(size=(n, d)).astype(np.float32)
    ref = O.knn_graph(x, d, k, metric)
    for prec in ["bf16", "fp32"]:
        res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


def test_duplicates_zero_rows_and_ties(ctx):
    rng = np.random.default_rng(3)
    n, d, k = 600, 12, 15
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[100:130] = x[5]          # 31 identical rows: rho = 0 rows, mean_all floor
    x[200] = 0.0
    x[201] = 0.0               # zero vectors (cosine distance 0 to all)
    x[300:310] = np.round(x[300:310])  # many exact distance ties
    x[301] = x[300] * 2.0      # parallel pair
    for metric in ["euclidean", "cosine"]:
        ref = O.knn_graph(x, d, k, metric)
        for prec in PRECS:
            res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
            assert_graph_equal(res, *ref)


def test_c2_sampled_rows_and_mode_agreement(ctx):
    x = DS.c2()
    n, d, k = x.shape[0], 128, 15
    out = {}
    for prec in ["bf16", "fp32", "exact"]:
        job = H.Job(ctx, [n], [d], [k], ["euclidean"], precision=prec)
        job.upload(0, x)
        job.run()
        job.sync()
        out[prec] = (job.download(0), job.counters())
        job.close()
    rows = np.random.default_rng(0).choice(n, 256, replace=False).astype(np.uint32)
    sub = O.exact_knn_rows(x, d, k, "euclidean", rows, nthreads=8)
    for prec, (res, cnt) in out.items():
        assert np.array_equal(res.knn.indices[rows], sub.indices), prec
        assert np.array_equal(res.knn.dists[rows].view(np.uint32), sub.dists.view(np.uint32)), prec
    base = out["exact"][0]
    for prec in ["bf16", "fp32"]:
        r = out[prec][0]
        assert np.array_equal(r.knn.indices, base.knn.indices)
        assert np.array_equal(r.knn.dists.view(np.uint32), base.knn.dists.view(np.uint32))
        assert np.array_equal(r.graph.vals.view(np.uint32), base.graph.vals.view(np.uint32))
    # the downstream stages against the oracle, on the GPU's kNN
    cal = O.smooth_knn(base.knn.dists, k)
    assert np.array_equal(base.calib.sigmas.view(np.uint32), cal.sigmas.view(np.uint32))
    g = O.fuzzy_simplicial_set(O.Knn(k, base.knn.indices, base.knn.dists), cal)
    assert np.array_equal(base.graph.rows, g.rows) and np.array_equal(base.graph.cols, g.cols)
    assert np.array_equal(base.graph.vals.view(np.uint32), g.vals.view(np.uint32))


def test_determinism_byte_identical(ctx):
    x = DS.c2(n=6000)
    a = H.knn_graph(x, 128, 15, precision="bf16", ctx=ctx)
    b = H.knn_graph(x, 128, 15, precision="bf16", ctx=ctx)
    for f in ["indices", "dists"]:
        assert getattr(a.knn, f).tobytes() == getattr(b.knn, f).tobytes()
    assert a.graph.vals.tobytes() == b.graph.vals.tobytes()
    assert a.graph.cols.tobytes() == b.graph.cols.tobytes()


def test_batch_equals_alone(ctx):
    rng = np.random.default_rng(4)
    specs = [(700, 32, 10, "euclidean"), (333, 64, 15, "cosine"), (1500, 128, 30, "euclidean"),
             (260, 256, 15, "cosine")]
    data = [rng.normal(size=(n, d)).astype(np.float32) for n, d, _, _ in specs]
    job = H.Job(ctx, [s[0] for s in specs], [s[1] for s in specs], [s[2] for s in specs],
                [s[3] for s in specs])
    for i, x in enumerate(data):
        job.upload(i, x)
    job.run()
    job.sync()
    for i, (n, d, k, m) in enumerate(specs):
        r = job.download(i)
        assert_graph_equal(r, *O.knn_graph(data[i], d, k, m))
    job.close()


def test_floor_and_maxima_over_replays_and_batches(ctx):
    # The global-mean floor's per-dataset flags are set during a run and
    # cleared for the next one: a batch where only some datasets need the
    # floor, replayed, and re-uploaded with the roles swapped.
    rng = np.random.default_rng(11)
    specs = [(600, 12, 15, "euclidean"), (900, 40, 10, "cosine"), (500, 12, 15, "euclidean")]
    data = [rng.normal(size=(n, d)).astype(np.float32) for n, d, _, _ in specs]
    data[0][100:130] = data[0][5]  # 31 identical rows: rho = 0, floor applies
    data[1][:40] = data[1][7]      # 40 parallel copies: cosine rho = 0 as well
    job = H.Job(ctx, [s[0] for s in specs], [s[1] for s in specs], [s[2] for s in specs],
                [s[3] for s in specs], use_graph=True)
    for phase in range(2):
        for i, x in enumerate(data):
            job.upload(i, x)
        refs = [O.knn_graph(x, d, k, m) for x, (_, d, k, m) in zip(data, specs)]
        for _ in range(3):
            job.run()
            job.sync()
            for i in range(len(specs)):
                assert_graph_equal(job.download(i), *refs[i])
        data[0], data[2] = rng.normal(size=(600, 12)).astype(np.float32), data[0][:500].copy()
    job.close()


def test_graph_replay(ctx):
    x = DS.c2(n=3000, d=64)
    ref = O.knn_graph(x, 64, 15, "euclidean")
    job = H.Job(ctx, [3000], [64], [15], ["euclidean"], use_graph=True)
    job.upload(0, x)
    for _ in range(3):
        job.run()
        job.sync()
        assert_graph_equal(job.download(0), *ref)
    y = DS.c2(seed=8, n=3000, d=64)
    job.upload(0, y)
    job.run()
    job.sync()
    assert_graph_equal(job.download(0), *O.knn_graph(y, 64, 15, "euclidean"))
    job.close()


# k = 15, 30, 50: lane groups of 16 and 32 (DPP folds) and 64 (permutes)
@pytest.mark.parametrize("k", [15, 30, 50])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_f64_calibration_vs_fp64_oracle(ctx, k, metric):
    x = DS.c1()
    job = H.Job(ctx, [1000], [50], [k], [metric], calib="f64")
    job.upload(0, x)
    job.run()
    job.sync()
    r = job.download(0)
    job.close()
    check_fp64_chain(r, x, 50, k, metric)
    # the same through the fused host entry
    f = H.knn_graph(x, 50, k, metric, calib="f64", ctx=ctx)
    assert f.knn.dists64.tobytes() == r.knn.dists64.tobytes()
    assert f.graph.vals.tobytes() == r.graph.vals.tobytes() and f.err64 == r.err64
    # the reference's f32 search stays within 1e-3 of the FP64 root
    ref32 = O.smooth_knn(r.knn.dists, k)
    assert (np.abs(ref32.sigmas - r.calib.sigmas64) / r.calib.sigmas64).max() < 1e-3
    # without the FP64 mode there is no statement and no f64 distances
    assert H.knn_graph(x, 50, k, metric, ctx=ctx).err64 is None


def test_f64_chain_on_escalated_and_fallback_rows(ctx):
    # rows the bf16 keys cannot certify take the bf16x3 re-sweep and the
    # exact scan: their f64 distances come from the re-rank of the re-sweep
    # and from the exact merge
    rng = np.random.default_rng(44)
    n, d, k = 2500, 40, 15
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[:300] = x[0] + rng.normal(size=(300, d)).astype(np.float32) * 1e-4  # a tight cluster
    for prec in ["bf16", "exact"]:
        r = H.knn_graph(x, d, k, "euclidean", precision=prec, calib="f64", ctx=ctx)
        check_fp64_chain(r, x, d, k, "euclidean")


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_heavy_ties_in_candidate_lists(ctx, metric):
    # 300 identical rows make equal selection keys straddle the lane-list
    # window; in cosine, 120 zero rows give -inf keys to every query
    rng = np.random.default_rng(12)
    n, d, k = 3000, 24, 15
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[1000:1300] = x[17]
    if metric == "cosine":
        x[2000:2120] = 0.0
    ref = O.knn_graph(x, d, k, metric)
    for prec in ["bf16", "fp32", "bf16x3"]:
        res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


def test_margin_and_split_invariance(ctx):
    # the candidate configuration changes only which keys are examined:
    # outputs are identical for every margin / split count
    x = DS.c2(n=5000, d=64)
    ref = H.knn_graph(x, 64, 15, precision="exact", ctx=ctx)
    for margin in [0, 5, 30, 60]:
        for nsplit in [1, 3, 5]:
            job = H.Job(ctx, [5000], [64], [15], ["euclidean"], margin=margin, nsplit=nsplit)
            job.upload(0, x)
            job.run()
            job.sync()
            r = job.download(0)
            job.close()
            assert np.array_equal(r.knn.indices, ref.knn.indices), (margin, nsplit)
            assert np.array_equal(r.graph.vals.view(np.uint32), ref.graph.vals.view(np.uint32))


def test_few_uncertified_rows_use_wide_fallback(ctx):
    # Row 0 has 200 neighbours at distance 1 (up to rounding) and those 200
    # rows are 1.41 +- 0.2 apart: their k-th and C-th selection keys differ
    # by less than the key error bound (which scales with the largest norms
    # of the dataset, here ~60), so these rows alone are not certified and
    # are recomputed by the exact scan. With so few rows the scan spreads
    # them over many column splits (73 here) and merges more than 1024
    # partial keys per row in rounds.
    rng = np.random.default_rng(21)
    n, d, k = 10000, 32, 15
    x = (rng.normal(size=(n, d)) * 3.0 + 10.0).astype(np.float32)
    c = np.zeros(d, np.float32)
    u = rng.normal(size=(200, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x[0] = c
    x[1:201] = (c + u).astype(np.float32)
    ref = O.knn_graph(x, d, k, "euclidean")
    for prec in ["bf16", "fp32"]:
        job = H.Job(ctx, [n], [d], [k], ["euclidean"], precision=prec)
        job.upload(0, x)
        job.run()
        job.sync()
        r = job.download(0)
        fb = job.counters()["fallback_rows"]
        job.close()
        assert_graph_equal(r, *ref)
        # those rows, and the occasional row whose sampled threshold listed
        # too few candidates to certify
        assert 1 <= fb <= 260, fb


def test_escalation_then_exact_scan_for_the_rest(ctx):
    # 20 copies of the pattern above in a larger cloud: ~4,000 rows fail the
    # bf16 bound, too many for the exact scan to be cheaper, so they are
    # re-swept in bf16x3 and re-ranked; the shell rows are then certified,
    # while the 20 centres (200 neighbours at distance 1 up to rounding) are
    # not and go on to the exact scan, which takes its rows from the
    # escalated-row list compacted in place.
    rng = np.random.default_rng(23)
    n, d, k = 30000, 32, 15
    x = (rng.normal(size=(n, d)) * 3.0 + 10.0).astype(np.float32)
    centres = []
    for ci in range(20):
        b = ci * 201
        c = (rng.normal(size=d) * 3.0 + 10.0).astype(np.float32)
        u = rng.normal(size=(200, d))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        x[b] = c
        x[b + 1:b + 201] = (c + u).astype(np.float32)
        centres.append(b)
    out = {}
    for prec in ["bf16", "exact"]:
        job = H.Job(ctx, [n], [d], [k], ["euclidean"], precision=prec)
        job.upload(0, x)
        job.run()
        job.sync()
        out[prec] = (job.download(0), job.counters())
        job.close()
    (r, cnt), (e, _) = out["bf16"], out["exact"]
    assert cnt["escalated_rows"] >= 1000, cnt
    assert 1 <= cnt["fallback_rows"] < cnt["escalated_rows"], cnt
    assert r.knn.indices.tobytes() == e.knn.indices.tobytes()
    assert r.knn.dists.tobytes() == e.knn.dists.tobytes()
    assert r.calib.sigmas.tobytes() == e.calib.sigmas.tobytes()
    assert r.graph.cols.tobytes() == e.graph.cols.tobytes()
    assert r.graph.vals.tobytes() == e.graph.vals.tobytes()
    rows = np.unique(np.concatenate([np.array(centres), centres[3] + np.arange(1, 40),
                                     rng.choice(n, 150, replace=False)])).astype(np.uint32)
    sub = O.exact_knn_rows(x, d, k, "euclidean", rows, nthreads=8)
    assert np.array_equal(r.knn.indices[rows], sub.indices)
    assert np.array_equal(r.knn.dists[rows].view(np.uint32), sub.dists.view(np.uint32))


@pytest.mark.parametrize("calib", ["ref_f32", "f64"])
def test_calibration_by_thread_matches_lane_groups(ctx, calib):
    # A job of >= 1024 rows per CU calibrates with a thread per row (its
    # distances in registers); each dataset alone is small enough for the
    # lane-group kernel. Both fold the same terms in the same order.
    rng = np.random.default_rng(31)
    ks = [10, 15, 30, 15]
    specs = [(7000, 8, ks[i % 4], "euclidean" if i % 3 else "cosine") for i in range(40)]
    assert sum(s[0] for s in specs) >= 1024 * 256
    data = []
    for n, d, _, _ in specs:
        c = rng.normal(size=(12, d)) * 4.0
        data.append((c[rng.integers(0, 12, n)] + rng.normal(size=(n, d))).astype(np.float32))
    data[5][:40] = data[5][3]  # rho = 0 rows: the floor
    job = H.Job(ctx, [s[0] for s in specs], [s[1] for s in specs], [s[2] for s in specs],
                [s[3] for s in specs], calib=calib)
    for i, x in enumerate(data):
        job.upload(i, x)
    job.run()
    job.sync()
    got = [job.download(i, with_sigma64=True) for i in range(len(specs))]
    job.close()
    for i, (n, d, k, m) in enumerate(specs):
        a = H.Job(ctx, [n], [d], [k], [m], calib=calib)
        a.upload(0, data[i])
        a.run()
        a.sync()
        alone = a.download(0, with_sigma64=True)
        a.close()
        r = got[i]
        assert r.knn.indices.tobytes() == alone.knn.indices.tobytes(), i
        assert r.calib.rhos.tobytes() == alone.calib.rhos.tobytes(), i
        assert r.calib.sigmas.tobytes() == alone.calib.sigmas.tobytes(), i
        assert r.calib.sigmas64.tobytes() == alone.calib.sigmas64.tobytes(), i
        assert r.graph.vals.tobytes() == alone.graph.vals.tobytes(), i
    if calib == "ref_f32":
        n, d, k, m = specs[5]
        assert_graph_equal(got[5], *O.knn_graph(data[5], d, k, m))


def test_c4_batch_matches_alone_and_oracle(ctx):
    # 48 datasets of the C4 mix (N log-uniform in [200, 8000], d in
    # {32, 64, 128, 256}, k in {10, 15, 30}, euclidean/cosine) in one job:
    # every dataset's graph equals its single-dataset run byte for byte, and
    # the smallest ones equal the oracle
    sets = DS.c4(count=48)
    job = H.Job(ctx, [s.n for s in sets], [s.d for s in sets], [s.k for s in sets],
                [s.metric for s in sets])
    for i, s in enumerate(sets):
        job.upload(i, s.data)
    job.run()
    job.sync()
    got = [job.download(i) for i in range(len(sets))]
    job.close()
    for i, s in enumerate(sets):
        alone = H.knn_graph(s.data, s.d, s.k, s.metric, ctx=ctx)
        r = got[i]
        assert r.knn.indices.tobytes() == alone.knn.indices.tobytes(), i
        assert r.knn.dists.tobytes() == alone.knn.dists.tobytes(), i
        assert r.calib.sigmas.tobytes() == alone.calib.sigmas.tobytes(), i
        assert r.graph.cols.tobytes() == alone.graph.cols.tobytes(), i
        assert r.graph.vals.tobytes() == alone.graph.vals.tobytes(), i
    for i in sorted(range(len(sets)), key=lambda j: sets[j].n * sets[j].d)[:6]:
        s = sets[i]
        assert_graph_equal(got[i], *O.knn_graph(s.data, s.d, s.k, s.metric))


def _serve_all(srv, ids, ctx=None, budget=0):
    """Drive the loop until every request finished; returns {id: result or exception}."""
    out = {}
    peak = 0
    while len(out) < len(ids):
        srv.progress()
        if ctx is not None:
            peak = max(peak, ctx.in_use)
            if budget:
                assert ctx.in_use <= budget
        for rid in ids:
            if rid in out:
                continue
            st = srv.state(rid)
            if st == srv.DONE or st < 0:
                try:
                    out[rid] = srv.fetch(rid)
                except Exception as e:  # noqa: BLE001 - the request's own error
                    out[rid] = e
    return out, peak


def _same(a, b):
    return (a.knn.indices.tobytes() == b.knn.indices.tobytes()
            and a.knn.dists.tobytes() == b.knn.dists.tobytes()
            and a.calib.sigmas.tobytes() == b.calib.sigmas.tobytes()
            and a.calib.rhos.tobytes() == b.calib.rhos.tobytes()
            and a.graph.rows.tobytes() == b.graph.rows.tobytes()
            and a.graph.cols.tobytes() == b.graph.cols.tobytes()
            and a.graph.vals.tobytes() == b.graph.vals.tobytes()
            and a.indptr.tobytes() == b.indptr.tobytes())


def test_server_results_same_alone_and_in_mix(ctx):
    # 14 mixed requests (C5 shapes, smaller N) submitted together: they are
    # batched into concurrent jobs; every result equals its run alone
    reqs = [DS.c5_request(100 + i, n_min=300, n_max=4000) for i in range(14)]
    srv = H.Server(ctx, max_jobs=3, max_batch=4)
    ids = [srv.submit(r.data, r.d, r.k, r.metric) for r in reqs]
    got, _ = _serve_all(srv, ids)
    srv.close()
    for rid, r in zip(ids, reqs):
        alone = H.knn_graph(r.data, r.d, r.k, r.metric, ctx=ctx)
        assert _same(got[rid], alone), rid


def test_server_budget_admission():
    # a context budget that holds about one request at a time: requests wait
    # for memory, the budget is never exceeded, a request larger than the
    # whole budget fails with MemoryError, the others are exact
    from holomap_amd import _lib
    reqs = [DS.c5_request(200 + i, n_min=500, n_max=3000) for i in range(6)]
    big = DS.c5_request(7, n_min=20000, n_max=20001)
    o = H._opts()
    est = []
    for r in reqs:
        a = np.array([r.n], np.uint64)
        b = np.array([r.d], np.uint32)
        c = np.array([r.k], np.uint32)
        m = np.array([0 if r.metric == "euclidean" else 1], np.int32)
        est.append(int(_lib.load().hm_job_estimate_bytes(1, a.ctypes.data, b.ctypes.data,
                                                          c.ctypes.data, m.ctypes.data,
                                                          ctypes.byref(o))))
    budget = int(max(est) * 1.3)
    bctx = H.Context(0, hbm_budget=budget)
    srv = H.Server(bctx, max_jobs=4, max_batch=8)
    ids = [srv.submit(r.data, r.d, r.k, r.metric) for r in reqs]
    big_id = srv.submit(big.data, big.d, big.k, big.metric)
    got, peak = _serve_all(srv, ids + [big_id], ctx=bctx, budget=budget)
    srv.close()
    assert isinstance(got[big_id], MemoryError)
    assert 0 < peak <= budget
    for rid, r in zip(ids, reqs):
        assert _same(got[rid], H.knn_graph(r.data, r.d, r.k, r.metric))
    bctx.close()


def test_fp32_escalation_certifies_tight_cosine_clusters(ctx):
    # 10 tight clusters (noise 0.5 around centres of norm ~23) under cosine:
    # neighbour similarities differ by less than the bf16 key error bound,
    # so rows are swept again with FP32 keys (enough of them that the
    # re-sweep is cheaper than the exact scan), which certify them; the
    # graph equals the oracle's
    rng = np.random.default_rng(31)
    n, d, k = 20000, 32, 30
    cen = rng.normal(0.0, 4.0, size=(10, d))
    x = (cen[rng.integers(0, 10, n)] + rng.normal(0.0, 0.5, size=(n, d))).astype(np.float32)
    job = H.Job(ctx, [n], [d], [k], ["cosine"])
    job.upload(0, x)
    job.run()
    job.sync()
    r = job.download(0)
    c = job.counters()
    job.close()
    assert_graph_equal(r, *O.knn_graph(x, d, k, "cosine"))
    assert c["escalated_rows"] > 0
    assert c["fallback_rows"] < c["escalated_rows"]


def test_autotune_keeps_outputs(ctx):
    # the autotuner changes the sweep configuration (splits, margin, item
    # order), never the outputs
    x = DS.c2(n=8000, d=64)
    ref = O.knn_graph(x, 64, 15, "euclidean")
    job = H.Job(ctx, [8000], [64], [15], ["euclidean"], autotune=True, use_graph=True)
    job.upload(0, x)
    for _ in range(2):
        job.run()
        job.sync()
        assert_graph_equal(job.download(0), *ref)
    job.close()


def oracle_graph_mt(x, d, k, metric, nthreads=8):
    """The oracle's graph with its kNN rows computed on several threads
    (rows are independent: the same bits as one thread)."""
    n = x.size // d
    knn = O.exact_knn_rows(x, d, k, metric, np.arange(n, dtype=np.uint32), nthreads=nthreads)
    cal = O.smooth_knn(knn.dists, k)
    return knn, cal, O.fuzzy_simplicial_set(knn, cal)


# rows longer than 512 floats: the re-rank reads the query row from global
# memory instead of staging it in LDS
@pytest.mark.parametrize("shape", [(1500, 513, 15), (2000, 640, 30), (2500, 784, 15)])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_long_rows_rerank(ctx, shape, metric):
    n, d, k = shape
    rng = np.random.default_rng(d)
    x = DS.gaussian_mixture(n, d, 12, 1.5, rng)
    x[7] = x[3]  # an exact duplicate: the index tie-break
    ref = oracle_graph_mt(x, d, k, metric)
    for prec in ["bf16", "fp32", "bf16x3"]:
        res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


# finite inputs whose squares leave the f32 range: |x| ~ 1e19 overflows the
# f32 keys and bias (||x||^2 ~ 1e40), |x| ~ 1e-20 underflows the products;
# the reference computes both in f64 (its distances stay finite)
@pytest.mark.parametrize("k", [15, 30])
def test_rerank_many_candidates_with_tied_distances(ctx, k):
    """Rows with more than 64 plausible candidates, most of them tied: 120
    identical rows (distance 0 to each other, the k lowest column indices
    win the ties) and a 70-row shell at one distance from a centre row."""
    rng = np.random.default_rng(31)
    n, d = 700, 16
    x = rng.normal(size=(n, d)).astype(np.float32) * 4.0
    x[100:220] = x[7]                    # 121 identical rows
    c = x[300].copy()
    for r in range(301, 371):            # 70 rows at distance 0.5 from row 300
        v = np.zeros(d, np.float32)
        v[r % d] = 0.5 if (r // d) % 2 == 0 else -0.5
        x[r] = c + v
    for metric in ["euclidean", "cosine"]:
        ref = oracle_graph_mt(x, d, k, metric)
        for prec in ["bf16", "fp32"]:
            res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
            assert_graph_equal(res, *ref)


def test_rerank_distances_at_f32_rounding_midpoints(ctx):
    """Candidate distances whose f64 value lies within 2^-44 of a midpoint
    between two f32 values: the re-rank's fast sums cannot decide their f32
    rounding, so these take the reference's sequential fold."""
    rng = np.random.default_rng(29)
    ncl, per, d, k = 40, 24, 128, 15
    x = np.zeros((ncl * per, d), np.float32)
    near = 0
    for c in range(ncl):
        rows = slice(c * per, (c + 1) * per)
        x[rows, 2:] = rng.integers(-64, 65, size=d - 2) * 8.0  # far-apart centres
        b = rng.uniform(0.5, 2.0, per - 1).astype(np.float32).astype(np.float64)
        ulp = np.spacing(b.astype(np.float32)).astype(np.float64)
        m = b + ulp / 2                      # midpoints between b and the next f32
        v1 = np.sqrt(m * m - b * b).astype(np.float32)
        x[c * per + 1:(c + 1) * per, 0] = b.astype(np.float32)
        x[c * per + 1:(c + 1) * per, 1] = v1
        dist = np.sqrt(b * b + v1.astype(np.float64) ** 2)   # from the centre row
        near += int(np.sum(np.abs(dist - m) < 2.0 ** -44 * m))
    assert near >= ncl * (per - 1) // 2   # most distances to a centre are undecided
    ref = oracle_graph_mt(x, d, k, "euclidean")
    for prec in ["bf16", "fp32"]:
        res = H.knn_graph(x, d, k, "euclidean", precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


@pytest.mark.parametrize("scale", [1e19, 1e-20, "mixed"])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_extreme_magnitudes(ctx, scale, metric):
    rng = np.random.default_rng(17)
    n, d, k = 1200, 24, 15
    x = DS.gaussian_mixture(n, d, 8, 3.0, rng).astype(np.float64)
    if scale == "mixed":
        x[:60] *= 1e19
        x[60:120] *= 1e-20
    else:
        x *= scale
    x = x.astype(np.float32)
    assert np.isfinite(x).all()
    ref = oracle_graph_mt(x, d, k, metric)
    for prec in ["bf16", "fp32"]:
        res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


def _knn_graph_like_rust_shim(ctx, x, d, k, metric):
    """hm_knn_graph called as the Rust shim calls it: host buffers sized as
    the reference's Vecs, cap = 2nk, no row pointer, no FP64 sigmas,
    default options."""
    from holomap_amd import _lib
    L = _lib.load()
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    n = x.size // d
    idx = np.empty(n * k, np.uint32)
    dist = np.empty(n * k, np.float32)
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    cap = 2 * n * k
    rows = np.empty(cap, np.uint32)
    cols = np.empty(cap, np.uint32)
    vals = np.empty(cap, np.float32)
    opt = _lib.Options()
    L.hm_default_options(ctypes.byref(opt))
    o = _lib.GraphOut()
    o.indices, o.dists = idx.ctypes.data, dist.ctypes.data
    o.sigmas, o.rhos = sig.ctypes.data, rho.ctypes.data
    o.rows, o.cols, o.vals = rows.ctypes.data, cols.ctypes.data, vals.ctypes.data
    o.indptr = None
    o.sigmas64 = None
    o.cap = cap
    m = _lib.HM_EUCLIDEAN if metric == "euclidean" else _lib.HM_COSINE
    st = L.hm_knn_graph(ctx.handle, x.ctypes.data, n, d, k, m, ctypes.byref(opt), ctypes.byref(o))
    nnz = int(o.nnz)
    return st, (idx.reshape(n, k), dist.reshape(n, k), sig, rho, rows[:nnz], cols[:nnz],
                vals[:nnz])


def test_hm_knn_graph_c_abi_as_the_rust_shim(ctx):
    cases = []
    for name in FIXTURES:
        fx, data = load_fixture(name)
        cases.append((data, fx["n_features"], fx["n_neighbors"], fx["metric"]))
    cases.append((DS.c1(), 50, 15, "euclidean"))
    cases.append((DS.c1(), 50, 15, "cosine"))
    for data, d, k, metric in cases:
        st, (idx, dist, sig, rho, rows, cols, vals) = _knn_graph_like_rust_shim(ctx, data, d, k,
                                                                                metric)
        assert st == 0, ctx.last_error()
        knn, cal, g = O.knn_graph(data, d, k, metric)
        assert np.array_equal(idx, knn.indices)
        assert np.array_equal(dist.view(np.uint32), knn.dists.view(np.uint32))
        assert np.array_equal(sig.view(np.uint32), cal.sigmas.view(np.uint32))
        assert np.array_equal(rho.view(np.uint32), cal.rhos.view(np.uint32))
        assert np.array_equal(rows, g.rows) and np.array_equal(cols, g.cols)
        assert np.array_equal(vals.view(np.uint32), g.vals.view(np.uint32))
    # the failure path the shim reports: a status and hm_last_error's message
    from holomap_amd import _lib
    x = np.ones((10, 3), np.float32)
    st, _ = _knn_graph_like_rust_shim(ctx, x, 3, 10, "euclidean")
    assert st == _lib.HM_ETOOFEW
    assert ctx.last_error()
    x[4, 1] = np.inf
    st, _ = _knn_graph_like_rust_shim(ctx, x, 3, 5, "euclidean")
    assert st == _lib.HM_ENONFINITE


# n_neighbors beyond the sweep's 64: every row of such a dataset is computed
# by the exact scan (buffers of 512 and 2048 keys per query for k > 64 and
# k > 448), the calibration runs a thread per row, and the symmetrisation's
# row groups take several out-edges per lane
@pytest.mark.parametrize("k", [65, 100, 200, 449, 1000])
def test_large_k_graph(ctx, k):
    rng = np.random.default_rng(k)
    n, d = 1300, 20
    x = DS.gaussian_mixture(n, d, 6, 3.0, rng)
    x[11] = x[10]
    metric = "euclidean" if k % 2 else "cosine"
    ref = oracle_graph_mt(x, d, k, metric)
    for prec in ["bf16", "exact"]:
        res = H.knn_graph(x, d, k, metric, precision=prec, ctx=ctx)
        assert_graph_equal(res, *ref)


def test_large_k_mixed_batch(ctx):
    # a batch mixing k <= 64 datasets (swept) with k > 64 ones (exact scan)
    rng = np.random.default_rng(41)
    specs = [(900, 16, 15, "euclidean"), (1100, 24, 80, "cosine"), (700, 40, 30, "cosine"),
             (1500, 8, 150, "euclidean")]
    data = [DS.gaussian_mixture(n, d, 5, 3.0, rng) for n, d, _, _ in specs]
    job = H.Job(ctx, [s[0] for s in specs], [s[1] for s in specs], [s[2] for s in specs],
                [s[3] for s in specs])
    for i, x in enumerate(data):
        job.upload(i, x)
    job.run()
    job.sync()
    for i, (n, d, k, m) in enumerate(specs):
        assert_graph_equal(job.download(i), *oracle_graph_mt(data[i], d, k, m))
    job.close()


def test_fuzzy_stage_with_large_k(ctx):
    # the stage-level symmetrisation with k = 80 (no kNN stage in the job)
    rng = np.random.default_rng(43)
    n, d, k = 900, 10, 80
    x = DS.gaussian_mixture(n, d, 4, 2.0, rng)
    knn = O.exact_knn_rows(x, d, k, "euclidean", np.arange(n, dtype=np.uint32), nthreads=8)
    cal = O.smooth_knn(knn.dists, k)
    g = O.fuzzy_simplicial_set(knn, cal)
    got = H.fuzzy_simplicial_set(H.Knn(k, knn.indices, knn.dists),
                                 H.SmoothKnn(cal.sigmas, cal.rhos), ctx=ctx)
    assert np.array_equal(got.rows, g.rows) and np.array_equal(got.cols, g.cols)
    assert np.array_equal(got.vals.view(np.uint32), g.vals.view(np.uint32))
    sc = H.smooth_knn(knn.dists, k, ctx=ctx)
    assert np.array_equal(sc.sigmas.view(np.uint32), cal.sigmas.view(np.uint32))


def test_k_beyond_the_limit_is_refused(ctx):
    from holomap_amd import _lib
    x = np.random.default_rng(2).normal(size=(_lib.HM_MAX_NEIGHBORS + 10, 4)).astype(np.float32)
    with pytest.raises(H.DeviceError, match="not supported"):
        H.knn_graph(x, 4, _lib.HM_MAX_NEIGHBORS + 1, ctx=ctx)
    # the serving loop refuses it at submit, so it never fails a batch
    srv = H.Server(ctx)
    with pytest.raises(H.DeviceError):
        srv.submit(x, 4, _lib.HM_MAX_NEIGHBORS + 1)
    ok = srv.submit(x[:300], 4, 15)
    while srv.state(ok) not in (srv.DONE,) and srv.state(ok) >= 0:
        srv.progress()
    r = srv.fetch(ok)
    assert_graph_equal(r, *O.knn_graph(x[:300], 4, 15, "euclidean"))
    srv.close()


def test_server_tight_cosine_budget():
    # a budget that holds one cosine request (by its metric-aware estimate):
    # every request is admitted in turn and none fails for memory
    from holomap_amd import _lib
    reqs = [DS.c5_request(300 + i, n_min=800, n_max=2500) for i in range(5)]
    reqs = [DS.Dataset(r.data / np.maximum(np.linalg.norm(r.data, axis=1, keepdims=True), 1e-30),
                       r.n, r.d, r.k, "cosine") for r in reqs]
    o = H._opts()
    est = []
    for r in reqs:
        a = np.array([r.n], np.uint64)
        b = np.array([r.d], np.uint32)
        c = np.array([r.k], np.uint32)
        m = np.array([_lib.HM_COSINE], np.int32)
        est.append(int(_lib.load().hm_job_estimate_bytes(1, a.ctypes.data, b.ctypes.data,
                                                          c.ctypes.data, m.ctypes.data,
                                                          ctypes.byref(o))))
    budget = max(est) + 4096
    bctx = H.Context(0, hbm_budget=budget)
    srv = H.Server(bctx, max_jobs=2, max_batch=4)
    ids = [srv.submit(r.data.astype(np.float32), r.d, r.k, r.metric) for r in reqs]
    got, peak = _serve_all(srv, ids, ctx=bctx, budget=budget)
    srv.close()
    assert 0 < peak <= budget
    for rid, r in zip(ids, reqs):
        assert not isinstance(got[rid], Exception), got[rid]
        assert _same(got[rid], H.knn_graph(r.data.astype(np.float32), r.d, r.k, r.metric))
    bctx.close()
