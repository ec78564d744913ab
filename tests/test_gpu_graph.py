"""GPU parity of the graph's first consumers (src/components.rs:13-39,
src/sparse.rs:55-93) against the oracle, bit for bit: component labels in
first-seen order, the normalised adjacency S and the degrees, and L x, on
the reference's fixture graphs, on graphs the GPU pipeline built (one and
several components) and on an isolated vertex; the job-level forms on the
resident graph equal the host-buffer forms."""
import json
import os

import numpy as np
import pytest

import holomap_amd as H
import oracle as O
from holomap_amd import datasets as DS
from hip_util import Hip

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def fixture_graph(name):
    with open(os.path.join(GOLD, f"stage2_{name}.json")) as f:
        fx = json.load(f)
    n = np.array(fx["data"]).reshape(-1, fx["n_features"]).shape[0]
    return H.FuzzyGraph(n, np.array(fx["graph_rows"], np.uint32),
                        np.array(fx["graph_cols"], np.uint32),
                        np.array(fx["graph_vals"], np.float64).astype(np.float32))


def islands(seed, sizes, d=8, gap=1e4):
    """Clusters far apart: the kNN graph (k < each size) splits into them."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(size=(m, d)) + gap * i for i, m in enumerate(sizes)]
    x = np.concatenate(parts).astype(np.float32)
    return x[rng.permutation(x.shape[0])]


def graphs(ctx):
    out = [("fixture-" + nm, fixture_graph(nm)) for nm in
           ["blobs_euclidean", "cosine", "duplicates_euclidean"]]
    out.append(("c1", H.knn_graph(DS.c1(), 50, 15, ctx=ctx).graph))
    out.append(("islands", H.knn_graph(islands(1, [300, 40, 500, 25, 100]), 8, 10, ctx=ctx).graph))
    out.append(("c2-8k", H.knn_graph(DS.c2(n=8000, d=64), 64, 15, ctx=ctx).graph))
    # an isolated vertex (no edges): degree -0.0, its own component
    out.append(("isolated", H.FuzzyGraph(4, np.array([0, 1, 1, 2], np.uint32),
                                         np.array([1, 0, 2, 1], np.uint32),
                                         np.array([0.5, 0.5, 1.0, 1.0], np.float32))))
    return out


def test_components_and_laplacian_vs_oracle(ctx):
    rng = np.random.default_rng(2)
    seen_multi = False
    for name, g in graphs(ctx):
        csr = H.Csr.from_fuzzy(g, ctx=ctx)
        cc = H.connected_components(csr, ctx=ctx)
        ref = O.connected_components(csr.indptr, csr.indices, g.n)
        assert cc.n_components == ref.n_components, name
        assert np.array_equal(cc.labels, ref.labels), name
        seen_multi |= cc.n_components > 1
        lap = H.Laplacian.new(g, ctx=ctx)
        s, d = O.laplacian_new(csr.indptr, csr.indices, g.vals, g.n)
        assert lap.s.vals.tobytes() == s.tobytes(), name
        assert lap.degrees.tobytes() == d.tobytes(), name
        for _ in range(2):
            x = rng.normal(size=g.n)
            y = lap.matvec(x, ctx=ctx)
            assert y.tobytes() == O.laplacian_matvec(csr.indptr, csr.indices, s, g.n, x).tobytes()
    assert seen_multi


def test_job_resident_graph_consumers(ctx):
    # a batch of three datasets; components and L x on each resident graph
    # equal the host-buffer forms on the downloaded graph
    sets = [(islands(3, [200, 60, 300]), 8, 12, "euclidean"),
            (DS.c2(n=3000, d=32), 32, 15, "cosine"),
            (islands(4, [500, 500]), 8, 30, "euclidean")]
    job = H.Job(ctx, [x.shape[0] for x, *_ in sets], [s[1] for s in sets],
                [s[2] for s in sets], [s[3] for s in sets])
    for i, (x, *_) in enumerate(sets):
        job.upload(i, x)
    job.run()
    job.sync()
    hip = Hip()
    rng = np.random.default_rng(6)
    try:
        for i, (x, d, k, m) in enumerate(sets):
            r = job.download(i)
            g = r.graph
            csr = H.Csr(g.n, r.indptr, g.cols, g.vals.astype(np.float64))
            ref = O.connected_components(r.indptr, g.cols, g.n)
            cc = job.connected_components(i)
            assert cc.n_components == ref.n_components and np.array_equal(cc.labels, ref.labels)
            s, _ = O.laplacian_new(r.indptr, g.cols, g.vals, g.n)
            x = rng.normal(size=g.n)
            dx, dy = hip.alloc(x.nbytes), hip.alloc(x.nbytes)
            hip.h2d(dx, x)
            job.laplacian_matvec(i, dx, dy)
            y = np.empty(g.n, np.float64)
            hip.d2h(y, dy)
            assert y.tobytes() == O.laplacian_matvec(r.indptr, g.cols, s, g.n, x).tobytes()
            assert csr.n == g.n
    finally:
        hip.close()
        job.close()


def test_csr_input_checks(ctx):
    bad = H.Csr(3, np.array([0, 1, 2, 3], np.uint64), np.array([1, 0, 7], np.uint32),
                np.ones(3))
    with pytest.raises(H.InvalidParameter):
        H.connected_components(bad, ctx=ctx)


def test_schedule_edges_vs_oracle(ctx):
    gs = graphs(ctx)
    for name, g in gs:
        for ne in [5, 200, 500, 2000]:
            s = H.schedule_edges(g, ne, ctx=ctx)
            r = O.schedule_edges(O.FuzzyGraph(g.n, g.rows, g.cols, g.vals), ne)
            assert s.head.tobytes() == r.head.tobytes(), (name, ne)
            assert s.tail.tobytes() == r.tail.tobytes(), (name, ne)
            assert s.epochs_per_sample.tobytes() == r.epochs_per_sample.tobytes(), (name, ne)
    # the resident graph of a job: the same schedule as from its download
    x = DS.c2(n=4000, d=32)
    job = H.Job(ctx, [4000], [32], [15], ["euclidean"])
    job.upload(0, x)
    job.run()
    job.sync()
    g = job.download(0).graph
    for ne in [7, 300]:
        s = job.schedule_edges(0, ne)
        r = O.schedule_edges(O.FuzzyGraph(g.n, g.rows, g.cols, g.vals), ne)
        assert s.head.tobytes() == r.head.tobytes() and s.tail.tobytes() == r.tail.tobytes()
        assert s.epochs_per_sample.tobytes() == r.epochs_per_sample.tobytes()
    job.close()
