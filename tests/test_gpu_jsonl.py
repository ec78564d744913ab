"""GPU parity of the JSON-lines front door: each served graph equals the
oracle's (l2_normalise of pipeline.rs:151-162, then the cosine graph with
k = n_neighbors, src/api.rs:147-149) bit for bit, with fit_transform's
errors in-band between them and every response in request order; and the
device l2_normalise against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import holomap_amd as H
import oracle as O
from holomap_amd import datasets as DS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "holomap_amd", "holomap_gpu_jsonl")


def oracle_graph(x, k):
    n = x.shape[0]
    knn = O.exact_knn_rows(x, x.shape[1], k, "cosine", np.arange(n, dtype=np.uint32), nthreads=8)
    cal = O.smooth_knn(knn.dists, k)
    return knn, cal, O.fuzzy_simplicial_set(knn, cal)


def test_device_l2_normalise(ctx):
    rng = np.random.default_rng(12)
    x = (rng.normal(size=(3000, 96)) * rng.uniform(1e-20, 1e19, size=(3000, 1))).astype(np.float32)
    x[10] = 0.0
    x[11] = 0.0
    x[11, 5] = 7.0
    got = H.l2_normalise(x, ctx=ctx)
    assert got.tobytes() == O.l2_normalise(x).tobytes()


def f32(vals):
    return np.array(vals, np.float64).astype(np.float32)


def test_front_door_graphs_and_errors_in_order():
    rng = np.random.default_rng(13)
    reqs = []
    for i, (n, d, k) in enumerate([(800, 16, 15), (2500, 64, 30), (300, 8, 10), (1200, 32, 15)]):
        x = DS.gaussian_mixture(n, d, 6, 3.0, rng)
        reqs.append(("ok", x, k))
        if i == 1:
            reqs.append(("few", x[:10], 15))
            reqs.append(("k1", x[:50], 1))
    lines = []
    for kind, x, k in reqs:
        lines.append(json.dumps({"protocol_version": 1, "vectors": x.astype(np.float64).tolist(),
                                 "params": {"n_components": 2, "n_neighbors": k,
                                            "min_cluster_size": 5, "seed": 7}}))
    r = subprocess.run([BIN], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    out = [json.loads(s) for s in r.stdout.splitlines()]
    assert len(out) == len(reqs)
    for (kind, x, k), o in zip(reqs, out):
        if kind == "few":
            assert o["error"] == ("holomap reduction failed: need more than n_neighbors=15 "
                                  "points, got 10")
            continue
        if kind == "k1":
            assert o["error"] == ("holomap reduction failed: invalid parameter: n_neighbors "
                                  "must be >= 2")
            continue
        assert "error" not in o, o.get("error")
        xn = O.l2_normalise(x)
        knn, cal, g = oracle_graph(xn, k)
        assert o["n"] == x.shape[0] and o["n_neighbors"] == k
        assert np.array_equal(np.array(o["knn_indices"], np.uint32), knn.indices)
        assert f32(o["knn_dists"]).tobytes() == knn.dists.tobytes()
        assert f32(o["sigmas"]).tobytes() == cal.sigmas.tobytes()
        assert f32(o["rhos"]).tobytes() == cal.rhos.tobytes()
        assert np.array_equal(np.array(o["graph"]["rows"], np.uint32), g.rows)
        assert np.array_equal(np.array(o["graph"]["cols"], np.uint32), g.cols)
        assert f32(o["graph"]["vals"]).tobytes() == g.vals.tobytes()
