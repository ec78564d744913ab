"""Pins the CPU oracle (oracle/hm_oracle.c) before it is trusted as the
checker of the GPU path.

Each oracle function is checked against something that would catch a
plausible mistake (dropped term, wrong sign/index, transposed operand):
  * the reference's golden fixtures (tests/golden/stage2_*.json),
  * the known answers of the reference's unit tests,
  * an independent numpy restatement / brute force on tiny inputs,
  * mathematical invariants (expf error bound vs. correctly rounded exp,
    the calibration equation, t-conorm symmetry).
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = ["blobs_euclidean", "cosine", "duplicates_euclidean"]


def load_fixture(name):
    with open(os.path.join(GOLD, f"stage2_{name}.json")) as f:
        fx = json.load(f)
    data = np.array(fx["data"], np.float64).astype(np.float32)
    return fx, data


def close(a, e):
    # src/fixture_parity.rs:18-24
    a = np.asarray(a, np.float64)
    e = np.asarray(e, np.float64)
    return np.all(np.abs(a - e) <= 1e-5 * (1.0 + np.abs(e)))


# ---------------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_knn_bit_exact(name):
    fx, data = load_fixture(name)
    knn = O.exact_knn(data, fx["n_features"], fx["n_neighbors"], fx["metric"])
    assert np.array_equal(knn.indices, np.array(fx["knn_indices"], np.uint32))
    # the fixture distances are float32 values written as decimals
    assert np.array_equal(knn.dists, np.array(fx["knn_dists"], np.float64).astype(np.float32))


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_calibration(name):
    fx, data = load_fixture(name)
    k = fx["n_neighbors"]
    knn = O.exact_knn(data, fx["n_features"], k, fx["metric"])
    cal = O.smooth_knn(knn.dists, k)
    assert np.array_equal(cal.rhos, np.array(fx["rhos"], np.float64).astype(np.float32))
    assert close(cal.sigmas, fx["sigmas"])


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_graph(name):
    fx, data = load_fixture(name)
    k = fx["n_neighbors"]
    knn, cal, g = O.knn_graph(data, fx["n_features"], k, fx["metric"])
    assert np.array_equal(g.rows, np.array(fx["graph_rows"], np.uint32))
    assert np.array_equal(g.cols, np.array(fx["graph_cols"], np.uint32))
    assert close(g.vals, fx["graph_vals"])


# ---------------------------------------------------------------- known answers
def test_metric_known_values():
    # src/metric.rs:91-101
    assert O.euclidean([0.0, 0.0], [3.0, 4.0]) == 5.0
    assert O.euclidean([1.0, 2.0, 3.0], [1.0, 2.0, 3.0]) == 0.0
    assert O.cosine([1.0, 0.0], [0.0, 1.0]) == 1.0
    assert O.cosine([1.0, 0.0], [2.0, 0.0]) == 0.0
    assert O.cosine([1.0, 0.0], [-3.0, 0.0]) == 2.0
    assert O.cosine([0.0, 0.0], [1.0, 2.0]) == 0.0  # NaN.max(0.0) = 0


def test_knn_known_answers():
    # src/knn.rs:64-104
    line = np.array([0.0, 1.0, 3.0, 7.0], np.float32)
    knn = O.exact_knn(line, 1, 3, "euclidean")
    assert knn.indices[1].tolist() == [1, 0, 2]
    assert knn.dists[1].tolist() == [0.0, 1.0, 2.0]
    assert knn.indices[3].tolist() == [3, 2, 1]
    assert knn.dists[3].tolist() == [0.0, 4.0, 6.0]
    dup = O.exact_knn(np.array([5.0, 9.0, 5.0], np.float32), 1, 2, "euclidean")
    assert dup.indices[0].tolist() == [0, 2]
    assert dup.indices[2].tolist() == [0, 2]
    assert dup.dists[2].tolist() == [0.0, 0.0]
    c = O.exact_knn(np.array([0.3, 0.7, 0.31, 0.69], np.float32), 2, 2, "cosine")
    assert c.dists[0, 0] == 0.0 and c.dists[1, 0] == 0.0
    assert c.dists[0, 1] > 0.0 and c.dists[1, 1] > 0.0
    a = O.exact_knn(np.array([1.0, 0.0, 100.0, 1.0, 0.1, 0.1], np.float32), 2, 3, "cosine")
    assert a.indices[0].tolist() == [0, 1, 2]


def test_smooth_knn_known_answers():
    # src/smooth_knn.rs:103-124: exact float32 values of the reference
    cal = O.smooth_knn(np.array([0, 1, 3, 0, 1, 2, 0, 2, 3], np.float32), 3)
    assert cal.rhos.tolist() == [1.0, 1.0, 2.0]
    assert cal.sigmas.tolist() == [3.72998046875, 1.864990234375, 1.864990234375]
    z = O.smooth_knn(np.zeros(4, np.float32), 2)
    assert z.rhos.tolist() == [0.0, 0.0] and z.sigmas.tolist() == [1.0, 1.0]


def test_fuzzy_known_answers():
    # src/fuzzy.rs:111-153
    knn = O.Knn(3, np.array([0, 1, 2, 1, 0, 2, 2, 1, 0], np.uint32).reshape(3, 3),
                np.array([0, 1, 3, 0, 1, 2, 0, 2, 3], np.float32).reshape(3, 3))
    cal = O.SmoothKnn(np.array([3.72998046875, 1.864990234375, 1.864990234375], np.float32),
                      np.array([1, 1, 2], np.float32))
    g = O.fuzzy_simplicial_set(knn, cal)
    w = O.expf(np.float32(-2.0) / np.float32(3.72998046875))
    fused = np.float32(np.float32(w + w) - np.float32(w * w))
    assert g.rows.tolist() == [0, 0, 1, 1, 2, 2]
    assert g.cols.tolist() == [1, 2, 0, 2, 0, 1]
    assert np.allclose(g.vals, [1.0, fused, 1.0, 1.0, fused, 1.0], atol=1e-6)
    knn2 = O.Knn(2, np.array([0, 1, 1, 2, 2, 1], np.uint32).reshape(3, 2),
                 np.array([0, 1, 0, 0.1, 0, 0.1], np.float32).reshape(3, 2))
    cal2 = O.SmoothKnn(np.ones(3, np.float32), np.array([0.5, 0.1, 0.1], np.float32))
    g2 = O.fuzzy_simplicial_set(knn2, cal2)
    w2 = O.expf(-0.5)
    assert g2.rows.tolist() == [0, 1, 1, 2]
    assert g2.cols.tolist() == [1, 0, 2, 1]
    assert g2.vals.tolist() == [w2, w2, 1.0, 1.0]


def test_csr_known_answer():
    # src/sparse.rs:117-124 (path graph 0-1-2)
    ip = O.csr_indptr(np.array([0, 1, 1, 2], np.uint32), 3)
    assert ip.tolist() == [0, 1, 3, 4]


# ---------------------------------------------------------------- independent restatements
def np_euclid(a, b):
    s = 0.0
    for x, y in zip(a.astype(np.float64), b.astype(np.float64)):
        t = x - y
        s = s + t * t
    return np.float32(np.sqrt(s))


def np_cosine(a, b):
    dot = na = nb = 0.0
    for x, y in zip(a.astype(np.float64), b.astype(np.float64)):
        dot += x * y
        na += x * x
        nb += y * y
    with np.errstate(invalid="ignore", divide="ignore"):
        v = 1.0 - dot / (np.sqrt(na) * np.sqrt(nb))
    return np.float32(0.0 if np.isnan(v) else max(v, 0.0))


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_brute_force_tiny(metric):
    rng = np.random.default_rng(5)
    n, d, k = 24, 7, 6
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[5] = x[2]          # exact duplicate: tie broken by index
    x[9] = 0.0           # zero vector (cosine: distance 0 to all)
    x[11] = 3.0 * x[4]   # parallel (cosine distance 0)
    f = np_euclid if metric == "euclidean" else np_cosine
    knn = O.exact_knn(x, d, k, metric)
    for i in range(n):
        dist = np.array([0.0 if j == i else f(x[i], x[j]) for j in range(n)], np.float32)
        order = np.lexsort((np.arange(n), dist))[:k]
        assert knn.indices[i].tolist() == order.tolist()
        assert np.array_equal(knn.dists[i], dist[order])


def test_knn_rows_matches_full():
    rng = np.random.default_rng(6)
    x = rng.normal(size=(300, 9)).astype(np.float32)
    full = O.exact_knn(x, 9, 8, "euclidean")
    rows = np.array([0, 17, 299, 150], np.uint32)
    sub = O.exact_knn_rows(x, 9, 8, "euclidean", rows, nthreads=2)
    assert np.array_equal(sub.indices, full.indices[rows])
    assert np.array_equal(sub.dists, full.dists[rows])


def test_expf_error_bound_and_specials():
    # The libm/FreeBSD method is documented < 1 ulp; a wrong coefficient,
    # constant or reduction step breaks this bound by orders of magnitude.
    rng = np.random.default_rng(7)
    bits = rng.integers(0x80000000, 0xC2D00000, size=20000, dtype=np.uint64).astype(np.uint32)
    xs = bits.view(np.float32)
    xs = np.concatenate([xs, np.linspace(-103.9, 0, 2001, dtype=np.float32)])
    L = O.lib()
    worst = 0.0
    for x in xs:
        y = np.float32(L.hmo_expf(float(x)))
        ref = np.exp(np.float64(x))
        if ref < 1.2e-38:  # subnormal results: absolute check
            assert abs(float(y) - ref) <= 1.5e-45
            continue
        ulp = np.spacing(np.float32(ref))
        worst = max(worst, abs(float(y) - ref) / float(ulp))
    assert worst < 1.0
    assert O.expf(0.0) == 1.0
    assert O.expf(-200.0) == 0.0
    # src/fmath.rs:155-161
    for x in [0.0, -1.0, 1.0, -7.5]:
        assert abs(float(O.expf(x)) - np.exp(x)) < 1e-6


def test_smooth_knn_solves_calibration_equation():
    rng = np.random.default_rng(8)
    x = rng.normal(size=(400, 6)).astype(np.float32)
    k = 10
    knn = O.exact_knn(x, 6, k, "euclidean")
    cal = O.smooth_knn(knn.dists, k)
    c64 = O.smooth_knn_f64(knn.dists, k)
    target = np.log2(k)
    for i in range(400):
        dd = knn.dists[i, 1:].astype(np.float64) - c64.rhos[i]
        ps = np.where(dd > 0, np.exp(-dd / c64.sigmas[i]), 1.0).sum()
        assert abs(ps - target) < 1e-9
    # the reference's float32 search stops within 1e-5 of the target
    rel = np.abs(cal.sigmas - c64.sigmas) / c64.sigmas
    assert rel.max() < 1e-3
    assert np.array_equal(cal.rhos.astype(np.float64), c64.rhos)


def test_fuzzy_invariants():
    rng = np.random.default_rng(9)
    x = rng.normal(size=(500, 5)).astype(np.float32)
    knn, cal, g = O.knn_graph(x, 5, 12, "euclidean")
    key = g.rows.astype(np.int64) * 500 + g.cols
    assert np.all(np.diff(key) > 0)                 # sorted, no duplicates
    assert np.all(g.rows != g.cols)                 # no self edges
    assert np.all((g.vals > 0) & (g.vals <= 1))
    m = {(int(r), int(c)): v for r, c, v in zip(g.rows, g.cols, g.vals)}
    for (r, c), v in m.items():
        assert m[(c, r)] == v                       # bitwise symmetric
    ip = O.csr_indptr(g.rows, 500)
    assert ip[-1] == g.rows.size and np.all(np.diff(ip.astype(np.int64)) >= 0)


# ---------------------------------------------------------------- components, Laplacian
def sym_csr(n, edges, w=None):
    """Symmetric, (row, col)-sorted CSR of an undirected edge list (the
    reference's test helper shape, src/components.rs:47-58)."""
    coo = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    rows = np.array([r for r, _ in coo], np.uint32)
    cols = np.array([c for _, c in coo], np.uint32)
    vals = np.ones(len(coo), np.float32) if w is None else np.array(
        [w[(min(r, c), max(r, c))] for r, c in coo], np.float32)
    return O.csr_indptr(rows, n), cols, vals


def test_components_known_answers():
    # src/components.rs:60-83
    c = O.connected_components(*sym_csr(3, [(0, 1), (1, 2)])[:2], 3)
    assert c.n_components == 1 and c.labels.tolist() == [0, 0, 0]
    c = O.connected_components(*sym_csr(4, [(0, 3), (1, 2)])[:2], 4)
    assert c.n_components == 2 and c.labels.tolist() == [0, 1, 1, 0]
    c = O.connected_components(*sym_csr(3, [(0, 1)])[:2], 3)
    assert c.n_components == 2 and c.labels.tolist() == [0, 0, 1]


def test_components_against_scipy_and_min_vertex_rank():
    # first-seen labels = rank of each component's smallest vertex; the
    # partition itself from scipy's connected_components
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components as cc
    rng = np.random.default_rng(3)
    for n, m in [(50, 30), (400, 300), (1000, 2000)]:
        e = rng.integers(0, n, size=(m, 2))
        e = [(int(a), int(b)) for a, b in e if a != b]
        ip, ix, _ = sym_csr(n, e)
        got = O.connected_components(ip, ix, n)
        a = sp.csr_matrix((np.ones(ix.size), ix.astype(np.int64), ip.astype(np.int64)), shape=(n, n))
        k, lab = cc(a, directed=False)
        assert got.n_components == k
        mins = {}
        for v in range(n):
            mins.setdefault(lab[v], v)
        order = {c: r for r, c in enumerate(sorted(mins, key=lambda c: mins[c]))}
        assert got.labels.tolist() == [order[lab[v]] for v in range(n)]


def p3():
    # src/sparse.rs:104-111: path 0-1-2, unit weights
    return sym_csr(3, [(0, 1), (1, 2)])


def test_laplacian_known_answers():
    # src/sparse.rs:113-146
    ip, ix, w = p3()
    assert ip.tolist() == [0, 1, 3, 4]
    s, d = O.laplacian_new(ip, ix, w, 3)
    assert d.tolist() == [1.0, 2.0, 1.0]
    y = O.laplacian_matvec(ip, ix, s, 3, np.array([1.0, np.sqrt(2.0), 1.0]))
    assert np.all(np.abs(y) < 1e-12)  # the trivial eigenvector D^1/2 1
    v1 = np.array([1.0, 0.0, -1.0])  # eigenvalue 1
    assert np.all(np.abs(O.laplacian_matvec(ip, ix, s, 3, v1) - v1) < 1e-12)
    # dense I - S against the product on basis vectors (src/sparse.rs:148-162)
    dense = np.eye(3)
    for i in range(3):
        for p in range(ip[i], ip[i + 1]):
            dense[i, ix[p]] -= s[p]
    for j in range(3):
        e = np.zeros(3)
        e[j] = 1.0
        assert np.all(np.abs(dense[:, j] - O.laplacian_matvec(ip, ix, s, 3, e)) < 1e-15)


def test_laplacian_against_scipy_and_spectrum():
    # S = D^-1/2 W D^-1/2 on a weighted symmetric graph: the product within
    # 1e-13 of scipy's, L's spectrum in [0, 2] with the trivial eigenvalue 0
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    n = 300
    e = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(1500, 2)) if a != b]
    e += [(i, i + 1) for i in range(n - 1)]  # connected
    w = {(min(a, b), max(a, b)): np.float32(rng.uniform(0.05, 1.0)) for a, b in e}
    ip, ix, vals = sym_csr(n, e, w)
    s, d = O.laplacian_new(ip, ix, vals, n)
    W = sp.csr_matrix((vals.astype(np.float64), ix.astype(np.int64), ip.astype(np.int64)),
                      shape=(n, n))
    deg = np.asarray(W.sum(1)).ravel()
    assert np.allclose(d, deg, rtol=1e-13, atol=0)
    Dm = sp.diags(1.0 / np.sqrt(deg))
    S = (Dm @ W @ Dm).tocsr()
    S.sort_indices()
    assert np.allclose(s, S.data, rtol=1e-13, atol=0)
    x = rng.normal(size=n)
    assert np.allclose(O.laplacian_matvec(ip, ix, s, n, x), x - S @ x, rtol=1e-12, atol=1e-12)
    ev = np.linalg.eigvalsh(np.eye(n) - S.toarray())
    assert ev.min() > -1e-10 and ev.max() < 2.0 + 1e-10 and abs(ev[0]) < 1e-10


def test_laplacian_isolated_vertex_degree_is_negative_zero():
    # Rust's f64 Sum starts at -0.0: an empty row has degree -0.0
    ip, ix, w = sym_csr(3, [(0, 1)])
    s, d = O.laplacian_new(ip, ix, w, 3)
    assert d[2] == 0.0 and np.signbit(d[2])
    x = np.array([1.0, 2.0, -0.0])
    y = O.laplacian_matvec(ip, ix, s, 3, x)
    assert y[2] == 0.0 and not np.signbit(y[2])  # -0.0 - (-0.0) = +0.0


def test_schedule_edges_known_answers():
    # src/sgd.rs:168-189, 230-242
    g = O.FuzzyGraph(4, np.array([0, 1, 2], np.uint32), np.array([1, 2, 3], np.uint32),
                     np.array([1.0, 0.5, 0.001], np.float32))
    s = O.schedule_edges(g, 500)
    assert s.head.tolist() == [0, 1] and s.tail.tolist() == [1, 2]
    assert s.epochs_per_sample.tolist() == [1.0, 2.0]
    g = O.FuzzyGraph(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32),
                     np.array([1.0, 0.01], np.float32))
    assert O.schedule_edges(g, 10).head.size == 2
    # n_epochs <= 10 prunes by the default schedule: 500 up to 10,000 points,
    # 200 above (src/sgd.rs:23-25): 0.004 survives 1/500, not 1/200
    for n, kept in [(100, 2), (10000, 2), (10001, 1)]:
        g = O.FuzzyGraph(n, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32),
                         np.array([1.0, 0.004], np.float32))
        assert O.schedule_edges(g, 5).head.size == kept, n


def test_schedule_edges_against_numpy():
    rng = np.random.default_rng(9)
    m = 5000
    g = O.FuzzyGraph(3000, rng.integers(0, 3000, m).astype(np.uint32),
                     rng.integers(0, 3000, m).astype(np.uint32),
                     (rng.random(m) ** 6).astype(np.float32))
    for ne in [3, 50, 500]:
        s = O.schedule_edges(g, ne)
        mx = np.float32(g.vals.max())
        prune = ne if ne > 10 else 500
        keep = ~(g.vals < np.float32(mx / np.float32(prune)))
        assert np.array_equal(s.head, g.rows[keep]) and np.array_equal(s.tail, g.cols[keep])
        assert np.array_equal(s.epochs_per_sample, np.float64(mx) / g.vals[keep].astype(np.float64))


def test_l2_normalise_against_numpy_sequential():
    # crates/holomap-clusterer/src/pipeline.rs:151-162: an f32 left fold of
    # the squares (numpy's cumsum is sequential), sqrt, f32 division; zero
    # rows unchanged
    rng = np.random.default_rng(4)
    x = rng.normal(size=(200, 37)).astype(np.float32) * np.float32(3.0)
    x[5] = 0.0
    x[6, :] = 0.0
    x[6, 3] = -2.5
    got = O.l2_normalise(x)
    for i in range(x.shape[0]):
        s = np.cumsum(x[i] * x[i], dtype=np.float32)[-1]
        norm = np.sqrt(s, dtype=np.float32)
        want = x[i] if norm == 0 else x[i] / norm
        assert got[i].tobytes() == want.astype(np.float32).tobytes(), i
    assert got[5].tobytes() == x[5].tobytes()
    assert got[6, 3] == -1.0
    assert O.l2_normalise(np.array([[3.0, 4.0]], np.float32)).tolist() == [[0.6000000238418579,
                                                                            0.800000011920929]]


# ---- the FP64 chain (oracle of the GPU's FP64 calibration mode) ----
def np_euclid64(a, b):
    s = 0.0
    for x, y in zip(a.astype(np.float64), b.astype(np.float64)):
        t = x - y
        s = s + t * t
    return np.sqrt(s)


def np_cosine64(a, b):
    dot = na = nb = 0.0
    for x, y in zip(a.astype(np.float64), b.astype(np.float64)):
        dot += x * y
        na += x * x
        nb += y * y
    with np.errstate(invalid="ignore", divide="ignore"):
        v = 1.0 - dot / (np.sqrt(na) * np.sqrt(nb))
    return 0.0 if np.isnan(v) else max(v, 0.0)


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_pair_dist64_is_the_unrounded_fold(metric):
    rng = np.random.default_rng(15)
    n, d, k = 40, 9, 7
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[5] = x[2]
    x[9] = 0.0
    knn = O.exact_knn(x, d, k, metric)
    d64 = O.pair_dist64(x, d, knn.indices, metric)
    # rounds to the kNN's f32 distances bit for bit
    assert np.array_equal(d64.astype(np.float32).view(np.uint32), knn.dists.view(np.uint32))
    # and is the sequential f64 fold itself (the row itself at 0)
    f = np_euclid64 if metric == "euclidean" else np_cosine64
    for i in range(n):
        for t in range(k):
            j = int(knn.indices[i, t])
            assert d64[i, t] == (0.0 if j == i else f(x[i], x[j]))


def test_smooth_knn_f64d_on_f64_distances():
    rng = np.random.default_rng(16)
    x = rng.normal(size=(300, 5)).astype(np.float32)
    x[40:45] = x[39]  # rows with zero distances
    k = 12
    knn = O.exact_knn(x, 5, k, "euclidean")
    d64 = O.pair_dist64(x, 5, knn.indices)
    c = O.smooth_knn_f64(d64, k)
    # on f32-representable inputs it is the widening variant bit for bit
    w = O.smooth_knn_f64(knn.dists.astype(np.float64), k)
    assert np.array_equal(w.sigmas, O.smooth_knn_f64(knn.dists, k).sigmas)
    # rho: the first positive f64 distance; sigma solves the equation in f64
    for i in range(300):
        pos = d64[i][d64[i] > 0]
        assert c.rhos[i] == (pos[0] if pos.size else 0.0)
        dd = d64[i, 1:] - c.rhos[i]
        with np.errstate(over="ignore"):
            ps = np.where(dd > 0, np.exp(-dd / c.sigmas[i]), 1.0).sum()
        floor = 1e-3 * d64[i].mean()
        assert abs(ps - np.log2(k)) < 1e-9 or c.sigmas[i] == pytest.approx(floor, rel=1e-12)
    # a mistake planted in the distances moves sigma
    bad = d64.copy()
    bad[:, 2] *= 1.001
    assert not np.array_equal(O.smooth_knn_f64(bad, k).sigmas, c.sigmas)


def test_fuzzy_f64_strengths_rounded_once():
    rng = np.random.default_rng(17)
    x = rng.normal(size=(200, 6)).astype(np.float32)
    k = 8
    knn = O.exact_knn(x, 6, k, "euclidean")
    d64 = O.pair_dist64(x, 6, knn.indices)
    c = O.smooth_knn_f64(d64, k)
    g = O.fuzzy_simplicial_set_f64(knn.indices, d64, c.sigmas, c.rhos)
    # numpy restatement: f64 strengths, rounded to f32, f32 t-conorm
    dd = d64 - c.rhos[:, None]
    w = np.where((dd <= 0) | (c.sigmas[:, None] == 0), 1.0, np.exp(-(dd / c.sigmas[:, None])))
    w32 = w.astype(np.float32)
    A = {}
    for i in range(200):
        for t in range(k):
            j = int(knn.indices[i, t])
            if j != i and w32[i, t] != 0:
                A[(i, j)] = w32[i, t]
    keys = sorted(set(A) | {(c_, r) for (r, c_) in A})
    exp_rows = np.array([r for r, _ in keys], np.uint32)
    exp_cols = np.array([c_ for _, c_ in keys], np.uint32)
    a = np.array([A.get(p, np.float32(0)) for p in keys], np.float32)
    b = np.array([A.get((p[1], p[0]), np.float32(0)) for p in keys], np.float32)
    exp_vals = (a + b) - a * b
    assert np.array_equal(g.rows, exp_rows) and np.array_equal(g.cols, exp_cols)
    # numpy's exp and libm's may differ by an ulp of f64: at most one f32 step apart
    assert np.max(np.abs(g.vals - exp_vals) / exp_vals) < 2.4e-7
