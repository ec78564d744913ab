"""The bench's FP64 error statement (CPU only).

bench.sigma_f64 is a numpy restatement of the FP64 bisection; fed the same
f32 distances it must agree with the oracle's hmo_smooth_knn_f64 (a
sequential C loop) to summation-order noise, including rows whose distances
are all zero (rho = 0, the all-rows floor) and rows with duplicates. The
whole statement, run on the oracle's own graph, must report distances within
the f32 rounding bound and weights whose support matches the FP64 union's.
"""
import os
import sys
from types import SimpleNamespace as NS

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402
import oracle as O  # noqa: E402
from holomap_amd import datasets as DS  # noqa: E402


@pytest.mark.parametrize("seed,k", [(0, 5), (1, 15), (2, 30)])
def test_sigma_f64_matches_fp64_oracle(seed, k):
    rng = np.random.default_rng(seed)
    x = DS.gaussian_mixture(700, 12, 5, 3.0, rng)
    x[100:140] = x[99]  # a block of duplicates: rows with several zero distances
    knn = O.exact_knn(x, 12, k)
    c64 = O.smooth_knn_f64(knn.dists, k)
    rho, sig = bench.sigma_f64(knn.dists.astype(np.float64), k)
    assert np.array_equal(rho, c64.rhos)
    assert (np.abs(sig - c64.sigmas) / c64.sigmas).max() < 1e-12


def test_sigma_f64_all_zero_rows_take_the_global_floor():
    d = np.zeros((4, 5))
    d[2:] = np.sort(np.random.default_rng(3).uniform(0.5, 2.0, (2, 5)), axis=1)
    d = d.astype(np.float32).astype(np.float64)  # the oracle reads f32 distances
    rho, sig = bench.sigma_f64(d, 5)
    c64 = O.smooth_knn_f64(d.astype(np.float32), 5)
    assert np.array_equal(rho, c64.rhos)
    assert np.allclose(sig, c64.sigmas, rtol=1e-12, atol=0)


def test_fp64_error_on_the_oracle_graph():
    x = DS.c1()
    k = 15
    knn, cal, g = O.knn_graph(x, 50, k, "euclidean")
    c64 = O.smooth_knn_f64(knn.dists, k)
    res = NS(knn=knn, calib=NS(sigmas=cal.sigmas, rhos=cal.rhos, sigmas64=c64.sigmas), graph=g)
    e = bench.fp64_error(x, k, res)
    assert e["rows"] == 1000 and e["edges"] == g.vals.size
    # f64 distances correctly rounded to f32
    assert e["distance_max_rel"] <= 2.0 ** -24 * (1 + 1e-9)
    assert e["edges_without_fp64_support"] == 0
    # the f32 chain's weights stay close to the FP64 chain's
    assert e["weight_max_abs"] < 1e-4
    assert e["sigma_max_rel"] < 1e-3 and e["sigma64_max_rel"] < 1e-3
