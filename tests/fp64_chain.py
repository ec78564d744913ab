"""Check of a result of the FP64 calibration mode against the oracle's FP64
chain (shared by the GPU tests)."""
import numpy as np
import pytest

import oracle as O


def check_fp64_chain(r, x, d, k, metric):
    """The f64 distances (the fold the f32 ones round) bit for bit, rho and
    sigma from them in f64, the graph from f64 strengths rounded once, and
    the library's own error statement against the host's."""
    ref64 = O.pair_dist64(x, d, r.knn.indices, metric)
    assert np.array_equal(r.knn.dists64, ref64)
    assert np.array_equal(r.knn.dists, r.knn.dists64.astype(np.float32))
    c64 = O.smooth_knn_f64(r.knn.dists64, k)
    assert np.array_equal(r.calib.rhos, c64.rhos.astype(np.float32))
    # device exp and libm exp may differ in the last bit of f64
    assert (np.abs(r.calib.sigmas64 - c64.sigmas) / c64.sigmas).max() < 1e-12
    assert np.array_equal(r.calib.sigmas, r.calib.sigmas64.astype(np.float32))
    g = O.fuzzy_simplicial_set_f64(r.knn.indices, r.knn.dists64, c64.sigmas, c64.rhos)
    assert np.array_equal(r.graph.rows, g.rows) and np.array_equal(r.graph.cols, g.cols)
    assert (np.abs(r.graph.vals - g.vals) / g.vals).max() < 2.4e-7
    # the statement: the largest relative error of the f32 outputs against
    # their f64 values, as the host computes it from the returned arrays
    e = r.err64
    assert e is not None
    nz = r.knn.dists64 > 0
    host_d = (np.abs(r.knn.dists.astype(np.float64)[nz] - r.knn.dists64[nz]) /
              r.knn.dists64[nz]).max() if nz.any() else 0.0
    host_s = (np.abs(r.calib.sigmas.astype(np.float64) - r.calib.sigmas64) /
              r.calib.sigmas64).max()
    assert e[0] == pytest.approx(float(np.float32(host_d)), rel=1e-6, abs=1e-30)
    assert e[1] == pytest.approx(float(np.float32(host_s)), rel=1e-6, abs=1e-30)
    # f32 rounding of normal values; subnormal strengths round coarser
    assert e[0] <= 2.0 ** -24 * 1.0001 and e[1] <= 2.0 ** -24 * 1.0001
    assert e[2] <= 2.0 ** -24 * 1.0001 or g.vals.min() < 1.2e-38
    return e
