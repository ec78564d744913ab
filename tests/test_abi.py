"""Host-side checks that need no GPU: the C-ABI library loads, exports every
symbol include/holomap_gpu.h declares, and the host validation mirrors
src/api.rs:113-144 and src/metric.rs:16-30."""
import ctypes
import os
import re

import numpy as np
import pytest

import holomap_amd as H
from holomap_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "holomap_gpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    L = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_version_and_options():
    L = _lib.load()
    assert b"gfx950" in L.hm_version()
    o = _lib.default_options()
    assert o.precision == _lib.HM_PREC_BF16 and o.calib == _lib.HM_CALIB_REF_F32
    assert o.margin == -1


def test_c_validate():
    L = _lib.load()
    bad = ctypes.c_uint64(0)
    x = np.ones(12, np.float32)
    assert L.hm_validate(x.ctypes.data, 12, 0, 2, ctypes.byref(bad)) == _lib.HM_EBADSHAPE
    assert L.hm_validate(x.ctypes.data, 12, 5, 2, ctypes.byref(bad)) == _lib.HM_EBADSHAPE
    assert L.hm_validate(x.ctypes.data, 12, 3, 1, ctypes.byref(bad)) == _lib.HM_EINVAL
    assert L.hm_validate(x.ctypes.data, 12, 3, 4, ctypes.byref(bad)) == _lib.HM_ETOOFEW
    x[7] = np.nan
    assert L.hm_validate(x.ctypes.data, 12, 3, 2, ctypes.byref(bad)) == _lib.HM_ENONFINITE
    assert bad.value == 7
    x[7] = 1.0
    assert L.hm_validate(x.ctypes.data, 12, 3, 2, ctypes.byref(bad)) == _lib.HM_OK


def test_python_validation_errors():
    x = np.ones(20, np.float32)
    with pytest.raises(H.BadShape):
        H.exact_knn(x, 3, 2)
    with pytest.raises(H.InvalidParameter):
        H.knn_graph(x, 2, 1)
    with pytest.raises(H.TooFewPoints):
        H.knn_graph(x, 2, 10)        # fit_transform rule: n > k
    with pytest.raises(H.TooFewPoints):
        H.exact_knn(x, 2, 11)        # stage rule: n >= k
    y = x.copy()
    y[3] = np.inf
    with pytest.raises(H.NonFiniteInput) as e:
        H.exact_knn(y, 2, 2)
    assert e.value.index == 3
    b = H.Holomap.builder(42).n_neighbors(1)
    with pytest.raises(H.InvalidParameter):
        b.graph(x, 2)
    with pytest.raises(H.InvalidParameter):
        H.Holomap.builder(42).spread(0.0).graph(x, 2)
    with pytest.raises(H.InvalidParameter):
        H.Holomap.builder(42).min_dist(-1.0).graph(x, 2)


def test_metric_parse_and_display():
    # src/metric.rs:81-89
    assert H.Metric.from_str("euclidean") is H.Metric.EUCLIDEAN
    assert H.Metric.from_str("Cosine") is H.Metric.COSINE
    with pytest.raises(H.InvalidParameter):
        H.Metric.from_str("manhattan")
    for m in H.Metric:
        assert H.Metric.from_str(str(m)) is m


def test_estimate_bytes_is_host_only():
    L = _lib.load()
    n = np.array([50000], np.uint64)
    d = np.array([784], np.uint32)
    k = np.array([15], np.uint32)
    o = _lib.default_options()
    eu = np.array([_lib.HM_EUCLIDEAN], np.int32)
    co = np.array([_lib.HM_COSINE], np.int32)
    b = L.hm_job_estimate_bytes(1, n.ctypes.data, d.ctypes.data, k.ctypes.data, eu.ctypes.data,
                                ctypes.byref(o))
    # at least the input, its padded copy and the bf16 operand
    assert b > 50000 * 784 * 10
    # cosine adds the normalised f32 operand (n_pad x kp floats); a NULL
    # metric array is counted as cosine
    bc = L.hm_job_estimate_bytes(1, n.ctypes.data, d.ctypes.data, k.ctypes.data, co.ctypes.data,
                                 ctypes.byref(o))
    assert bc - b == 50048 * 800 * 4
    bn = L.hm_job_estimate_bytes(1, n.ctypes.data, d.ctypes.data, k.ctypes.data, None,
                                 ctypes.byref(o))
    assert bn == bc
