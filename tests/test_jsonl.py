"""The JSON-lines front door (holomap_amd/holomap_gpu_jsonl), the checks and
in-band errors of the clusterer's protocol (crates/holomap-clusterer/src/
main.rs:10-31, pipeline.rs:79-115, protocol.rs:67-93), which need no GPU:
one response line per request line, in order, blank lines skipped, every
failure answered in-band with the reference's message."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "holomap_amd", "holomap_gpu_jsonl")
PARAMS = {"n_components": 2, "n_neighbors": 15, "min_cluster_size": 5, "seed": 42}


def run(lines, timeout=60):
    if not os.path.exists(BIN):
        pytest.skip("front door not built (python -m holomap_amd.build)")
    r = subprocess.run([BIN], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stderr
    return [json.loads(x) for x in r.stdout.splitlines()]


def req(vectors, version=1, **params):
    p = dict(PARAMS)
    p.update(params)
    return json.dumps({"protocol_version": version, "vectors": vectors, "params": p})


def test_protocol_errors_in_band_and_in_order():
    out = run([req([[0.1]], version=2), "", "   ", "{not json", req([]),
               req([[0.1, 0.2], [0.3]]), req([[], []]),
               json.dumps({"protocol_version": 1, "vectors": [[1.0]]}),
               req([[1.0, 2.0], [3.0, 4.0]], n_components=0)])
    assert len(out) == 7
    assert all(o["protocol_version"] == 1 for o in out)
    assert out[0]["error"] == "unsupported protocol_version 2"
    assert out[1]["error"].startswith("bad request: ")
    assert out[2]["error"] == "empty input"
    assert out[3]["error"] == "vector dimensions inconsistent"  # protocol.rs tests
    assert out[4]["error"] == "zero-dimensional vectors"
    assert out[5]["error"] == "bad request: missing field `params`"
    for o in out[:6]:
        assert o["assignments"] == [] and "probabilities" not in o
    # no reduction: no graph
    assert "error" not in out[6] and out[6]["n"] == 2 and out[6]["graph"]["rows"] == []


def test_field_types_are_checked():
    bad = [json.dumps({"protocol_version": "1", "vectors": [[1.0]], "params": PARAMS}),
           json.dumps({"protocol_version": 1, "vectors": [[1.0, "x"]], "params": PARAMS}),
           json.dumps({"protocol_version": 1, "vectors": [[1.0]],
                       "params": dict(PARAMS, n_neighbors=-1)}),
           json.dumps({"protocol_version": 1, "vectors": [[1.0]],
                       "params": {k: v for k, v in PARAMS.items() if k != "seed"}})]
    out = run(bad)
    assert len(out) == 4
    for o in out:
        assert o["error"].startswith("bad request: ")
    assert "seed" in out[3]["error"]
