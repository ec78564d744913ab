"""kNN-graph construction throughput on one MI355X.

One step = the whole graph build of one dataset with the input already
resident in HBM: exact kNN (BF16 MFMA candidate sweep + f64 re-rank +
certification + exact fallback), smooth-kNN calibration and fuzzy-union
symmetrisation into COO + CSR row pointer. Default workload: N=20,000 x
d=128 float32, a mixture of 50 anisotropic Gaussians with 2 % duplicated
rows, euclidean, k=15 (synthetic, seeded). The plain FP32 GPU path is timed
beside it, and the reference's CPU path (the oracle, single thread) on a
bounded sample.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN-graph pts/s BF16+rerank vs FP32 (bit-exact idx); p50/p99 fit latency"
PEAK = {"bf16": 2.5e15, "bf16x3": 2.5e15, "fp32": 157.3e12}
HBM_PEAK = 8.0e12  # bytes/s


def workload(name):
    from holomap_amd import datasets as DS
    if name == "c2":
        return DS.c2(), 128, 15, "euclidean", (
            "N=20000 d=128 k=15 euclidean; 50 anisotropic Gaussians (centres N(0,9I), "
            "diagonal scales U[0.5,2]), 2% exact duplicates, seed 7")
    if name == "c3":
        return DS.c3(), 784, 15, "euclidean", (
            "N=50000 d=784 k=15 euclidean; 28x28 image-like rows, ~20% non-zero, "
            "values k/255, margin 2k")
    if name == "c1":
        return DS.c1(), 50, 15, "euclidean", "N=1000 d=50 k=15 euclidean; 10 Gaussians, seed 42"
    raise SystemExit(f"unknown workload {name}")


def run_path(ctx, x, d, k, metric, precision, steps, warmup):
    import holomap_amd as H
    n = x.shape[0]
    # per-stage device times: HIP events on the job's stream around each stage
    # of eager runs of an identically configured (autotuned) job; run first,
    # so the timed job's kernels are the last ones of the path
    tj = H.Job(ctx, [n], [d], [k], [metric], precision=precision, timing=True, autotune=True)
    tj.upload(0, x)
    tj.run()
    tj.sync()
    reps = 10
    acc = None
    for _ in range(reps):
        tj.run()
        tj.sync()
        t = tj.times()
        acc = t if acc is None else {kk: acc[kk] + t[kk] for kk in acc}
    tj.close()
    stage = {kk: v / reps for kk, v in acc.items()}
    # timed loop: the whole pipeline replayed from a captured HIP graph; the
    # first (untimed) run autotunes the sweep configuration
    job = H.Job(ctx, [n], [d], [k], [metric], precision=precision, use_graph=True, autotune=True)
    job.upload(0, x)
    for _ in range(max(warmup, 1)):
        job.run()
    job.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        job.run()
    job.sync()
    dt = (time.perf_counter() - t0) / steps
    res = job.download(0)
    counters = job.counters()
    job.close()
    return dt, stage, counters, res


def cpu_baseline(x, d, k, metric, res, budget_s=12.0):
    import oracle as O
    n = x.shape[0]
    rng = np.random.default_rng(0)
    # calibrate the sample size on 20 rows, then time ~budget_s of kNN rows
    probe = rng.choice(n, 20, replace=False).astype(np.uint32)
    t0 = time.perf_counter()
    O.exact_knn_rows(x, d, k, metric, probe, nthreads=1)
    per = (time.perf_counter() - t0) / 20
    m = int(max(20, min(n, budget_s / max(per, 1e-6))))
    rows = rng.choice(n, m, replace=False).astype(np.uint32)
    t0 = time.perf_counter()
    sub = O.exact_knn_rows(x, d, k, metric, rows, nthreads=1)
    t_knn = time.perf_counter() - t0
    assert np.array_equal(sub.indices, res.knn.indices[rows])
    t0 = time.perf_counter()
    cal = O.smooth_knn(res.knn.dists, k)
    O.fuzzy_simplicial_set(O.Knn(k, res.knn.indices, res.knn.dists), cal)
    t_rest = time.perf_counter() - t0
    per_point = t_knn / m + t_rest / n
    return {"value": 1.0 / per_point, "unit": "pts/s", "cores": 1, "kind": "port",
            "sample": f"oracle exact_knn on {m} random query rows x all {n} columns "
                      f"({t_knn:.1f} s, 1 thread) + full smooth_knn/fuzzy on the GPU's kNN "
                      f"({t_rest:.2f} s); pts/s = 1/(t_knn/{m} + t_rest/{n})"}


def sweep_traffic(workload_name, precision):
    """HBM bytes per sweep launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload_name}_{precision}_gram_sweep_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def target_check(ctx, steps):
    """The N=50,000 envelope workload: BF16 vs FP32 GPU path and the CPU port."""
    x, d, k, metric, desc = workload("c3")
    n = x.shape[0]
    dt, stage, cnt, res = run_path(ctx, x, d, k, metric, "bf16", steps, 2)
    dt32, st32, c32, res32 = run_path(ctx, x, d, k, metric, "fp32", max(3, steps // 2), 1)
    same = bool(np.array_equal(res32.knn.indices, res.knn.indices) and
                np.array_equal(res32.graph.vals.view(np.uint32), res.graph.vals.view(np.uint32)))
    cb = cpu_baseline(x, d, k, metric, res)
    return {"workload": f"c3: {desc}", "bf16_ms_per_step": dt * 1e3, "bf16_pts_per_s": n / dt,
            "fp32_ms_per_step": dt32 * 1e3, "speedup_vs_fp32": dt32 / dt,
            "identical_graph": same, "fallback_rows": cnt["fallback_rows"],
            "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
            "cpu_baseline": cb, "speedup_vs_cpu": (n / dt) / cb["value"]}


def batch_bench(ctx, steps):
    """C4: 256 datasets (N log-uniform in [200, 8000], d in {32, 64, 128,
    256}, k in {10, 15, 30}, euclidean/cosine) in one job: one launch per
    stage for the whole batch, replayed from a captured hipGraph."""
    import holomap_amd as H
    from holomap_amd import datasets as DS
    sets = DS.c4()
    job = H.Job(ctx, [s.n for s in sets], [s.d for s in sets], [s.k for s in sets],
                [s.metric for s in sets], use_graph=True, autotune=True)
    for i, s in enumerate(sets):
        job.upload(i, s.data)
    t0 = time.perf_counter()
    job.run()  # the first run autotunes the sweep configuration, then captures
    job.sync()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        job.run()
    job.sync()
    dt = (time.perf_counter() - t0) / steps
    cnt = job.counters()
    job.close()
    rows = sum(s.n for s in sets)
    return {"workload": "c4: 256 datasets, N log-uniform [200,8000], d in {32,64,128,256}, "
                        "k in {10,15,30}, metric 50/50, seed 11",
            "datasets": len(sets), "rows": rows, "ms_per_step": dt * 1e3,
            "graphs_per_s": len(sets) / dt, "pts_per_s": rows / dt,
            "fallback_rows": cnt["fallback_rows"], "escalated_rows": cnt["escalated_rows"],
            "autotune_s": t_tune, "tuned_nsplit": cnt["nsplit"]}


def serve_bench(ctx, lambdas, n_req, pool=16, seed=0):
    """Serving loop (hm_server_*): Poisson arrivals at each rate, requests of
    the mixed C5 shapes (N log-uniform in [500, 50000], d in {64, 384, 768},
    k in {15, 30}, Gaussian-mixture euclidean or unit-normalised cosine),
    drawn from a pool of `pool` seeded datasets. Latency = arrival -> result
    on the host (fetch done)."""
    import holomap_amd as H
    from holomap_amd import datasets as DS
    reqs = [DS.c5_request(1000 + i) for i in range(pool)]
    rng = np.random.default_rng(seed)
    out = []
    for lam in lambdas:
        srv = H.Server(ctx, max_jobs=4, max_batch=16)
        arr = np.cumsum(rng.exponential(1.0 / lam, n_req))
        pick = rng.integers(0, pool, n_req)
        lat, pend, pts = [], {}, 0
        i = 0
        t0 = time.perf_counter()
        while len(lat) < n_req:
            now = time.perf_counter() - t0
            while i < n_req and arr[i] <= now:
                r = reqs[pick[i]]
                pend[srv.submit(r.data, r.d, r.k, r.metric)] = (arr[i], r.n)
                i += 1
            srv.progress()
            for rid in list(pend):
                if srv.state(rid) == srv.DONE:
                    srv.fetch(rid)
                    a, n = pend.pop(rid)
                    lat.append(time.perf_counter() - t0 - a)
                    pts += n
            if not pend and i < n_req:
                time.sleep(max(0.0, min(arr[i] - (time.perf_counter() - t0), 0.002)))
        span = time.perf_counter() - t0
        srv.close()
        lat = np.array(lat)
        out.append({"lambda_req_s": lam, "requests": n_req,
                    "p50_ms": float(np.percentile(lat, 50) * 1e3),
                    "p99_ms": float(np.percentile(lat, 99) * 1e3),
                    "throughput_req_s": n_req / span, "throughput_pts_s": pts / span})
    return {"workload": "c5 serving: Poisson arrivals, N log-uniform [500,50000], d in "
                        "{64,384,768}, k in {15,30}, euclidean mixture / cosine unit rows; "
                        f"{pool} pooled seeded datasets; max_jobs 4, max_batch 16",
            "rates": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="bf16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the N=50,000 (d=784) BF16 vs FP32 vs CPU check")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop figures")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-dataset batch figures")
    ap.add_argument("--serve-rates", default="50,200",
                    help="Poisson arrival rates (req/s) of the serving figures")
    ap.add_argument("--serve-requests", type=int, default=120,
                    help="requests per arrival rate")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB (reads and writes that "
                         "many bytes once) to calibrate the HBM byte counters, then exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe * (1 << 20)}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the kNN-graph path here runs on a single device "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import holomap_amd as H
    x, d, k, metric, desc = workload(args.workload)
    n = x.shape[0]
    ctx = H.Context(0)
    # the boundary's host-buffer form (hm_knn_graph: X from host memory in,
    # kNN, sigma/rho and the COO back to host memory, job set-up included);
    # measured first, so that the timed job's sweeps are the run's last ones
    H.knn_graph(x, d, k, metric, precision=args.precision, ctx=ctx)  # warm the memory pool
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        H.knn_graph(x, d, k, metric, precision=args.precision, ctx=ctx)
    t_host = (time.perf_counter() - t0) / reps
    dt, stage, cnt, res = run_path(ctx, x, d, k, metric, args.precision, args.steps, args.warmup)
    out = {
        "metric": METRIC,
        "value": n / dt,
        "unit": "pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "N": n, "d": d, "k": k,
                   "metric": metric, "precision": args.precision, "candidates_per_row": cnt["candidates"],
                   "nsplit": cnt["nsplit"], "autotune": True, "parallelism": "single GPU"},
        "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
        "fallback_rows": cnt["fallback_rows"],
        "nnz": cnt["nnz"],
    }
    # achieved bandwidth of the other stages against their algorithmic bytes
    # (HIP-event stage times of the eager runs; these stages are bound by
    # dependent memory round trips, not by HBM bandwidth):
    #  top-k (re-rank): the listed keys (C per row, 8 bytes), the k winners'
    #    rows and the query row (d floats each), the k outputs (idx + dist);
    #  calibration: k distances in, sigma, rho and the f64 sigma out;
    #  symmetrisation: kNN idx + dist and sigma/rho in, the COO (row, col,
    #    value) and the row pointer out.
    C = cnt["candidates"]
    sb = {"topk": (n * (C * 8 + (k + 1) * d * 4 + k * 8), "rerank"),
          "calibration": (n * (k * 4 + 16), "calib"),
          "symmetrisation": (n * k * 8 + n * 8 + cnt["nnz"] * 12 + (n + 1) * 8, "fuzzy")}
    out["stage_bandwidth"] = {}
    for name, (nbytes, st) in sb.items():
        t = stage[st] * 1e-3
        gbs = nbytes / t / 1e9 if t > 0 else 0.0
        out["stage_bandwidth"][name] = {"bytes": int(nbytes), "ms": round(stage[st], 4),
                                        "GB/s": round(gbs, 1), "frac_of_hbm": gbs * 1e9 / HBM_PEAK}
    # roofline of the dominant kernel, the candidate sweep (MFMA-bound). It is
    # launched twice per step (threshold pass, collect pass); each launch does
    # the algorithmic 2*N*N*d flops of one Gram matrix. The sweep stage time is
    # HIP events on the job's stream around both launches (eager runs).
    flops = 2.0 * n * n * d
    t_launch = stage["sweep"] * 1e-3 / 2
    peak = PEAK.get(args.precision, PEAK["bf16"])
    ach = flops / t_launch if t_launch > 0 else 0.0
    # bf16 rows of up to 384 bytes (d <= 192) go to the query-resident sweep
    qres = args.precision == "bf16" and 2 * (-(-d // 64) * 64) <= 384
    kname = "hm::gram_sweep_qres_kernel" if qres else "hm::gram_sweep_kernel"
    out["roofline"] = {"bound": "mfma", "kernel": f"{kname} (2 launches/step)",
                       "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                       "frac": ach / peak, "traffic": sweep_traffic(args.workload, args.precision),
                       "flops_per_launch": flops, "launch_ms": t_launch * 1e3}
    out["pcie_inclusive"] = {"ms_per_call": t_host * 1e3, "pts_per_s": n / t_host,
                             "entry": "hm_knn_graph (host buffers in and out, per-call job)"}
    if not args.no_fp32 and args.precision != "fp32":
        dt32, st32, c32, res32 = run_path(ctx, x, d, k, metric, "fp32", max(3, args.steps // 2), 2)
        same = bool(np.array_equal(res32.knn.indices, res.knn.indices) and
                    np.array_equal(res32.graph.vals.view(np.uint32), res.graph.vals.view(np.uint32)))
        out["fp32_path"] = {"ms_per_step": dt32 * 1e3, "pts_per_s": n / dt32,
                            "sweep_ms": st32["sweep"], "fallback_rows": c32["fallback_rows"],
                            "identical_graph": same}
        out["speedup_vs_fp32"] = dt32 / dt
    if not args.no_cpu:
        cb = cpu_baseline(x, d, k, metric, res)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    if not args.no_target and args.workload != "c3":
        out["target_n50k"] = target_check(ctx, 10)
    if not args.no_batch:
        out["batch_c4"] = batch_bench(ctx, 10)
    if not args.no_serve:
        rates = [float(v) for v in args.serve_rates.split(",") if v]
        out["serving"] = serve_bench(ctx, rates, args.serve_requests)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
