"""kNN-graph construction throughput on one MI355X.

One step = the whole graph build of one dataset with the input already
resident in HBM: exact kNN (BF16 MFMA candidate sweep + f64 re-rank +
certification + exact fallback), smooth-kNN calibration and fuzzy-union
symmetrisation into COO + CSR row pointer. Default workload: N=20,000 x
d=128 float32, a mixture of 50 anisotropic Gaussians with 2 % duplicated
rows, euclidean, k=15 (synthetic, seeded). The plain FP32 GPU path is timed
beside it, and the reference's CPU path (the oracle, one thread and all the
process's cores) on a bounded sample. Further legs: the N=50,000 x d=784
image-like workload (BF16 vs FP32 vs CPU, FP64 calibration and its error
against an FP64 restatement of the chain), the 256-dataset batch and the
serving loop under Poisson arrivals at six rates.

Prints one JSON line on stdout; progress lines go to stderr.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# a fixed sweep configuration instead of the autotuner's (--nsplit), for
# profiles whose every dispatch has the same configuration
FIXED = {"nsplit": 0}


def run_path(ctx, x, d, k, metric, precision, steps, warmup, margin=-1, calib="ref_f32"):
    """margin >= 0 pins the candidate count C = k + margin (the autotuner then
    only picks column splits and item order); -1 lets it also try k, 2k, 3k."""
    import holomap_amd as H
    n = x.shape[0]
    kw = dict(precision=precision, margin=margin, calib=calib, autotune=FIXED["nsplit"] == 0,
              nsplit=FIXED["nsplit"])
    # per-stage device times: HIP events on the job's stream around each stage
    # of eager runs of an identically configured (autotuned) job; run first,
    # so the timed job's kernels are the last ones of the path
    tj = H.Job(ctx, [n], [d], [k], [metric], timing=True, **kw)
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
    job = H.Job(ctx, [n], [d], [k], [metric], use_graph=True, **kw)
    job.upload(0, x)
    for _ in range(max(warmup, 1)):
        job.run()
    job.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        job.run()
    job.sync()
    dt = (time.perf_counter() - t0) / steps
    res = job.download(0, with_sigma64=calib == "f64")
    counters = job.counters()
    job.close()
    return dt, stage, counters, res


def sweep_config(cnt, n, k):
    """The sweep configuration a run used (dataset 0)."""
    return {"C_target": cnt["target"], "margin": cnt["target"] - k,
            "listed_keys_per_row": round(cnt["listed"] / n, 2),
            "list_slots_per_row": cnt["candidates"], "nsplit": cnt["nsplit"],
            "sample_stride": cnt["sample_stride"]}


def sweep_roofline(n, d, precision, stage, workload_name):
    """Roofline of the dominant kernel, the candidate sweep's collect pass
    (every query row against every column: the algorithmic 2*N*N*d flops of
    one Gram matrix per launch). Its duration is the sweep stage minus the
    threshold pass (HIP events on the job's stream, eager runs). The
    threshold pass sweeps a sample of the columns; per_step_frac counts both
    passes against the one Gram a step needs."""
    flops = 2.0 * n * n * d
    t_collect = (stage["sweep"] - stage["threshold"]) * 1e-3
    t_sweep = stage["sweep"] * 1e-3
    peak = PEAK.get(precision, PEAK["bf16"])
    ach = flops / t_collect if t_collect > 0 else 0.0
    rowbytes = 2 * (-(-d // 64) * 64)  # bf16 operand row
    if precision == "bf16" and rowbytes <= 256:
        kname = "hm::gram_sweep_qreg_kernel<1>"
    elif precision == "bf16" and rowbytes <= 384:
        kname = "hm::gram_sweep_qres_kernel<hm::QresCfgT<8, 3, 384>, 1>"
    elif precision == "bf16":
        kname = "hm::gram_sweep_kernel<0, 1, 256>"
    else:
        kname = f"hm::gram_sweep_kernel<{ {'fp32': 1, 'bf16x3': 2}.get(precision, 1)}, 1, 128>"
    return {"bound": "mfma", "kernel": f"{kname} (collect pass, 1 launch/step)",
            "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": ach / peak, "traffic": sweep_traffic(workload_name, precision),
            "flops_per_launch": flops, "launch_ms": t_collect * 1e3,
            "threshold_pass_ms": stage["threshold"],
            "per_step_frac": (flops / t_sweep) / peak if t_sweep > 0 else 0.0}


def stage_bandwidth(n, d, k, cnt, stage):
    """Achieved bandwidth of the other stages against their algorithmic
    (unique) bytes, HIP-event stage times of the eager runs. These stages are
    bound by dependent memory round trips, not by HBM bandwidth.
      top-k (re-rank): the dataset (every candidate row is a row of X), the
        listed keys (8 bytes each), the k outputs (index + distance);
      calibration: k distances in; sigma, rho, the f64 sigma and the
        directed edges (key, payload, weight, source row: 16 bytes) out;
      symmetrisation: the directed edges in, the COO (row, col, value) and
        the row pointer out."""
    E = n * k
    sb = {"topk": (n * d * 4 + cnt["listed"] * 8 + E * 8, "rerank"),
          "calibration": (E * 4 + n * 16 + E * 16, "calib"),
          "symmetrisation": (E * 16 + cnt["nnz"] * 12 + (n + 1) * 8, "fuzzy")}
    out = {}
    for name, (nbytes, st) in sb.items():
        t = stage[st] * 1e-3
        gbs = nbytes / t / 1e9 if t > 0 else 0.0
        out[name] = {"bytes": int(nbytes), "ms": round(stage[st], 4), "GB/s": round(gbs, 1),
                     "frac_of_hbm": gbs * 1e9 / HBM_PEAK}
    return out


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_baseline(x, d, k, metric, res, budget_s=12.0, all_core_s=6.0):
    """The oracle (a C restatement of the reference's CPU path) timed on the
    host: kNN on a bounded sample of query rows (all N columns each; rows are
    independent), the calibration and symmetrisation in full. One thread,
    then every core of the process's CPU share (at most 16) on another
    sample."""
    import oracle as O
    n = x.shape[0]
    rng = np.random.default_rng(0)
    model, nproc, affinity = cpu_info()
    threads = max(1, min(16, affinity))

    def sample_rate(nthreads, budget):
        # calibrate the sample size on a probe, then time ~budget s of rows
        probe = rng.choice(n, 20 * nthreads, replace=False).astype(np.uint32)
        t0 = time.perf_counter()
        O.exact_knn_rows(x, d, k, metric, probe, nthreads=nthreads)
        per = (time.perf_counter() - t0) / probe.size
        m = int(max(20 * nthreads, min(n, budget / max(per, 1e-7))))
        rows = rng.choice(n, m, replace=False).astype(np.uint32)
        t0 = time.perf_counter()
        sub = O.exact_knn_rows(x, d, k, metric, rows, nthreads=nthreads)
        t = time.perf_counter() - t0
        assert np.array_equal(sub.indices, res.knn.indices[rows])
        assert np.array_equal(sub.dists.view(np.uint32), res.knn.dists[rows].view(np.uint32))
        return m, t

    m1, t1 = sample_rate(1, budget_s)
    t0 = time.perf_counter()
    cal = O.smooth_knn(res.knn.dists, k)
    O.fuzzy_simplicial_set(O.Knn(k, res.knn.indices, res.knn.dists), cal)
    t_rest = time.perf_counter() - t0
    mt, tt = sample_rate(threads, all_core_s)
    one = 1.0 / (t1 / m1 + t_rest / n)
    # the calibration and symmetrisation stay serial in the reference
    allc = 1.0 / (tt / mt + t_rest / n)
    return {"value": one, "unit": "pts/s", "cores": 1, "kind": "port",
            "sample": f"oracle exact_knn on {m1} random query rows x all {n} columns "
                      f"({t1:.1f} s, 1 thread) + full smooth_knn/fuzzy on the GPU's kNN "
                      f"({t_rest:.2f} s); pts/s = 1/(t_knn/{m1} + t_rest/{n}); the sampled "
                      f"rows' indices and distance bits equal the GPU's",
            "cpu_model": model, "nproc": nproc, "affinity_cpus": affinity,
            "all_core": {"value": allc, "cores": threads,
                         "sample": f"oracle exact_knn on {mt} rows with {threads} threads "
                                   f"({tt:.1f} s) + the same serial smooth_knn/fuzzy"}}


def sigma_f64(dist, k):
    """rho and sigma of src/smooth_knn.rs:39-77 in f64 for every row of the
    (n, k) f64 distances: rho = the first positive distance (0 if none),
    sigma = the bisection on sum_{t>=1} exp(-(d_t - rho) / sigma) = log2(k)
    from [0, inf) starting at 1, run until the midpoint stops moving, then
    floored at 1e-3 x the row's mean (the all-rows mean when rho = 0)."""
    n = dist.shape[0]
    pos = dist > 0
    rho = np.where(pos.any(1), dist[np.arange(n), np.argmax(pos, 1)], 0.0)
    target = np.log2(k)
    dd = dist[:, 1:] - rho[:, None]
    lo, hi, mid = np.zeros(n), np.full(n, np.inf), np.ones(n)
    act = np.ones(n, bool)
    for _ in range(2200):
        a = np.flatnonzero(act)
        if a.size == 0:
            break
        m, da = mid[a], dd[a]
        with np.errstate(over="ignore", invalid="ignore"):
            psum = np.where(da > 0, np.exp(-(da / m[:, None])), 1.0).sum(1)
        eq, gt = psum == target, psum > target
        nlo = np.where(gt, lo[a], m)
        nhi = np.where(gt, m, hi[a])
        nmid = np.where(gt | ~np.isinf(nhi), 0.5 * (nlo + nhi), m * 2.0)
        stop = eq | (nmid == m) | np.isinf(nmid)
        lo[a], hi[a] = nlo, nhi
        mid[a] = np.where(stop, m, nmid)
        act[a[stop]] = False
    floor = np.where(rho > 0, 1e-3 * dist.mean(1), 1e-3 * dist.mean())
    return rho, np.where(mid < floor, floor, mid)


def fp64_error(x, k, res):
    """Distance, sigma and edge-weight error of a result against an FP64
    restatement of the chain computed here in numpy: f64 distances of the
    result's neighbour pairs, the bisection of src/smooth_knn.rs:39-77 in f64
    on them (run until the interval stops shrinking), f64 membership
    strengths (src/fuzzy.rs:36-41) and the f64 fuzzy union
    a + b - a * b at the result's coordinates."""
    import scipy.sparse as sp
    n = x.shape[0]
    idx = res.knn.indices.astype(np.int64)
    d64 = np.empty((n, k))
    for s0 in range(0, n, 1024):
        e0 = min(n, s0 + 1024)
        diff = x[idx[s0:e0]].astype(np.float64) - x[s0:e0, None, :].astype(np.float64)
        d64[s0:e0] = np.sqrt(np.einsum("ijk,ijk->ij", diff, diff))
    d32 = res.knn.dists.astype(np.float64)
    nz = d64 > 0
    dist_rel = float(np.max(np.abs(d32[nz] - d64[nz]) / d64[nz])) if nz.any() else 0.0
    assert np.all(d32[~nz] == 0)
    rho, sig64 = sigma_f64(d64, k)
    srel = np.abs(res.calib.sigmas.astype(np.float64) - sig64) / sig64
    out = {"rows": n, "distance_max_rel": dist_rel, "sigma_max_rel": float(srel.max())}
    if res.calib.sigmas64 is not None:
        out["sigma64_max_rel"] = float((np.abs(res.calib.sigmas64 - sig64) / sig64).max())
    # directed f64 strengths, then the union at the result's coordinates
    ddr = d64 - rho[:, None]
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        w = np.where((ddr <= 0) | (sig64[:, None] == 0), 1.0, np.exp(-(ddr / sig64[:, None])))
    keep = (idx != np.arange(n)[:, None]) & (w != 0)
    r = np.repeat(np.arange(n), k)[keep.ravel()]
    A = sp.csr_matrix((w.ravel()[keep.ravel()], (r, idx.ravel()[keep.ravel()])), shape=(n, n))
    S = (A + A.T - A.multiply(A.T)).tocsr()
    v64 = np.asarray(S[res.graph.rows.astype(np.int64), res.graph.cols.astype(np.int64)]).ravel()
    pos = v64 > 0
    v32 = res.graph.vals.astype(np.float64)
    out.update({"edges": int(v32.size), "edges_without_fp64_support": int((~pos).sum()),
                "weight_max_rel": float((np.abs(v32[pos] - v64[pos]) / v64[pos]).max()),
                "weight_max_abs": float(np.abs(v32[pos] - v64[pos]).max())})
    return out


def sweep_traffic(workload_name, precision):
    """HBM bytes per collect-pass launch from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload_name}_gram_sweep_traffic.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{precision}]:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def target_check(ctx, steps, with_cpu=True):
    """The N=50,000 envelope workload as configured: BF16 candidates with
    margin 2k (pinned: the autotuner picks only column splits and item
    order), FP64 smooth-kNN calibration, the FP32 GPU path and the CPU port
    beside it, and the result's error against an FP64 restatement."""
    x, d, k, metric, desc = workload("c3")
    n = x.shape[0]
    log("c3: bf16 path")
    dt, stage, cnt, res = run_path(ctx, x, d, k, metric, "bf16", steps, 2, margin=2 * k,
                                   calib="f64")
    log("c3: fp32 path")
    dt32, st32, c32, res32 = run_path(ctx, x, d, k, metric, "fp32", max(3, steps // 2), 1,
                                      margin=2 * k, calib="f64")
    same = bool(np.array_equal(res32.knn.indices, res.knn.indices) and
                np.array_equal(res32.graph.vals.view(np.uint32), res.graph.vals.view(np.uint32)))
    log("c3: fp64 error")
    out = {"workload": f"c3: {desc}; FP64 calibration", "bf16_ms_per_step": dt * 1e3,
           "bf16_pts_per_s": n / dt, "sweep": sweep_config(cnt, n, k),
           "fp32_ms_per_step": dt32 * 1e3, "fp32_sweep": sweep_config(c32, n, k),
           "speedup_vs_fp32": dt32 / dt, "identical_graph": same,
           "fallback_rows": cnt["fallback_rows"], "escalated_rows": cnt["escalated_rows"],
           "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
           "roofline": sweep_roofline(n, d, "bf16", stage, "c3"),
           "fp32_roofline": sweep_roofline(n, d, "fp32", st32, "c3"),
           "stage_bandwidth": stage_bandwidth(n, d, k, cnt, stage),
           # the library's own statement (largest relative error of the f32
           # distances, sigmas and directed strengths against its f64 values)
           # and an independent FP64 chain computed here in numpy
           "fp64_error": {"library_statement": dict(zip(["distance", "sigma", "strength"],
                                                        res.err64 or (None,) * 3)),
                          **fp64_error(x, k, res)}}
    if with_cpu:
        log("c3: cpu baseline")
        cb = cpu_baseline(x, d, k, metric, res)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = (n / dt) / cb["value"]
        out["speedup_vs_cpu_all_core"] = (n / dt) / cb["all_core"]["value"]
    return out


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


def serve_bench(ctx, plan, pool=32, seed=0):
    """Serving loop (hm_server_*): Poisson arrivals at each rate of `plan`
    ((rate req/s, requests) pairs), requests of the mixed shapes (N
    log-uniform in [500, 50000], d in {64, 384, 768}, k in {15, 30},
    Gaussian-mixture euclidean or unit-normalised cosine), drawn at random
    from a pool of `pool` seeded datasets. Latency = arrival -> result on the
    host (fetch done): upload, queueing, device work and download."""
    import holomap_amd as H
    from holomap_amd import datasets as DS
    reqs = [DS.c5_request(1000 + i) for i in range(pool)]
    rng = np.random.default_rng(seed)
    out = []
    for lam, n_req in plan:
        log(f"serving: {n_req} requests at {lam:g} req/s")
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
                    "max_ms": float(lat.max() * 1e3),
                    "throughput_req_s": n_req / span, "throughput_pts_s": pts / span})
    sizes = np.array([r.n for r in reqs])
    return {"workload": "c5 serving: Poisson arrivals, N log-uniform [500,50000], d in "
                        "{64,384,768}, k in {15,30}, euclidean mixture / cosine unit rows; "
                        f"requests drawn from a pool of {pool} seeded datasets (N from "
                        f"{sizes.min()} to {sizes.max()}, mean {sizes.mean():.0f}); "
                        "max_jobs 4, max_batch 16",
            "pool": pool, "rates": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="bf16")
    ap.add_argument("--nsplit", type=int, default=0,
                    help="column splits of the sweep, fixed (no autotuning); 0: autotuned")
    ap.add_argument("--calib", default="ref_f32", choices=["ref_f32", "f64"],
                    help="calibration mode of the timed path")
    ap.add_argument("--margin", type=int, default=-1,
                    help="candidates beyond k per row (C = k + margin); -1: 2k, autotuned")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the N=50,000 (d=784) BF16 vs FP32 vs CPU check")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop figures")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-dataset batch figures")
    ap.add_argument("--serve-plan", default="5:50,10:100,20:200,50:1000,100:1000,200:10000",
                    help="serving figures: rate:requests pairs (Poisson arrivals, req/s)")
    ap.add_argument("--serve-pool", type=int, default=32,
                    help="seeded request datasets the serving figures draw from")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB (reads and writes that "
                         "many bytes once) to calibrate the HBM byte counters, then exit")
    args = ap.parse_args()
    FIXED["nsplit"] = args.nsplit
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
    # the boundary's host-buffer form (one hm_knn_graph call: X from host
    # memory in, kNN, sigma/rho, the COO and the row pointer back to host
    # memory, job set-up included); measured first, so that the timed job's
    # sweeps are the run's last ones
    log(f"{args.workload}: hm_knn_graph host-buffer calls")
    H.knn_graph(x, d, k, metric, precision=args.precision, ctx=ctx)  # warm the memory pool
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        H.knn_graph(x, d, k, metric, precision=args.precision, ctx=ctx)
    t_host = (time.perf_counter() - t0) / reps
    log(f"{args.workload}: timed path")
    dt, stage, cnt, res = run_path(ctx, x, d, k, metric, args.precision, args.steps, args.warmup,
                                   margin=args.margin, calib=args.calib)
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
                   "metric": metric, "precision": args.precision, "calib": args.calib,
                   "autotune": args.nsplit == 0,
                   "parallelism": "single GPU", **sweep_config(cnt, n, k)},
        "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
        "fallback_rows": cnt["fallback_rows"],
        "escalated_rows": cnt["escalated_rows"],
        "nnz": cnt["nnz"],
    }
    out["stage_bandwidth"] = stage_bandwidth(n, d, k, cnt, stage)
    out["roofline"] = sweep_roofline(n, d, args.precision, stage, args.workload)
    out["pcie_inclusive"] = {"ms_per_call": t_host * 1e3, "pts_per_s": n / t_host,
                             "entry": "hm_knn_graph (host buffers in and out, per-call job)"}
    if not args.no_fp32 and args.precision != "fp32":
        log(f"{args.workload}: fp32 path")
        # same margin and calibration mode as the line's own run, so the
        # graphs compare like for like
        dt32, st32, c32, res32 = run_path(ctx, x, d, k, metric, "fp32", max(3, args.steps // 2), 2,
                                          margin=args.margin, calib=args.calib)
        same = bool(np.array_equal(res32.knn.indices, res.knn.indices) and
                    np.array_equal(res32.graph.vals.view(np.uint32), res.graph.vals.view(np.uint32)))
        out["fp32_path"] = {"ms_per_step": dt32 * 1e3, "pts_per_s": n / dt32,
                            "sweep_ms": st32["sweep"], "fallback_rows": c32["fallback_rows"],
                            "sweep": sweep_config(c32, n, k),
                            "roofline": sweep_roofline(n, d, "fp32", st32, args.workload),
                            "identical_graph": same}
        out["speedup_vs_fp32"] = dt32 / dt
    if not args.no_cpu:
        log(f"{args.workload}: cpu baseline")
        cb = cpu_baseline(x, d, k, metric, res)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    if not args.no_target and args.workload != "c3":
        out["target_n50k"] = target_check(ctx, 10, with_cpu=not args.no_cpu)
    if not args.no_batch:
        log("c4: batch")
        out["batch_c4"] = batch_bench(ctx, 10)
    if not args.no_serve:
        plan = [(float(a), int(b)) for a, b in
                (v.split(":") for v in args.serve_plan.split(",") if v)]
        out["serving"] = serve_bench(ctx, plan, pool=args.serve_pool)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
