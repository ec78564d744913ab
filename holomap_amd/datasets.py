"""Seeded synthetic workloads for the kNN-graph path (numpy, deterministic).

  c1  1,000 x 50     mixture of 10 isotropic Gaussians, centres ~ N(0, 4^2 I),
                     noise sigma = 1, equal clusters, seed 42, euclidean, k = 15
  c2  20,000 x 128   mixture of 50 anisotropic Gaussians, centres ~ N(0, 3^2 I),
                     per-cluster diagonal scales ~ U[0.5, 2], 2 % of the rows
                     overwritten by exact copies of other rows, seed 7, k = 15
  c3  50,000 x 784   28x28 image-like rows: a random mask blurred by a 3x3 box
                     filter and thresholded to ~20 % non-zero pixels, non-zeros
                     uniform integers 1..255 scaled by 1/255; k = 15
  c4  256 datasets, N log-uniform in [200, 8000], d in {32, 64, 128, 256},
      5-40 clusters (sigma 1, centres ~ N(0, 4^2 I)), k in {10, 15, 30},
      euclidean / cosine 50/50
  c5  serving requests: N log-uniform in [500, 50000], d in {64, 384, 768},
      Gaussian mixture (euclidean) or unit-normalised N(0, I) rows (cosine),
      k in {15, 30}, per-request seed
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np


def gaussian_mixture(n, d, n_clusters, center_std, rng, scales=None, noise=1.0):
    centers = rng.normal(0.0, center_std, size=(n_clusters, d))
    labels = np.arange(n) % n_clusters
    rng.shuffle(labels)
    sc = noise if scales is None else scales[labels]
    x = centers[labels] + rng.normal(0.0, 1.0, size=(n, d)) * sc
    return x.astype(np.float32)


def c1(seed: int = 42) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return gaussian_mixture(1000, 50, 10, 4.0, rng)


def c2(seed: int = 7, n: int = 20000, d: int = 128) -> np.ndarray:
    rng = np.random.default_rng(seed)
    scales = rng.uniform(0.5, 2.0, size=(50, d))
    x = gaussian_mixture(n, d, 50, 3.0, rng, scales=scales)
    m = n // 50  # 2 % duplicated rows
    perm = rng.permutation(n)
    tgt, src = perm[:m], perm[m:2 * m]
    x[tgt] = x[src]
    return x


def c3(seed: int = 3, n: int = 50000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 784), np.float32)
    step = 5000
    for s in range(0, n, step):
        b = min(step, n - s)
        mask = (rng.random((b, 30, 30)) < 0.10).astype(np.float32)
        blur = np.zeros((b, 28, 28), np.float32)
        for dy in range(3):
            for dx in range(3):
                blur += mask[:, dy:dy + 28, dx:dx + 28]
        blur = blur.reshape(b, 784) + rng.random((b, 784), dtype=np.float32) * 0.01
        thr = np.quantile(blur, 0.8, axis=1, keepdims=True)
        nz = blur > thr
        vals = rng.integers(1, 256, size=(b, 784)).astype(np.float32) / np.float32(255.0)
        out[s:s + b] = np.where(nz, vals, np.float32(0.0))
    return out


@dataclass
class Dataset:
    data: np.ndarray
    n: int
    d: int
    k: int
    metric: str


def c4(seed: int = 11, count: int = 256):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(8000)))))
        d = int(rng.choice([32, 64, 128, 256]))
        k = int(rng.choice([10, 15, 30]))
        metric = "euclidean" if rng.random() < 0.5 else "cosine"
        nc = int(rng.integers(5, 41))
        x = gaussian_mixture(n, d, nc, 4.0, rng)
        out.append(Dataset(x, n, d, k, metric))
    return out


def c5_request(seed: int, n_min: int = 500, n_max: int = 50000) -> Dataset:
    rng = np.random.default_rng(seed)
    n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
    d = int(rng.choice([64, 384, 768]))
    k = int(rng.choice([15, 30]))
    if rng.random() < 0.5:
        nc = int(rng.integers(5, 41))
        x = gaussian_mixture(n, d, nc, 4.0, rng)
        metric = "euclidean"
    else:
        x = rng.normal(0.0, 1.0, size=(n, d))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        x = x.astype(np.float32)
        metric = "cosine"
    return Dataset(x, n, d, k, metric)


def sha256(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x, np.float32).tobytes()).hexdigest()
