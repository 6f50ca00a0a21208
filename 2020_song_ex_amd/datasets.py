"""Seeded synthetic workloads and vector-file I/O.

Workloads (all seeded, numpy PCG64):

* ``gmm_int``  -- Gaussian mixture with centres U[0,255]^dim, isotropic sigma,
  clipped to [0,255] and rounded to integers (SIFT-like).  Squared L2
  distances of such vectors are exact in float32 while 255^2 * dim < 2^24
  (dim <= 258), so every summation order gives the same distances.
* ``beta_mix`` -- centres U[0,1]^dim, each point = clip(centre + 0.25 *
  (Beta(2,5) - 2/7), 0, 1) per coordinate (high-dimensional float data).
* ``vmf_mix``  -- unit-norm von Mises-Fisher mixture: uniform centres on the
  sphere, concentration kappa, sampled with Wood's (1994) rejection method.

File formats: ``.fvecs`` / ``.ivecs`` (int32 dim + row, parser.h:10-39,
run_compare.py:149-189) and ``.npy`` int32 ground truth (bench_gpu.cu:163-219).
"""
from __future__ import annotations

import struct

import numpy as np


def _centres_for(rng, n, n_centres):
    return rng.integers(0, n_centres, size=n)


def gmm_int(n, nq, dim=128, n_centres=1000, sigma=25.0, seed=0, n_extra=0):
    """(base [n, dim], queries [nq, dim]) float32 integer-valued in [0, 255].
    n_extra > 0 also returns that many further draws of the same mixture
    (drawn after the queries, so base and queries do not change)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 255.0, size=(n_centres, dim)).astype(np.float32)

    def draw(m):
        out = np.empty((m, dim), np.float32)
        step = 1 << 16
        for s in range(0, m, step):
            e = min(m, s + step)
            lab = _centres_for(rng, e - s, n_centres)
            x = centres[lab] + rng.normal(0.0, sigma, size=(e - s, dim)).astype(np.float32)
            np.clip(x, 0.0, 255.0, out=x)
            out[s:e] = np.rint(x)
        return out

    base = draw(n)
    queries = draw(nq)  # held-out draws from the same mixture
    if n_extra:
        return base, queries, draw(n_extra)
    return base, queries


def beta_mix(n, nq, dim=960, n_centres=500, scale=0.25, seed=0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(n_centres, dim)).astype(np.float32)
    mean = 2.0 / 7.0

    def draw(m):
        out = np.empty((m, dim), np.float32)
        step = 1 << 14
        for s in range(0, m, step):
            e = min(m, s + step)
            lab = _centres_for(rng, e - s, n_centres)
            x = centres[lab] + scale * (rng.beta(2.0, 5.0, size=(e - s, dim)).astype(np.float32) - mean)
            np.clip(x, 0.0, 1.0, out=x)
            out[s:e] = x
        return out

    return draw(n), draw(nq)


def _vmf_w(rng, kappa, dim, m):
    """Wood (1994): the component along the mean direction, m samples."""
    b = (-2.0 * kappa + np.sqrt(4.0 * kappa * kappa + (dim - 1.0) ** 2)) / (dim - 1.0)
    x0 = (1.0 - b) / (1.0 + b)
    c = kappa * x0 + (dim - 1.0) * np.log(1.0 - x0 * x0)
    out = np.empty(m, np.float64)
    todo = np.arange(m)
    while todo.size:
        z = rng.beta((dim - 1.0) / 2.0, (dim - 1.0) / 2.0, size=todo.size)
        w = (1.0 - (1.0 + b) * z) / (1.0 - (1.0 - b) * z)
        u = rng.uniform(size=todo.size)
        ok = kappa * w + (dim - 1.0) * np.log(1.0 - x0 * w) - c >= np.log(u)
        out[todo[ok]] = w[ok]
        todo = todo[~ok]
    return out


def vmf_mix(n, nq, dim=256, n_centres=2000, kappa=50.0, seed=0):
    """(base, queries) unit-norm float32 from a vMF mixture."""
    rng = np.random.default_rng(seed)
    mu = rng.normal(size=(n_centres, dim))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)

    def draw(m):
        out = np.empty((m, dim), np.float32)
        step = 1 << 15
        for s in range(0, m, step):
            e = min(m, s + step)
            lab = _centres_for(rng, e - s, n_centres)
            mus = mu[lab]
            w = _vmf_w(rng, kappa, dim, e - s)
            v = rng.normal(size=(e - s, dim))
            v -= np.sum(v * mus, axis=1, keepdims=True) * mus  # tangent direction
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            x = w[:, None] * mus + np.sqrt(np.maximum(0.0, 1.0 - w * w))[:, None] * v
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            out[s:e] = x.astype(np.float32)
        return out

    return draw(n), draw(nq)


def hyperplane_codes(vecs, n_bits=256, seed=0, planes=None):
    """sign(R x) > 0 packed into uint64 [n, n_bits/64]; R seeded Gaussian [n_bits, dim]."""
    vecs = np.asarray(vecs, dtype=np.float32)
    if planes is None:
        planes = np.random.default_rng(seed).normal(size=(n_bits, vecs.shape[1])).astype(np.float32)
    words = n_bits // 64
    out = np.zeros((vecs.shape[0], words), np.uint64)
    step = 1 << 15
    weights = (np.uint64(1) << np.arange(64, dtype=np.uint64))
    for s in range(0, vecs.shape[0], step):
        bits = (vecs[s:s + step] @ planes.T) > 0
        for w in range(words):
            out[s:s + step, w] = np.bitwise_or.reduce(
                np.where(bits[:, 64 * w:64 * (w + 1)], weights, np.uint64(0)), axis=1)
    return out, planes


# ------------------------------------------------------------------ configs

def config1(seed=0):
    """10,000 x 128 integer GMM (50 centres, sigma 20), 100 queries; out-degree 16,
    k=10, queue 64, Bloom 2^15."""
    base, q = gmm_int(10_000, 100, dim=128, n_centres=50, sigma=20.0, seed=seed)
    return {"base": base, "queries": q, "block": 16, "k": 10, "ef": 64, "bloom_bits": 1 << 15}


def config2(seed=0, n=1_000_000, nq=10_000, n_extra=0):
    """1,000,000 x 128 integer GMM (1,000 centres, sigma 25), 10,000 held-out queries;
    out-degree 32; k in {10, 100} x queue in {32, 64, 128, 256, 512}.
    n_extra: further draws (training / validation queries for the learned search)."""
    out = gmm_int(n, nq, dim=128, n_centres=1000, sigma=25.0, seed=seed, n_extra=n_extra)
    cfg = {"base": out[0], "queries": out[1], "block": 32, "ks": (10, 100), "efs": (32, 64, 128, 256, 512)}
    if n_extra:
        cfg["extra"] = out[2]
    return cfg


def config3(seed=0, n=1_000_000, nq=1_000):
    """1,000,000 x 960 values in [0, 1] around 500 centres perturbed by Beta(2, 5) noise, 1,000
    queries; out-degree 32, L2, k=10 (the high-dimensional plain and persistent kernels)."""
    base, q = beta_mix(n, nq, dim=960, n_centres=500, seed=seed)
    return {"base": base, "queries": q, "block": 32, "k": 10}


def config4(seed=0, n=1_000_000, nq=10_000, n_extra=0):
    """1,000,000 unit vectors (dim 256, 2,000-centre vMF mixture, kappa 50) hashed to 256-bit
    codes (uint64 [., 4]) by seeded Gaussian hyperplanes; 10,000 query codes; out-degree 32,
    Hamming, k=10 (the Hamming path with learned entry and termination).
    n_extra: further query codes (training / validation)."""
    vec, qv = vmf_mix(n, nq + n_extra, dim=256, n_centres=2000, kappa=50.0, seed=seed)
    codes, planes = hyperplane_codes(vec, seed=seed + 1)
    del vec
    qcodes, _ = hyperplane_codes(qv, planes=planes)
    cfg = {"base": codes, "queries": qcodes[:nq], "block": 32, "k": 10}
    if n_extra:
        cfg["extra"] = qcodes[nq:]
    return cfg


def config5(seed=0, n=1_200_000, nq=10_000, n_extra=0):
    """1,200,000 unit vectors (dim 200, 5,000-centre vMF mixture, kappa 30) searched by inner
    product; out-degree 32; requests with k in {1, 10, 100} and queue budgets in {64, 256}
    (the serving workload).  n_extra: further query draws."""
    base, q = vmf_mix(n, nq + n_extra, dim=200, n_centres=5000, kappa=30.0, seed=seed)
    cfg = {"base": base, "queries": q[:nq], "block": 32, "ks": (1, 10, 100), "efs": (64, 256)}
    if n_extra:
        cfg["extra"] = q[nq:]
    return cfg


# ------------------------------------------------------------------ file I/O

def read_fvecs(path, limit=None, dtype="<f4"):
    """Rows of an .fvecs/.ivecs file (int32 dim, then dim values), parser.h:10-39."""
    raw = np.fromfile(path, dtype=np.int32)
    if raw.size == 0:
        return np.empty((0, 0), np.float32 if dtype == "<f4" else np.int32)
    dim = int(raw[0])
    rows = raw.size // (dim + 1)
    if limit is not None:
        rows = min(rows, limit)
    mat = raw[: rows * (dim + 1)].reshape(rows, dim + 1)[:, 1:]
    return mat.view(np.float32).copy() if dtype == "<f4" else mat.copy()


def read_ivecs(path, limit=None):
    return read_fvecs(path, limit, dtype="<i4")


def write_fvecs(path, arr):
    """run_compare.py:188-196."""
    arr = np.asarray(arr, dtype=np.float32)
    n, dim = arr.shape
    with open(path, "wb") as f:
        for i in range(n):
            f.write(struct.pack("<i", dim))
            f.write(arr[i].tobytes())
