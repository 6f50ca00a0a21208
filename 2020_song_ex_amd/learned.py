"""Learned entry, scheduling and early termination for the graph search.

Three 2 x 64 ReLU MLPs (FP32), trained on the device by libsong's HIP trainer
(song_mlp_train).  The query's 16 leading principal components (of a base
sample, standardised) feed the first two, which run in a planning pass before
the search (song_plan.hip); the third runs fused inside the search kernel:

* entry network: projection -> scores over 64 pivot nodes (nodes nearest to
  k-means centroids of the base); the best n_seeds pivots replace the
  reference's evenly spaced seeds (kernalGraph.h:215-216).  Label: the pivot
  nearest to the query.
* cost network: projection -> the query's standardised log2 hop count (squared
  error).  The persistent queue runs the queries most expensive first, so the
  long searches overlap the short ones instead of trailing after them.
* stop network: 8 features of where the search stands -- hop count, popped
  distance minus the k-th best, spread of the top k, hops since the top k last
  changed, frontier entries still closer than the k-th best, recent top-k
  insertions, frontier fill, recent improvement of the k-th best (distances
  divided by a per-index scale; song_search_impl.h) -> probability that the
  k-th best distance is already final.  Labels come from traces of the search
  itself (song_learn_trace): at a consulted hop t the label is 1 when the k-th
  best distance at t equals the one at the end of the search.

The plain kernel stays the fallback: a NaN entry score keeps the reference
seeds, a NaN cost sorts as zero and a NaN stop logit never stops, so a broken
model returns the plain search's result.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .index import Index

P_DIM, N_PIVOTS, N_FEATURES, HIDDEN = 16, 64, 8, 64
STOP_FLOATS = HIDDEN * N_FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN  # + b3
LOSS_DEFAULT, LOSS_MSE = 0, 1


class SongLearnedModel(ctypes.Structure):
    _fields_ = [("proj_dim", ctypes.c_int32), ("mu", ctypes.c_void_p), ("proj", ctypes.c_void_p),
                ("proj_scale", ctypes.c_void_p), ("entry_w1", ctypes.c_void_p), ("entry_b1", ctypes.c_void_p),
                ("entry_w2", ctypes.c_void_p), ("entry_b2", ctypes.c_void_p), ("entry_w3", ctypes.c_void_p),
                ("entry_b3", ctypes.c_void_p), ("pivots", ctypes.c_void_p), ("stop_weights", ctypes.c_void_p),
                ("stop_b3", ctypes.c_float), ("dist_scale", ctypes.c_float), ("cost_weights", ctypes.c_void_p)]


def _bind():
    L = N.lib()
    if not hasattr(L, "_learned_bound"):
        L.song_index_set_model.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.song_learn_trace.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        L.song_mlp_train.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_float, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.song_mlp_forward.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L._learned_bound = True
    return L


# ------------------------------------------------------------------ MLP helpers

def mlp_init(in_dim, out_dim, seed=0):
    """He-initialised flat weights: w1 [64 x in] b1 w2 [64 x 64] b2 w3 [out x 64] b3."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2.0 / in_dim), size=(HIDDEN, in_dim))
    w2 = rng.normal(0, np.sqrt(2.0 / HIDDEN), size=(HIDDEN, HIDDEN))
    w3 = rng.normal(0, np.sqrt(1.0 / HIDDEN), size=(out_dim, HIDDEN))
    parts = [w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3.ravel(), np.zeros(out_dim)]
    return np.concatenate(parts).astype(np.float32)


def mlp_split(w, in_dim, out_dim):
    o = 0
    out = []
    for shape in [(HIDDEN, in_dim), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (out_dim, HIDDEN), (out_dim,)]:
        n = int(np.prod(shape))
        out.append(w[o:o + n].reshape(shape))
        o += n
    return out


def mlp_forward_numpy(w, X, in_dim, out_dim):
    """Plain FP32 reference forward pass (float64 accumulation)."""
    w1, b1, w2, b2, w3, b3 = [a.astype(np.float64) for a in mlp_split(w, in_dim, out_dim)]
    h = np.maximum(X.astype(np.float64) @ w1.T + b1, 0)
    h = np.maximum(h @ w2.T + b2, 0)
    return h @ w3.T + b3


def mlp_train(X, y, in_dim, out_dim, epochs=10, batch=1024, lr=2e-3, seed=0, weights=None, device=0,
              loss=LOSS_DEFAULT):
    """song_mlp_train: Adam on the device; returns (weights, per-epoch loss)."""
    L = _bind()
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    w = mlp_init(in_dim, out_dim, seed) if weights is None else np.array(weights, dtype=np.float32)
    curve = np.zeros(epochs, np.float32)
    N.check(L.song_mlp_train(device, N.ptr(X), N.ptr(y), X.shape[0], in_dim, out_dim, loss, epochs, batch, lr,
                             seed, N.ptr(w), N.ptr(curve)), "song_mlp_train")
    return w, curve


def mlp_forward(w, X, in_dim, out_dim, device=0):
    """The search kernel's own inference code on given inputs (song_mlp_forward)."""
    L = _bind()
    X = np.ascontiguousarray(X, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    out = np.empty((X.shape[0], out_dim), np.float32)
    N.check(L.song_mlp_forward(device, N.ptr(X), X.shape[0], in_dim, out_dim, N.ptr(w), N.ptr(out)),
            "song_mlp_forward")
    return out


# ------------------------------------------------------------------ the model

class LearnedModel:
    """Arrays of a learned model; set_on() uploads them to an index."""

    def __init__(self, proj_dim, mu, proj, proj_scale, entry_w, pivots, stop_w, dist_scale, cost_w=None):
        self.proj_dim = int(proj_dim)
        self.mu = np.ascontiguousarray(mu, np.float32)
        self.proj = np.ascontiguousarray(proj, np.float32)
        self.proj_scale = np.ascontiguousarray(proj_scale, np.float32)
        self.entry_w = np.ascontiguousarray(entry_w, np.float32)   # flat 16 -> 64 layout
        self.pivots = np.ascontiguousarray(pivots, np.int32)
        self.stop_w = np.ascontiguousarray(stop_w, np.float32)     # flat 8 -> 1 layout (b3 last)
        self.dist_scale = float(dist_scale)
        self.cost_w = None if cost_w is None else np.ascontiguousarray(cost_w, np.float32)  # flat 16 -> 1

    def set_on(self, index: Index):
        L = _bind()
        ew = mlp_split(self.entry_w, P_DIM, N_PIVOTS)
        ew = [np.ascontiguousarray(a, np.float32) for a in ew]
        self._keep = ew
        m = SongLearnedModel()
        m.proj_dim = self.proj_dim
        m.mu = self.mu.ctypes.data
        m.proj = self.proj.ctypes.data
        m.proj_scale = self.proj_scale.ctypes.data
        m.entry_w1, m.entry_b1, m.entry_w2, m.entry_b2, m.entry_w3, m.entry_b3 = [a.ctypes.data for a in ew]
        m.pivots = self.pivots.ctypes.data
        m.stop_weights = self.stop_w.ctypes.data
        m.stop_b3 = float(self.stop_w[STOP_FLOATS])
        m.dist_scale = self.dist_scale
        m.cost_weights = self.cost_w.ctypes.data if self.cost_w is not None else None
        N.check(L.song_index_set_model(index.handle, ctypes.byref(m)), "song_index_set_model")

    def save(self, path):
        extra = {} if self.cost_w is None else {"cost_w": self.cost_w}
        np.savez(path, proj_dim=self.proj_dim, mu=self.mu, proj=self.proj, proj_scale=self.proj_scale,
                 entry_w=self.entry_w, pivots=self.pivots, stop_w=self.stop_w, dist_scale=self.dist_scale, **extra)

    @classmethod
    def load(cls, path):
        z = np.load(path)
        return cls(int(z["proj_dim"]), z["mu"], z["proj"], z["proj_scale"], z["entry_w"], z["pivots"],
                   z["stop_w"], float(z["dist_scale"]), z["cost_w"] if "cost_w" in z.files else None)

    def project(self, queries, metric):
        """The entry network's input, as the kernel computes it (float64 here)."""
        if metric == N.METRIC_HAMMING:
            bits = np.unpackbits(np.ascontiguousarray(queries, np.uint64).view(np.uint8), axis=1,
                                 bitorder="little").astype(np.float64) * 2 - 1
            return (bits @ self.proj.T.astype(np.float64)) * self.proj_scale
        q = np.zeros((queries.shape[0], self.proj_dim), np.float64)
        q[:, :queries.shape[1]] = queries
        return ((q - self.mu) @ self.proj.T.astype(np.float64)) * self.proj_scale


def _as_rows(vecs, metric, proj_dim):
    if metric == N.METRIC_HAMMING:
        return np.unpackbits(np.ascontiguousarray(vecs, np.uint64).view(np.uint8), axis=1,
                             bitorder="little").astype(np.float32) * 2 - 1
    out = np.zeros((vecs.shape[0], proj_dim), np.float32)
    out[:, :vecs.shape[1]] = vecs
    return out


def kmeans_pivots(index: Index, base, n_pivots=N_PIVOTS, sample=50_000, iters=12, seed=0, pool=1.0):
    """Nodes nearest to k-means centroids of a base sample (distinct ids), chosen among the first
    pool * n nodes.  The NSW build inserts nodes in id order and early nodes keep the long-range
    edges of the sparse graph they joined, so a small pool gives seeds from which the greedy
    search moves fast (node 0, the first reference seed, is the extreme case)."""
    rng = np.random.default_rng(seed)
    n = base.shape[0]
    idx = rng.choice(n, size=min(sample, n), replace=False)
    X = _as_rows(base[idx], index.metric, index.row_stride if index.metric != N.METRIC_HAMMING else 256)
    C = X[rng.choice(X.shape[0], size=n_pivots, replace=X.shape[0] < n_pivots)].copy()
    for _ in range(iters):
        d = (X * X).sum(1)[:, None] + (C * C).sum(1)[None, :] - 2 * X @ C.T
        lab = d.argmin(1)
        for c in range(n_pivots):
            m = lab == c
            if m.any():
                C[c] = X[m].mean(0)
    if index.metric == N.METRIC_HAMMING:
        bits = (C > 0).astype(np.uint8)
        cq = np.packbits(bits, axis=1, bitorder="little").view(np.uint64)
    else:
        cq = C[:, :base.shape[1]]
    n_pool = max(n_pivots * 8, min(n, int(round(pool * n))))
    if n_pool >= n:
        ids, _, _ = index.bruteforce(cq, 8)
    else:
        sub = Index(base[:n_pool], np.full((n_pool, 2), -1, np.int32), metric=index.metric, device=index.device)
        ids, _, _ = sub.bruteforce(cq, 8)
        sub.close()
    chosen = []
    for row in ids:
        for v in row:
            if v >= 0 and v not in chosen:
                chosen.append(int(v))
                break
    while len(chosen) < n_pivots:
        v = int(rng.integers(0, n))
        if v not in chosen:
            chosen.append(v)
    return np.array(chosen, np.int32)


def fit(index: Index, base, train_queries, *, k=10, ef=256, n_seeds=4, bloom_bits=0, stop_min_hops=8,
        stop_interval=2, max_evals=128, entry_epochs=20, stop_epochs=4, cost_epochs=20, cost_ef=None,
        cost_threshold=2.5, pivot_pool=1.0, seed=0, log=None, trace_entry=True, trace_ef=None,
        cost_search=None):
    """Train the three networks for this index; returns the LearnedModel (already set on the index).
    The cost network learns the hop counts of the learned search (entry + stop at cost_threshold) at
    queue size cost_ef (default ef).  The stop network learns from traces of the search at queue
    size trace_ef (default ef), from the learned entry points (trace_entry) or the reference seeds:
    the features depend on both, so they are best taken as the stop will be used.  cost_search:
    the search options (Index.search keywords) of the run whose hop counts the cost network
    learns, when the queue is to be ordered for another configuration than the default one."""
    say = log or (lambda *_: None)
    metric = index.metric
    proj_dim = 256 if metric == N.METRIC_HAMMING else index.row_stride
    rng = np.random.default_rng(seed)
    train_queries = np.ascontiguousarray(train_queries)
    nq = train_queries.shape[0]

    pivots = kmeans_pivots(index, base, seed=seed, pool=pivot_pool)
    say(f"pivots: {len(set(pivots.tolist()))} distinct nodes")
    if metric == N.METRIC_HAMMING:
        mu = np.zeros(proj_dim, np.float32)
    else:
        mu = np.zeros(proj_dim, np.float32)
        sample = base[rng.choice(base.shape[0], size=min(100_000, base.shape[0]), replace=False)]
        mu[:base.shape[1]] = sample.mean(0)
    # projection: the 16 leading principal directions of a base sample
    sample = _as_rows(base[rng.choice(base.shape[0], size=min(50_000, base.shape[0]), replace=False)],
                      metric, proj_dim).astype(np.float64)
    sample -= mu
    cov = sample.T @ sample / sample.shape[0]
    evals, evecs = np.linalg.eigh(cov)
    proj = evecs[:, ::-1][:, :P_DIM].T.astype(np.float32)
    if metric != N.METRIC_HAMMING:
        proj[:, base.shape[1]:] = 0.0
    model = LearnedModel(proj_dim, mu, proj, np.ones(P_DIM, np.float32), mlp_init(P_DIM, N_PIVOTS, seed),
                         pivots, mlp_init(N_FEATURES, 1, seed + 1), 1.0)
    z = model.project(train_queries, metric)
    model.proj_scale = (1.0 / np.maximum(z.std(0), 1e-12)).astype(np.float32)
    z = model.project(train_queries, metric).astype(np.float32)

    # entry labels: the pivot nearest to each query (GPU brute force over the pivots)
    pv = base[pivots]
    pindex = Index(pv, np.full((N_PIVOTS, 2), -1, np.int32), metric=metric, device=index.device)
    lab, _, _ = pindex.bruteforce(train_queries, 1)
    pindex.close()
    ew, eloss = mlp_train(z, lab[:, 0].astype(np.float32), P_DIM, N_PIVOTS, epochs=entry_epochs,
                          lr=2e-3, seed=seed)
    model.entry_w = ew
    acc = (mlp_forward_numpy(ew, z, P_DIM, N_PIVOTS).argmax(1) == lab[:, 0]).mean()
    say(f"entry network: loss {eloss[0]:.3f} -> {eloss[-1]:.3f}, nearest-pivot accuracy {acc:.3f}")

    # distance scale: median k-th best distance of the plain search
    probe = index.search(train_queries[: min(2000, nq)], k, ef, bloom_bits=bloom_bits, n_seeds=n_seeds)
    kth = probe.dists[:, k - 1]
    scale = float(np.median(np.abs(kth[np.isfinite(kth)]))) or 1.0
    model.dist_scale = scale
    model.set_on(index)

    # stop traces (from the learned entry points or the reference seeds)
    dq = index.queries_to_device(train_queries)
    d_tr = N.DeviceBuffer(nq * max_evals * (N_FEATURES + 1) * 4)
    d_cnt = N.DeviceBuffer(nq * 4)
    d_fin = N.DeviceBuffer(nq * 4)
    o = N.make_opts(bloom_bits, 0, n_seeds, 0, 0, 0, 0)
    o.learned = 1 if trace_entry else 0
    o.stop_min_hops = stop_min_hops
    o.stop_interval = stop_interval
    L = _bind()
    N.check(L.song_learn_trace(index.handle, dq.ptr, nq, k, trace_ef or ef, ctypes.byref(o), max_evals, d_tr.ptr,
                               d_cnt.ptr, d_fin.ptr), "song_learn_trace")
    tr = d_tr.download(np.empty((nq, max_evals, N_FEATURES + 1), np.float32))
    cnt = d_cnt.download(np.empty(nq, np.int32))
    fin = d_fin.download(np.empty(nq, np.float32))
    rows = np.arange(max_evals)[None, :] < cnt[:, None]
    feats = tr[rows][:, :N_FEATURES]
    dk = tr[rows][:, N_FEATURES]
    final = np.repeat(fin, cnt)
    y = (dk == final).astype(np.float32)
    say(f"stop traces: {feats.shape[0]} hops from {nq} queries, {y.mean():.3f} already final")
    sw, sloss = mlp_train(feats, y, N_FEATURES, 1, epochs=stop_epochs, lr=2e-3, seed=seed + 1)
    model.stop_w = sw
    say(f"stop network: loss {sloss[0]:.3f} -> {sloss[-1]:.3f}")
    model.set_on(index)

    # cost labels: standardised log2 hop counts of the learned search itself
    copts = dict(learned=3, stop_threshold=cost_threshold, stop_min_hops=stop_min_hops, stop_interval=stop_interval)
    if cost_search:
        copts = {kk: v for kk, v in cost_search.items() if kk not in ("ef", "mode", "waves_per_block")}
        copts["learned"] = copts.get("learned", 2) & 3  # the labelling run itself keeps input order
    sopts = {"bloom_bits": bloom_bits, "n_seeds": n_seeds, **copts}  # cost_search's own values win
    r = index.search(train_queries, k, cost_ef or ef, **sopts)
    hops = np.log2(np.maximum(r.trace[:, 2], 1).astype(np.float64))
    mean, std = hops.mean(), max(hops.std(), 1e-6)
    target = ((hops - mean) / std).astype(np.float32)
    cw, closs = mlp_train(z, target, P_DIM, 1, epochs=cost_epochs, lr=2e-3, seed=seed + 2, loss=LOSS_MSE)
    pred = mlp_forward_numpy(cw, z, P_DIM, 1)[:, 0]
    corr = float(np.corrcoef(pred, target)[0, 1]) if target.std() > 0 else 0.0
    say(f"cost network: loss {closs[0]:.3f} -> {closs[-1]:.3f}, correlation with log2 hops {corr:.3f}")
    model.cost_w = cw
    model.set_on(index)
    return model
