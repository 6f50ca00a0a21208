"""Autotuner for the search launch: picks the configuration with the highest
measured throughput whose recall@k on a validation set meets a target.

Knobs: kernel mode (plain / persistent), persistent grid (waves per compute
unit), queries per workgroup, Bloom filter size, queue size and -- for the
learned search -- the stop threshold.  Every candidate is timed on the device
(HIP events, inputs resident in HBM, best of a few launches) and its recall
measured against exact ground truth.  Candidates a few percent apart swap
places from one pass to the next, so the leading ones that meet the recall
target are timed again, interleaved over several rounds, and the pick is the
one with the lowest median launch time.  The pick is deterministic given those
measurements (ties go to the earlier candidate in the grid order).  Bloom size and threshold change
results, so each candidate's recall is its own; the plain-kernel parity tests
cover every Bloom size.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


def recall_at(ids, gt, k):
    hits = 0
    for a, g in zip(ids[:, :k], gt[:, :k]):
        hits += len(set(a.tolist()) & set(g.tolist()))
    return hits / (ids.shape[0] * k)


@dataclass
class Candidate:
    params: dict
    recall: float = 0.0
    kernel_ms: float = 0.0
    qps: float = 0.0
    expanded_mean: float = 0.0


@dataclass
class TuneResult:
    best: Candidate | None
    tried: list = field(default_factory=list)

    def table(self):
        return [{**c.params, "recall": round(c.recall, 4), "kernel_ms": round(c.kernel_ms, 4),
                 "qps": round(c.qps, 1), "expanded_mean": round(c.expanded_mean, 2)} for c in self.tried]


def measure(index, dq, nq, k, params, gt, d_ids, d_tr, repeats=3):
    """Run one configuration: (recall, best kernel ms of `repeats`, mean expansions)."""
    p = dict(params)
    ef = p.pop("ef")
    index.search_device(dq, nq, k, ef, d_ids, None, d_tr, **p)
    ids = d_ids.download(np.empty((nq, k), np.int32))
    tr = d_tr.download(np.empty((nq, 12), np.int32))
    best = float("inf")
    for _ in range(repeats):
        tm = index.search_device(dq, nq, k, ef, d_ids, None, None, **p)
        best = min(best, tm["kernel_ms"])
    return recall_at(ids, gt, k), best, float(tr[:, 2].mean())


def candidates(grid: dict):
    """The grid's configurations in the order they are tried (product over the keys in order).  A
    value that is a dict sets several knobs together (its items replace the key)."""
    names = list(grid)
    out = []
    for values in itertools.product(*[grid[n] for n in names]):
        c = {}
        for n, v in zip(names, values):
            if isinstance(v, dict):
                c.update(v)
            else:
                c[n] = v
        out.append(c)
    return out


def pick(tried, target_recall):
    """The candidate with the highest throughput among those whose recall meets the target; on equal
    throughput the earlier one.  None if no candidate meets it."""
    best = None
    for c in tried:
        if c.recall >= target_recall and (best is None or c.qps > best.qps):
            best = c
    return best


def finalists(tried, target_recall, n):
    """The `n` fastest candidates whose recall meets the target, fastest first (on equal
    throughput the earlier one)."""
    ok = [(i, c) for i, c in enumerate(tried) if c.recall >= target_recall]
    ok.sort(key=lambda ic: (-ic[1].qps, ic[0]))
    return [c for _, c in ok[:n]]


def pick_final(finals, times):
    """Among `finals` (in their order), the one whose launch times `times[i]` have the lowest
    median; on equal medians the earlier one."""
    best, best_ms = None, float("inf")
    for c, t in zip(finals, times):
        ms = float(np.median(t))
        if ms < best_ms:
            best, best_ms = c, ms
    return best


def tune(index, queries, gt, k, grid: dict, target_recall, repeats=3, n_final=4, final_rounds=7,
         log=None) -> TuneResult:
    """Grid search over `grid` (name -> list of values, including "ef"), then `final_rounds`
    interleaved launches of the `n_final` leading candidates (0: the grid pass decides)."""
    say = log or (lambda *_: None)
    nq = queries.shape[0]
    dq = index.queries_to_device(queries)
    d_ids = N.DeviceBuffer(nq * k * 4)
    d_tr = N.DeviceBuffer(nq * 48)
    res = TuneResult(None)
    for params in candidates(grid):
        rec, ms, exp = measure(index, dq, nq, k, params, gt, d_ids, d_tr, repeats)
        c = Candidate(params, rec, ms, nq / (ms / 1e3), exp)
        res.tried.append(c)
        say(f"tune {params} recall {rec:.4f} kernel {ms:.3f} ms qps {c.qps:.0f}")
    res.best = pick(res.tried, target_recall)
    finals = finalists(res.tried, target_recall, n_final)
    if len(finals) > 1 and final_rounds > 0:
        times = [[] for _ in finals]
        for _ in range(final_rounds):
            for c, t in zip(finals, times):
                p = dict(c.params)
                ef = p.pop("ef")
                t.append(index.search_device(dq, nq, k, ef, d_ids, None, None, **p)["kernel_ms"])
        res.best = pick_final(finals, times)
        for c, t in zip(finals, times):
            say(f"final {c.params} median {float(np.median(t)):.3f} ms")
    return res
