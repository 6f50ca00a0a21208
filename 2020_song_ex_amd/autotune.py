"""Autotuner for the search launch: picks the configuration with the highest
measured throughput whose recall@k on a validation set meets a target.

Knobs: kernel mode (plain / persistent), persistent grid (waves per compute
unit), queries per workgroup, Bloom filter size, queue size and -- for the
learned search -- the stop threshold.  Every candidate is timed on the device
(HIP events, inputs resident in HBM) and its recall measured against exact
ground truth; the pick is deterministic given those measurements (ties go to
the earlier candidate in the grid order).  Bloom size and threshold change
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
    """The grid's configurations in the order they are tried (product over the keys in order)."""
    names = list(grid)
    return [dict(zip(names, values)) for values in itertools.product(*[grid[n] for n in names])]


def pick(tried, target_recall):
    """The candidate with the highest throughput among those whose recall meets the target; on equal
    throughput the earlier one.  None if no candidate meets it."""
    best = None
    for c in tried:
        if c.recall >= target_recall and (best is None or c.qps > best.qps):
            best = c
    return best


def tune(index, queries, gt, k, grid: dict, target_recall, repeats=3, log=None) -> TuneResult:
    """Grid search over `grid` (name -> list of values, including "ef")."""
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
    return res
