"""The oracle: a CPU restatement of the reference's SONG search, for checking.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The product library never loads it.

Two native libraries sit behind it:

* ``liboracle.so``  -- ``song_oracle.c``, the C restatement of
  ``kernalGraph.h:184-299`` / ``hashKernelGraph.h:238-352`` (search),
  ``gpuHeap.h`` (sets), ``bloomFilter.h`` (visited set), ``graph.h:33-115``
  (CPU search) and brute-force ground truth.
* ``_ref/libsongref.so`` -- the reference's own ``graph.h`` / ``bitHash.h`` /
  ``distance.h`` compiled from its sources by ``oracle/Makefile`` (built in the
  development container where the reference checkout exists; the built file
  travels with the tree).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
METRIC = {"l2": 0, "ip": 1, "hamming": 2}
TRACE_FIELDS = (
    "seed_count", "visited_nodes", "expanded_nodes", "neighbor_slots_scanned",
    "candidate_nodes", "distance_evals", "frontier_pushes", "result_updates",
    "peak_frontier", "peak_results", "result_count", "early_stopped",
)

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_lib = None
_ref = None


def build() -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_bloom_pos.argtypes = [ctypes.c_int, _I64, _I64]
        L.or_bloom_pos.restype = ctypes.c_int
        L.or_bloom_positions.argtypes = [_P, ctypes.c_int, _I64, _P]
        L.or_bloom_run.argtypes = [_I64, _P, ctypes.c_int, _P, ctypes.c_int, _P, _P]
        L.or_set_run.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_int, _P, _P, _P, _P, _P, _P]
        L.or_set_run.restype = ctypes.c_int
        L.or_l2_sequential.argtypes = [_P, _P, ctypes.c_int]
        L.or_l2_sequential.restype = ctypes.c_float
        L.or_row_dist_f32.argtypes = [ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int]
        L.or_row_dist_f32.restype = ctypes.c_float
        L.or_hamming256.argtypes = [_P, _P]
        L.or_hamming256.restype = ctypes.c_int
        L.or_seed_id.argtypes = [ctypes.c_int, _I64, ctypes.c_int]
        L.or_seed_id.restype = _I64
        L.or_search.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                ctypes.c_int, _P, _P, _P]
        L.or_search.restype = ctypes.c_int
        L.or_search_capped.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.or_search_capped.restype = ctypes.c_int
        L.or_search_learned.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64, ctypes.c_int,
                                        ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _P, _P, _P]
        L.or_search_learned.restype = ctypes.c_int
        L.or_stop_logits.argtypes = [_P, ctypes.c_float, _P, ctypes.c_int, _P]
        L.or_learned_trace.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64, ctypes.c_int,
                                       ctypes.c_int, _P, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.or_learned_trace.restype = ctypes.c_int
        L.or_cpu_search.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
        L.or_cpu_search.restype = ctypes.c_int
        L.or_bruteforce.argtypes = [_P, _I64, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _P, _P]
        L.or_bruteforce.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libsongref.so"))


def ref():
    """The reference's own CPU code (graph.h, bitHash.h), compiled in place."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libsongref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libsongref.so is not built (needs the reference sources)")
        R = ctypes.CDLL(path)
        for f in ("ref_block_size", "ref_fixed_degree", "ref_search_degree",
                  "ref_ef_construction", "ref_n_multiprobe"):
            getattr(R, f).restype = ctypes.c_int
        R.ref_l2.argtypes = [_P, _P, ctypes.c_int]
        R.ref_l2.restype = ctypes.c_float
        R.ref_hamming.argtypes = [_P, _P]
        R.ref_hamming.restype = ctypes.c_int
        R.ref_build_graph.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P]
        R.ref_build_graph.restype = ctypes.c_int
        R.ref_graph_search.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        R.ref_graph_search.restype = ctypes.c_double
        R.ref_bithash_project.argtypes = [ctypes.c_int, ctypes.c_uint, _P, ctypes.c_int, _P]
        R.ref_bithash_matrix.argtypes = [ctypes.c_int, ctypes.c_uint, _P]
        _ref = R
    return _ref


# --------------------------------------------------------------------------- oracle calls

def search(data, adj, queries, k, ef, *, block, metric="l2", bloom_bits=1 << 14, n_seeds=4,
           group=8, ref_order=False, exact_visited=False, unbounded=False, entries=None, max_hops=0,
           result_cap=0):
    """Re-execute the reference GPU search (kernalGraph.h:184-299) on the CPU.

    ``data``: float32 [n, dim] (l2/ip) or uint64 [n, 4] (hamming).
    Returns (ids [nq, k] best-first, dists [nq, k], trace [nq, 12]).
    """
    m = METRIC[metric]
    if m == 2:
        data = np.ascontiguousarray(data, dtype=np.uint64)
        queries = np.ascontiguousarray(queries, dtype=np.uint64)
        dim = 256
    else:
        data = np.ascontiguousarray(data, dtype=np.float32)
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        dim = data.shape[1]
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    n, nq = data.shape[0], queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    d = np.empty((nq, k), np.float32)
    tr = np.empty((nq, 12), np.int32)
    ne = 0
    if entries is not None:
        entries = np.ascontiguousarray(entries, dtype=np.int32)
        ne = entries.shape[1]
    rc = lib().or_search_capped(_ptr(data), _ptr(adj), n, dim, block, m, _ptr(queries), nq, k, ef,
                                bloom_bits, n_seeds, group, int(ref_order), int(exact_visited),
                                int(unbounded), _ptr(entries), ne, int(max_hops), int(result_cap), _ptr(ids),
                                _ptr(d), _ptr(tr))
    if rc != 0:
        raise ValueError("or_search rejected its arguments")
    return ids, d, tr


def search_learned(data, adj, queries, k, ef, *, block, stop_w, dist_scale, threshold, min_hops, interval,
                   confirm=1, max_hops=0, result_cap=0, late_hops=0, late_interval=0, metric="l2",
                   bloom_bits=1 << 14, n_seeds=4, group=8, entries=None, entry_late_hops=0):
    """The search with the learned stop network (the MI355X kernel's learned bit 1) re-executed on the
    CPU: stop_w is the flat network (w1 [64 x 8], b1, w2 [64 x 64], b2, w3, then b3).  entries
    [nq, n] replace the reference seeds, or with entry_late_hops > 0 join the search at that hop
    (which starts at the n reference seeds).
    Returns (ids [nq, k] best-first, dists [nq, k], trace [nq, 12])."""
    m = METRIC[metric]
    if m == 2:
        data = np.ascontiguousarray(data, dtype=np.uint64)
        queries = np.ascontiguousarray(queries, dtype=np.uint64)
        dim = 256
    else:
        data = np.ascontiguousarray(data, dtype=np.float32)
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        dim = data.shape[1]
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    w = np.ascontiguousarray(stop_w, dtype=np.float32)
    n, nq = data.shape[0], queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    d = np.empty((nq, k), np.float32)
    tr = np.empty((nq, 12), np.int32)
    ne = 0
    if entries is not None:
        entries = np.ascontiguousarray(entries, dtype=np.int32)
        ne = entries.shape[1]
    rc = lib().or_search_learned(_ptr(data), _ptr(adj), n, dim, block, m, _ptr(queries), nq, k, ef, bloom_bits,
                                 n_seeds, group, _ptr(entries), ne, int(max_hops), int(result_cap), _ptr(w),
                                 float(w[4800]), float(dist_scale), float(threshold), int(min_hops), int(interval),
                                 int(confirm), int(late_hops), int(late_interval), int(entry_late_hops), _ptr(ids),
                                 _ptr(d), _ptr(tr))
    if rc != 0:
        raise ValueError("or_search_learned rejected its arguments")
    return ids, d, tr


def learned_trace(data, adj, queries, k, ef, *, block, dist_scale, min_hops, interval, max_evals, late_hops=0,
                  late_interval=0, metric="l2", bloom_bits=1 << 14, n_seeds=4, group=8):
    """The learned search's trace mode re-executed on the CPU: per consulted hop the 8 features
    and the k-th best distance.  Returns (rows [nq, max_evals, 9], counts [nq], final_dk [nq])."""
    m = METRIC[metric]
    dt = np.uint64 if m == 2 else np.float32
    data = np.ascontiguousarray(data, dtype=dt)
    queries = np.ascontiguousarray(queries, dtype=dt)
    dim = 256 if m == 2 else data.shape[1]
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    rows = np.zeros((nq, max_evals, 9), np.float32)
    cnt = np.empty(nq, np.int32)
    fin = np.empty(nq, np.float32)
    rc = lib().or_learned_trace(_ptr(data), _ptr(adj), data.shape[0], dim, block, m, _ptr(queries), nq, k, ef,
                                bloom_bits, n_seeds, group, None, 0, float(dist_scale), int(min_hops), int(interval),
                                int(late_hops), int(late_interval), int(max_evals), _ptr(rows), _ptr(cnt), _ptr(fin))
    if rc != 0:
        raise ValueError("or_learned_trace rejected its arguments")
    return rows, cnt, fin


def stop_logits(stop_w, features):
    """The stop network's logits for feature rows [n, 8], in the kernel's operation order."""
    w = np.ascontiguousarray(stop_w, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float32)
    out = np.empty(f.shape[0], np.float32)
    lib().or_stop_logits(_ptr(w), float(w[4800]), _ptr(f), f.shape[0], _ptr(out))
    return out


def cpu_search(data, adj, queries, k, ef, *, block, n_seeds=4):
    """Graph::search (graph.h:33-115) restated; ids worst-first like the reference."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    tr = np.empty((nq, 12), np.int32)
    lib().or_cpu_search(_ptr(data), _ptr(adj), data.shape[0], data.shape[1], block,
                        _ptr(queries), nq, k, ef, n_seeds, _ptr(ids), _ptr(tr))
    return ids, tr


def bruteforce(base, queries, k, metric="l2"):
    """Exact k-NN, ties by smaller id; (ids, dists as float64)."""
    m = METRIC[metric]
    if m == 2:
        base = np.ascontiguousarray(base, dtype=np.uint64)
        queries = np.ascontiguousarray(queries, dtype=np.uint64)
        dim = 256
    else:
        base = np.ascontiguousarray(base, dtype=np.float32)
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        dim = base.shape[1]
    nq = queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    d = np.empty((nq, k), np.float64)
    lib().or_bruteforce(_ptr(base), base.shape[0], _ptr(queries), nq, dim, m, k, _ptr(ids), _ptr(d))
    return ids, d


def bloom_positions(ids, nbits):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.empty((ids.size, 4), np.int32)
    lib().or_bloom_positions(_ptr(ids), ids.size, nbits, _ptr(out))
    return out


def bloom_run(words, inserted, probes):
    ins = np.ascontiguousarray(inserted, dtype=np.int32)
    pr = np.ascontiguousarray(probes, dtype=np.int32)
    hit = np.empty(pr.size, np.int32)
    bits = np.empty(words, np.uint64)
    lib().or_bloom_run(words, _ptr(ins), ins.size, _ptr(pr), pr.size, _ptr(hit), _ptr(bits))
    return hit.astype(bool), bits


def set_run(cap, ops, vals, ids):
    ops = np.ascontiguousarray(ops, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    n = ops.size
    ret = np.empty(n, np.int32)
    val = np.empty(n, np.float32)
    fd = np.empty(max(cap, 1), np.float32)
    fi = np.empty(max(cap, 1), np.int32)
    dr = np.empty(max(cap, 1) + 1, np.int32)
    dd = np.empty(max(cap, 1) + 1, np.float32)
    sz = lib().or_set_run(cap, _ptr(ops), _ptr(vals), _ptr(ids), n, _ptr(ret), _ptr(val),
                          _ptr(fd), _ptr(fi), _ptr(dr), _ptr(dd))
    return ret, val, fd[:sz].copy(), fi[:sz].copy(), dr, dd


def row_dist(metric, q, x, group):
    q = np.ascontiguousarray(q, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().or_row_dist_f32(METRIC[metric], _ptr(q), _ptr(x), q.size, group))


def l2_sequential(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().or_l2_sequential(_ptr(a), _ptr(b), a.size))


# --------------------------------------------------------------------------- reference calls

def ref_build_graph(data):
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, dim = data.shape
    adj = np.empty((n, ref().ref_block_size()), np.int32)
    ref().ref_build_graph(_ptr(data), n, dim, _ptr(adj))
    return adj


def ref_graph_search(data, adj, queries, k, ef, n_threads=1, per_query=False):
    """The reference's Graph::search; returns (ids worst-first, trace, wall_ms, per_query_ms)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    nq = queries.shape[0]
    ids = np.empty((nq, k), np.int32)
    tr = np.empty((nq, 12), np.int32)
    pq = np.empty(nq, np.float64) if per_query else None
    ms = ref().ref_graph_search(_ptr(data), _ptr(adj), data.shape[0], data.shape[1],
                                _ptr(queries), nq, k, ef, n_threads, _ptr(ids), _ptr(tr), _ptr(pq))
    return ids, tr, ms, pq


def ref_l2(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(ref().ref_l2(_ptr(a), _ptr(b), a.size))


def ref_bithash(vecs, seed=42):
    vecs = np.ascontiguousarray(vecs, dtype=np.float32)
    n, dim = vecs.shape
    out = np.empty((n, 8), np.uint32)
    ref().ref_bithash_project(dim, seed, _ptr(vecs), n, _ptr(out))
    return out


def ref_bithash_matrix(dim, seed=42):
    out = np.empty((256, dim), np.float32)
    ref().ref_bithash_matrix(dim, seed, _ptr(out))
    return out
