"""Python front end over the C ABI: NSW build, device index, search.

Every call goes to libsong.so; there is no other compute path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

TRACE_FIELDS = (
    "seed_count", "visited_nodes", "expanded_nodes", "neighbor_slots_scanned",
    "candidate_nodes", "distance_evals", "frontier_pushes", "result_updates",
    "peak_frontier", "peak_results", "result_count", "early_stopped",
)

# config.h:9-24 of the reference: BLOCK_SIZE 32, FIXED_DEGREE 31,
# SEARCH_DEGREE 15, EF_CONSTRUCTION 150.
REF_BLOCK, REF_FIXED, REF_SEARCH, REF_EF_CONSTRUCTION = 32, 31, 15, 150


def degrees_for_block(block: int):
    """Out-degree D -> (block, fixed, search) = (D, D-1, D/2-1); D=32 gives config.h's 32/31/15."""
    return block, block - 1, max(1, block // 2 - 1)


def _metric(m) -> int:
    return N.METRICS[m] if isinstance(m, str) else int(m)


def _vec_array(vecs, metric: int) -> np.ndarray:
    if metric == N.METRIC_HAMMING:
        return np.ascontiguousarray(vecs, dtype=np.uint64)
    return np.ascontiguousarray(vecs, dtype=np.float32)


def nsw_build(vecs, metric="l2", block=REF_BLOCK, fixed_deg=None, search_deg=None,
              ef_construction=REF_EF_CONSTRUCTION, n_threads=1, batch_div=0, seq_prefix=1,
              exact_order=True, verbose=False) -> np.ndarray:
    """CPU NSW build (build_kernel_graph, kernalGraph.h:140-149); returns adj [n, block]."""
    m = _metric(metric)
    v = _vec_array(vecs, m)
    n = v.shape[0]
    dim = 256 if m == N.METRIC_HAMMING else v.shape[1]
    _, f, s = degrees_for_block(block)
    fixed_deg = f if fixed_deg is None else fixed_deg
    search_deg = s if search_deg is None else search_deg
    adj = np.empty((n, block), np.int32)
    N.check(N.lib().song_nsw_build(N.ptr(v), n, dim, m, block, fixed_deg, search_deg, ef_construction,
                                   n_threads, batch_div, seq_prefix, int(bool(exact_order)),
                                   int(bool(verbose)), N.ptr(adj)), "song_nsw_build")
    return adj


def cpu_search(vecs, adj, queries, k, ef, *, n_threads=1, n_seeds=0, best_first=True):
    """The reference's CPU search, Graph::search (graph.h:33-115), through song_cpu_search on
    the host.  Returns (ids [nq, k], dists, trace [nq, 12], per-query ms [nq], search ms); ids
    best first unless best_first=False (Graph::search's own worst-first order)."""
    v = np.ascontiguousarray(vecs, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, v.shape[1])
    a = np.ascontiguousarray(adj, dtype=np.int32)
    nq = q.shape[0]
    ids = np.empty((nq, k), np.int32)
    d = np.empty((nq, k), np.float32)
    tr = np.empty((nq, 12), np.int32)
    per = np.empty(nq, np.float64)
    total = ctypes.c_double()
    N.check(N.lib().song_cpu_search(N.ptr(v), N.ptr(a), v.shape[0], v.shape[1], a.shape[1], N.ptr(q), nq, k,
                                    ef, n_seeds, n_threads, N.ptr(ids), N.ptr(d), N.ptr(tr), N.ptr(per),
                                    ctypes.byref(total)), "song_cpu_search")
    if best_first:
        # worst first with -1 padding at the end -> best first, padding kept at the end
        cnt = tr[:, 10]
        for i in range(nq):
            c = int(cnt[i])
            ids[i, :c] = ids[i, :c][::-1].copy()
            d[i, :c] = d[i, :c][::-1].copy()
    return ids, d, tr, per, total.value


def graph_save(path, adj, metric="l2", dim=0):
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    N.check(N.lib().song_graph_save(os.fsencode(path), N.ptr(adj), adj.shape[0], adj.shape[1],
                                    _metric(metric), dim), "song_graph_save")


def graph_load(path):
    n = ctypes.c_int64()
    b, m, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    L = N.lib()
    N.check(L.song_graph_load_header(os.fsencode(path), ctypes.byref(n), ctypes.byref(b),
                                     ctypes.byref(m), ctypes.byref(d)), "song_graph_load_header")
    adj = np.empty((n.value, b.value), np.int32)
    N.check(L.song_graph_load(os.fsencode(path), N.ptr(adj), n.value, b.value), "song_graph_load")
    return adj, {"metric": m.value, "dim": d.value}


def bithash_matrix(dim: int, seed: int = 42) -> np.ndarray:
    """BitHash's projection matrix (bitHash.h:22-26), [256, dim]."""
    out = np.empty((256, dim), np.float32)
    N.check(N.lib().song_bithash_matrix(dim, seed, N.ptr(out)), "song_bithash_matrix")
    return out


def bithash_project(R, vecs, n_threads=1) -> np.ndarray:
    """BitHash::project of each row (bitHash.h:32-44) as uint64 [n, 4]."""
    R = np.ascontiguousarray(R, dtype=np.float32)
    v = np.ascontiguousarray(vecs, dtype=np.float32)
    out = np.empty((v.shape[0], 4), np.uint64)
    N.check(N.lib().song_bithash_project(N.ptr(R), R.shape[1], N.ptr(v), v.shape[0], n_threads,
                                         N.ptr(out)), "song_bithash_project")
    return out


def bithash_project_device(R, vecs, device=0):
    """The same codes computed on the GPU (song_bithash_project_device): (uint64 [n, 4], kernel ms)."""
    R = np.ascontiguousarray(R, dtype=np.float32)
    v = np.ascontiguousarray(vecs, dtype=np.float32)
    n, dim = v.shape
    d_R = N.DeviceBuffer(R.nbytes, device).upload(R)
    d_v = N.DeviceBuffer(max(v.nbytes, 4), device).upload(v)
    d_c = N.DeviceBuffer(max(n * 32, 8), device)
    ms = ctypes.c_double(0.0)
    N.check(N.lib().song_bithash_project_device(d_R.ptr, dim, d_v.ptr, n, device, d_c.ptr, ctypes.byref(ms)),
            "song_bithash_project_device")
    return d_c.download(np.empty((n, 4), np.uint64)), ms.value


@dataclass
class SearchResult:
    ids: np.ndarray
    dists: np.ndarray | None
    trace: np.ndarray | None          # [nq, 12] int32, TRACE_FIELDS order
    timing: dict
    batches: list = field(default_factory=list)


class Index:
    """Device-resident vectors + adjacency (KernelGraph::upload_to_gpu, kernalGraph.h:38-43)."""

    def __init__(self, vecs, adj, metric="l2", device=0):
        self.metric = _metric(metric)
        v = _vec_array(vecs, self.metric)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        self.n = v.shape[0]
        self.dim = 256 if self.metric == N.METRIC_HAMMING else v.shape[1]
        self.block = a.shape[1]
        self.device = device
        h = ctypes.c_void_p()
        N.check(N.lib().song_index_create(N.ptr(v), N.ptr(a), self.n, self.dim, self.block,
                                          self.metric, device, ctypes.byref(h)), "song_index_create")
        self._h = h
        rs = ctypes.c_int32()
        N.check(N.lib().song_index_row_stride(self._h, ctypes.byref(rs)), "song_index_row_stride")
        self.row_stride = rs.value

    @property
    def handle(self):
        return self._h

    def set_visit_table(self, enabled):
        """Build (True) or free (False) the per-slot visited-filter positions the search reads
        instead of hashing each neighbour id (built by default for rows of at most 32 slots)."""
        N.check(N.lib().song_index_set_visit_table(self._h, 1 if enabled else 0), "song_index_set_visit_table")

    def set_wide_offsets(self, enabled):
        """Address vector rows with 64-bit offsets (the path of indexes above 2 GiB) on this index too."""
        N.check(N.lib().song_index_set_wide_offsets(self._h, 1 if enabled else 0), "song_index_set_wide_offsets")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().song_index_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _queries(self, q):
        q = _vec_array(q, self.metric)
        if q.ndim != 2 or q.shape[1] != (4 if self.metric == N.METRIC_HAMMING else self.dim):
            raise ValueError(f"queries must be [nq, {self.dim if self.metric != 2 else 4}]")
        return q

    def search(self, queries, k, ef, *, bloom_bits=0, batch=0, n_seeds=0, waves_per_block=0,
               lanes_per_row=0, mode=0, waves_per_cu=0, learned=0, stop_threshold=0.0, stop_min_hops=0,
               stop_interval=1, stop_max_hops=0, stop_confirm=0, result_cap=0, stop_late_hops=0, stop_late_interval=0,
               entry_late_hops=0, dists=True, trace=True) -> SearchResult:
        """gpu_search (kernalGraph.h:307-382): host queries in, best-first ids out."""
        q = self._queries(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        d = np.empty((nq, k), np.float32) if dists else None
        tr = np.empty((nq, 12), np.int32) if trace else None
        o = N.make_opts(bloom_bits, batch, n_seeds, waves_per_block, lanes_per_row, mode, waves_per_cu,
                        learned, stop_threshold, stop_min_hops, stop_interval, stop_max_hops, stop_confirm,
                        result_cap, stop_late_hops, stop_late_interval, entry_late_hops)
        tm = N.SongTiming()
        bsz = batch if batch > 0 else max(nq, 1)
        nb = max(1, -(-nq // bsz))
        bt = (N.SongBatchTiming * nb)()
        N.check(N.lib().song_search(self._h, N.ptr(q), nq, k, ef, ctypes.byref(o), N.ptr(ids), N.ptr(d),
                                    N.ptr(tr), ctypes.byref(tm), bt), "song_search")
        timing = {f: getattr(tm, f) for f, _ in N.SongTiming._fields_}
        batches = [{f: getattr(b, f) for f, _ in N.SongBatchTiming._fields_ if f != "pad_"}
                   for b in bt[: tm.n_batches]]
        return SearchResult(ids, d, tr, timing, batches)

    def bruteforce(self, queries, k):
        """Exact k-NN on the device over the resident vectors: (ids, dists, kernel_ms)."""
        q = self._queries(queries)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.int32)
        d = np.empty((nq, k), np.float32)
        ms = ctypes.c_double(0.0)
        N.check(N.lib().song_index_bruteforce(self._h, N.ptr(q), nq, k, N.ptr(ids), N.ptr(d),
                                              ctypes.byref(ms)), "song_index_bruteforce")
        return ids, d, ms.value

    def queries_to_device(self, queries) -> N.DeviceBuffer:
        q = self._queries(queries)
        h = ctypes.c_void_p()
        N.check(N.lib().song_queries_to_device(self._h, N.ptr(q), q.shape[0], ctypes.byref(h)),
                "song_queries_to_device")
        esz = 8 if self.metric == N.METRIC_HAMMING else 4
        return N.DeviceBuffer(q.shape[0] * self.row_stride * esz, self.device, handle=h)

    def search_device(self, dq: N.DeviceBuffer, nq, k, ef, d_ids: N.DeviceBuffer,
                      d_dists: N.DeviceBuffer | None = None, d_trace: N.DeviceBuffer | None = None, *,
                      bloom_bits=0, batch=0, n_seeds=0, waves_per_block=0, lanes_per_row=0, mode=0,
                      waves_per_cu=0, learned=0, stop_threshold=0.0, stop_min_hops=0, stop_interval=1,
                      stop_max_hops=0, stop_confirm=0, result_cap=0, stop_late_hops=0, stop_late_interval=0,
                      entry_late_hops=0,
                      no_alloc=False) -> dict:
        """Search with every buffer resident on the device; returns the timing dict.
        no_alloc: through song_search_device_static (no device allocation; its scratch is a
        static array per GPU, so calls on one GPU must not overlap)."""
        o = N.make_opts(bloom_bits, batch, n_seeds, waves_per_block, lanes_per_row, mode, waves_per_cu,
                        learned, stop_threshold, stop_min_hops, stop_interval, stop_max_hops, stop_confirm,
                        result_cap, stop_late_hops, stop_late_interval, entry_late_hops)
        tm = N.SongTiming()
        fn = N.lib().song_search_device_static if no_alloc else N.lib().song_search_device
        N.check(fn(self._h, dq.ptr, nq, k, ef, ctypes.byref(o), d_ids.ptr,
                   d_dists.ptr if d_dists else None, d_trace.ptr if d_trace else None, ctypes.byref(tm)),
                "song_search_device_static" if no_alloc else "song_search_device")
        return {f: getattr(tm, f) for f, _ in N.SongTiming._fields_}

    def search_device_mixed(self, dq: N.DeviceBuffer, nq, d_ks: N.DeviceBuffer, d_efs: N.DeviceBuffer,
                            k_max, ef_max, d_ids: N.DeviceBuffer, d_dists: N.DeviceBuffer | None = None,
                            d_trace: N.DeviceBuffer | None = None, *, bloom_bits=0, n_seeds=0,
                            waves_per_block=0, lanes_per_row=0, mode=0, waves_per_cu=0, learned=0,
                            stop_threshold=0.0, stop_min_hops=0, stop_interval=1, stop_max_hops=0,
                            stop_confirm=0, result_cap=0, stop_late_hops=0, stop_late_interval=0,
                            entry_late_hops=0) -> dict:
        """Per-query k / queue size in one launch (rows of k_max ids)."""
        o = N.make_opts(bloom_bits, 0, n_seeds, waves_per_block, lanes_per_row, mode, waves_per_cu,
                        learned, stop_threshold, stop_min_hops, stop_interval, stop_max_hops, stop_confirm,
                        result_cap, stop_late_hops, stop_late_interval, entry_late_hops)
        tm = N.SongTiming()
        N.check(N.lib().song_search_device_mixed(self._h, dq.ptr, nq, d_ks.ptr, d_efs.ptr, k_max, ef_max,
                                                 ctypes.byref(o), d_ids.ptr,
                                                 d_dists.ptr if d_dists else None,
                                                 d_trace.ptr if d_trace else None, ctypes.byref(tm)),
                "song_search_device_mixed")
        return {f: getattr(tm, f) for f, _ in N.SongTiming._fields_}
