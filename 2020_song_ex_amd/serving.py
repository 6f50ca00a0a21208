"""Continuous-batching serving loop (libsong's song_server_*): requests queue
from any thread, a native dispatcher admits them into launch slots as slots
and device memory free up, and each request's latency is measured from its
arrival to its completion."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .index import Index

_I32, _I64, _P = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p


class SongServerOpts(ctypes.Structure):
    _fields_ = [("max_batch", _I32), ("n_streams", _I32), ("memory_budget", _I64), ("k_max", _I32),
                ("ef_max", _I32), ("search", N.SongSearchOpts), ("resident", _I32), ("idle_ms", _I32)]


class SongServerStats(ctypes.Structure):
    _fields_ = [("submitted", _I64), ("completed", _I64), ("in_flight", _I64), ("launches", _I64),
                ("deferred", _I64), ("peak_bytes", _I64), ("memory_budget", _I64), ("slot_capacity", _I32),
                ("pad_", _I32)]


def _bind():
    L = N.lib()
    if not hasattr(L, "_server_bound"):
        L.song_server_create.argtypes = [_P, _P, ctypes.POINTER(_P)]
        L.song_server_submit.argtypes = [_P, _P, _I32, _I32, ctypes.POINTER(_I64)]
        L.song_server_poll.argtypes = [_P, _I64, _P, _P, _P]
        L.song_server_wait.argtypes = [_P, _I64, _I32, _P, _P, _P]
        L.song_server_stats_get.argtypes = [_P, _P]
        L.song_server_destroy.argtypes = [_P]
        L.song_server_destroy.restype = None
        L.song_server_run_poisson.argtypes = [_P, _P, _I32, _P, _P, ctypes.c_double, ctypes.c_uint64, _P, _P,
                                              _P, _P]
        L._server_bound = True
    return L


class Server:
    def __init__(self, index: Index, *, max_batch=1024, n_streams=4, memory_budget=0, k_max=100, ef_max=512,
                 resident=False, idle_ms=0, **search):
        L = _bind()
        self.index = index
        self.k_max = k_max
        o = SongServerOpts()
        o.max_batch, o.n_streams, o.memory_budget, o.k_max, o.ef_max = max_batch, n_streams, memory_budget, \
            k_max, ef_max
        o.search = N.make_opts(**search)
        o.resident, o.idle_ms = int(bool(resident)), idle_ms
        h = _P()
        N.check(L.song_server_create(index.handle, ctypes.byref(o), ctypes.byref(h)), "song_server_create")
        self._h = h

    def _rows(self, q):
        """Queries in the index's device row layout (dim padded to the row stride)."""
        if self.index.metric == N.METRIC_HAMMING:
            return np.ascontiguousarray(q, np.uint64).reshape(-1, 4)
        q = np.asarray(q, np.float32).reshape(-1, self.index.dim)
        out = np.zeros((q.shape[0], self.index.row_stride), np.float32)
        out[:, :self.index.dim] = q
        return out

    def submit(self, query, k, ef):
        row = self._rows(query)[0]
        t = _I64()
        N.check(_bind().song_server_submit(self._h, N.ptr(row), k, ef, ctypes.byref(t)), "song_server_submit")
        return t.value

    def wait(self, ticket, k, timeout_ms=60000):
        ids = np.empty(k, np.int32)
        d = np.empty(k, np.float32)
        lat = ctypes.c_double()
        rc = _bind().song_server_wait(self._h, ticket, timeout_ms, N.ptr(ids), N.ptr(d), ctypes.byref(lat))
        if rc < 0:
            N.check(rc, "song_server_wait")
        return (ids, d, lat.value) if rc == 1 else None

    def poll(self, ticket, k):
        ids = np.empty(k, np.int32)
        d = np.empty(k, np.float32)
        lat = ctypes.c_double()
        rc = _bind().song_server_poll(self._h, ticket, N.ptr(ids), N.ptr(d), ctypes.byref(lat))
        if rc < 0:
            N.check(rc, "song_server_poll")
        return (ids, d, lat.value) if rc == 1 else None

    def stats(self):
        st = SongServerStats()
        N.check(_bind().song_server_stats_get(self._h, ctypes.byref(st)), "song_server_stats_get")
        return {f: getattr(st, f) for f, _ in SongServerStats._fields_ if f != "pad_"}

    def run_poisson(self, queries, ks, efs, rate_qps, seed=0):
        """Poisson arrivals at rate_qps; returns (ids [nq, k_max], dists, latency_ms [nq], wall_ms)."""
        rows = self._rows(queries)
        nq = rows.shape[0]
        ks = np.ascontiguousarray(ks, np.int32)
        efs = np.ascontiguousarray(efs, np.int32)
        ids = np.empty((nq, self.k_max), np.int32)
        d = np.empty((nq, self.k_max), np.float32)
        lat = np.empty(nq, np.float64)
        wall = ctypes.c_double()
        N.check(_bind().song_server_run_poisson(self._h, N.ptr(rows), nq, N.ptr(ks), N.ptr(efs), rate_qps, seed,
                                                N.ptr(ids), N.ptr(d), N.ptr(lat), ctypes.byref(wall)),
                "song_server_run_poisson")
        return ids, d, lat, wall.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().song_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
