"""ctypes binding of libsong.so (the C ABI declared in include/song.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsong.so")

METRIC_L2, METRIC_IP, METRIC_HAMMING = 0, 1, 2
METRICS = {"l2": METRIC_L2, "ip": METRIC_IP, "hamming": METRIC_HAMMING}

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


class SongTiming(ctypes.Structure):
    _fields_ = [("alloc_ms", ctypes.c_double), ("h2d_ms", ctypes.c_double),
                ("memset_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("d2h_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("n_batches", _I32), ("ef_used", _I32)]


class SongBatchTiming(ctypes.Structure):
    _fields_ = [("batch_index", _I32), ("batch_offset", _I32), ("batch_size", _I32), ("pad_", _I32),
                ("h2d_ms", ctypes.c_double), ("memset_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("d2h_ms", ctypes.c_double)]


class SongSearchOpts(ctypes.Structure):
    _fields_ = [("bloom_bits", _I32), ("batch", _I32), ("n_seeds", _I32),
                ("waves_per_block", _I32), ("lanes_per_row", _I32), ("mode", _I32),
                ("waves_per_cu", _I32), ("learned", _I32), ("stop_threshold", ctypes.c_float),
                ("stop_min_hops", _I32), ("stop_interval", _I32), ("stop_max_hops", _I32),
                ("stop_confirm", _I32), ("result_cap", _I32), ("stop_late_hops", _I32),
                ("stop_late_interval", _I32), ("entry_late_hops", _I32)]


EXPORTS = (
    "song_abi_version", "song_last_error", "song_device_count", "song_nsw_build",
    "song_graph_save", "song_graph_load_header", "song_graph_load", "song_bithash_matrix",
    "song_bithash_project", "song_bithash_project_device", "song_index_create", "song_index_destroy",
    "song_index_row_stride", "song_index_set_visit_table", "song_index_set_wide_offsets",
    "song_search", "song_search_device", "song_dev_alloc", "song_dev_free", "song_memcpy_h2d",
    "song_memcpy_d2h", "song_queries_to_device", "song_index_bruteforce",
    "song_search_device_mixed", "song_dev_copy", "song_index_set_model",
    "song_learn_trace", "song_mlp_train", "song_mlp_forward", "song_server_create", "song_server_submit",
    "song_server_poll", "song_server_wait", "song_server_stats_get", "song_server_destroy",
    "song_server_run_poisson", "song_search_device_static",
    "song_cpu_search", "song_dev_gather",
)

_lib = None


class SongError(RuntimeError):
    pass


def lib():
    """Load libsong.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SongError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                        "(make -C 2020_song_ex_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    L.song_abi_version.restype = ctypes.c_int
    L.song_last_error.restype = ctypes.c_char_p
    L.song_device_count.argtypes = [_P]
    L.song_nsw_build.argtypes = [_P, _I64, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I64,
                                 _I32, _I32, _P]
    L.song_graph_save.argtypes = [ctypes.c_char_p, _P, _I64, _I32, _I32, _I32]
    L.song_graph_load_header.argtypes = [ctypes.c_char_p, _P, _P, _P, _P]
    L.song_graph_load.argtypes = [ctypes.c_char_p, _P, _I64, _I32]
    L.song_bithash_matrix.argtypes = [_I32, ctypes.c_uint32, _P]
    L.song_bithash_project.argtypes = [_P, _I32, _P, _I64, _I32, _P]
    L.song_bithash_project_device.argtypes = [_P, _I32, _P, _I64, _I32, _P, _P]
    L.song_index_create.argtypes = [_P, _P, _I64, _I32, _I32, _I32, _I32, ctypes.POINTER(_P)]
    L.song_index_destroy.argtypes = [_P]
    L.song_index_destroy.restype = None
    L.song_index_row_stride.argtypes = [_P, _P]
    L.song_index_set_visit_table.argtypes = [_P, _I32]
    L.song_index_set_wide_offsets.argtypes = [_P, _I32]
    L.song_search.argtypes = [_P, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P]
    L.song_search_device.argtypes = [_P, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P]
    L.song_cpu_search.argtypes = [_P, _P, _I64, _I32, _I32, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P,
                                  _P]
    L.song_search_device_static.argtypes = [_P, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P]
    L.song_dev_alloc.argtypes = [_I32, _I64, ctypes.POINTER(_P)]
    L.song_dev_free.argtypes = [_P]
    L.song_memcpy_h2d.argtypes = [_P, _P, _I64]
    L.song_memcpy_d2h.argtypes = [_P, _P, _I64]
    L.song_queries_to_device.argtypes = [_P, _P, _I32, ctypes.POINTER(_P)]
    L.song_index_bruteforce.argtypes = [_P, _P, _I32, _I32, _P, _P, _P]
    L.song_dev_copy.argtypes = [_P, _P, _I64, _P]
    L.song_dev_gather.argtypes = [_P, _P, _I64, _I32, _P, _I64, _P]
    L.song_search_device_mixed.argtypes = [_P, _P, _I32, _P, _P, _I32, _I32, _P, _P, _P, _P, _P]
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().song_last_error().decode(errors="replace")
        raise SongError(f"{what} failed ({rc}): {msg}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def device_count() -> int:
    c = _I32(0)
    check(lib().song_device_count(ctypes.byref(c)), "song_device_count")
    return c.value


MODE_PLAIN, MODE_PERSISTENT = 0, 1


def make_opts(bloom_bits=0, batch=0, n_seeds=0, waves_per_block=0, lanes_per_row=0, mode=0,
              waves_per_cu=0, learned=0, stop_threshold=0.0, stop_min_hops=0, stop_interval=1,
              stop_max_hops=0, stop_confirm=0, result_cap=0, stop_late_hops=0, stop_late_interval=0,
              entry_late_hops=0):
    o = SongSearchOpts()
    o.entry_late_hops = entry_late_hops
    o.result_cap = result_cap
    o.stop_late_hops = stop_late_hops
    o.stop_late_interval = stop_late_interval
    o.learned = learned
    o.stop_threshold = stop_threshold
    o.stop_min_hops = stop_min_hops
    o.stop_interval = stop_interval
    o.stop_max_hops = stop_max_hops
    o.stop_confirm = stop_confirm
    o.mode = mode
    o.waves_per_cu = waves_per_cu
    o.bloom_bits = bloom_bits
    o.batch = batch
    o.n_seeds = n_seeds
    o.waves_per_block = waves_per_block
    o.lanes_per_row = lanes_per_row
    return o


class DeviceBuffer:
    """A plain device allocation owned by Python (freed on close / GC)."""

    def __init__(self, nbytes: int, device: int = 0, handle=None):
        self.nbytes = int(nbytes)
        if handle is None:
            h = _P()
            check(lib().song_dev_alloc(device, self.nbytes, ctypes.byref(h)), "song_dev_alloc")
            handle = h
        self.handle = handle

    @property
    def ptr(self):
        return self.handle

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().song_memcpy_h2d(self.handle, ptr(arr), arr.nbytes), "song_memcpy_h2d")
        return self

    def download(self, arr: np.ndarray) -> np.ndarray:
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        check(lib().song_memcpy_d2h(ptr(arr), self.handle, arr.nbytes), "song_memcpy_d2h")
        return arr

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().song_dev_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
