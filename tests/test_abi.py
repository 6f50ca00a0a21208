"""The C ABI: every symbol include/song.h declares is exported by libsong.so,
and argument errors come back as status codes (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "song.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(song_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(song):
    lib = song._native.lib()
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(song._native.EXPORTS)


def test_abi_version(song):
    assert song._native.lib().song_abi_version() == 2


def test_oracle_libraries_load(oracle):
    assert oracle.lib().or_bloom_pos(0, 0, 16384) >= 0


def test_index_rejects_bad_adjacency(song):
    """Checked on the host before any device call."""
    data = np.zeros((10, 4), np.float32)
    adj = np.full((10, 8), -1, np.int32)
    adj[3, 2] = 10  # == n: out of range
    with pytest.raises(song.SongError, match="adjacency"):
        song.Index(data, adj)
    adj[3, 2] = -2
    with pytest.raises(song.SongError, match="adjacency"):
        song.Index(data, adj)


def test_index_rejects_bad_shapes(song):
    with pytest.raises(song.SongError):
        song.Index(np.zeros((10, 4), np.float32), np.full((10, 65), -1, np.int32))
    with pytest.raises(song.SongError):
        song.Index(np.zeros((10, 4097), np.float32), np.full((10, 8), -1, np.int32))


def test_null_arguments(song):
    L = song._native.lib()
    assert L.song_index_row_stride(None, None) == -1
    assert L.song_search(None, None, 0, 10, 10, None, None, None, None, None, None) == -1
    assert b"null" in L.song_last_error()
    assert L.song_graph_load_header(b"/nonexistent/graph.bin", ctypes.byref(ctypes.c_int64()),
                                    ctypes.byref(ctypes.c_int32()), ctypes.byref(ctypes.c_int32()),
                                    ctypes.byref(ctypes.c_int32())) == -4


def test_device_count_does_not_fail(song):
    assert song.device_count() >= 0


def test_autotuner_pick_is_deterministic():
    """The tuner tries the grid in product order and picks, among candidates whose recall meets
    the target, the highest throughput -- the earlier one on a tie; none when none meets it."""
    import importlib
    at = importlib.import_module("2020_song_ex_amd.autotune")
    grid = {"ef": [64, 128], "mode": [0, 1], "bloom_bits": [1 << 13]}
    order = at.candidates(grid)
    assert order == [{"ef": 64, "mode": 0, "bloom_bits": 8192}, {"ef": 64, "mode": 1, "bloom_bits": 8192},
                     {"ef": 128, "mode": 0, "bloom_bits": 8192}, {"ef": 128, "mode": 1, "bloom_bits": 8192}]
    C = at.Candidate
    tried = [C(order[0], recall=0.85, qps=9e6), C(order[1], recall=0.91, qps=5e6),
             C(order[2], recall=0.95, qps=5e6), C(order[3], recall=0.97, qps=3e6)]
    assert at.pick(tried, 0.90) is tried[1]  # fastest above the target; the tie keeps the earlier
    assert at.pick(tried, 0.80) is tried[0]
    assert at.pick(tried, 0.96) is tried[3]
    assert at.pick(tried, 0.99) is None
    assert at.pick(list(reversed(tried)), 0.90) is tried[2]  # order decides ties
    # a dict value sets several knobs together
    assert at.candidates({"ef": [64], "late": [{"a": 1, "b": 2}, {"a": 0, "b": 0}]}) == [
        {"ef": 64, "a": 1, "b": 2}, {"ef": 64, "a": 0, "b": 0}]


def test_autotuner_finalists():
    """The finalists are the fastest candidates that meet the recall target, fastest first;
    the final pick is the lowest median of their interleaved launch times."""
    import importlib
    at = importlib.import_module("2020_song_ex_amd.autotune")
    C = at.Candidate
    tried = [C({"ef": 1}, recall=0.85, qps=9e6), C({"ef": 2}, recall=0.91, qps=5e6),
             C({"ef": 3}, recall=0.95, qps=6e6), C({"ef": 4}, recall=0.97, qps=5e6),
             C({"ef": 5}, recall=0.99, qps=1e6)]
    f = at.finalists(tried, 0.90, 3)
    assert f == [tried[2], tried[1], tried[3]]  # 6e6, then the 5e6 tie in grid order
    assert at.finalists(tried, 0.999, 3) == []
    assert at.finalists(tried, 0.90, 10) == [tried[2], tried[1], tried[3], tried[4]]
    # one slow outlier does not decide: medians 2.0, 1.9, 2.1
    times = [[2.0, 2.0, 1.5, 2.2, 2.0], [1.9, 3.0, 1.9, 1.8, 1.9], [2.1, 2.1, 2.1, 2.1, 2.1]]
    assert at.pick_final(f, times) is tried[1]
    assert at.pick_final(f, [[1.0], [1.0], [0.5]]) is tried[3]
    assert at.pick_final(f, [[1.0], [1.0], [1.0]]) is tried[2]  # equal medians keep the earlier


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_index_rejects_non_finite_vectors(song, bad):
    """Checked on the host before any device call: a NaN distance would order above every
    finite key, where the reference's float scans treat it by position."""
    data = np.zeros((10, 4), np.float32)
    data[7, 2] = bad
    with pytest.raises(song.SongError, match="finite"):
        song.Index(data, np.full((10, 8), -1, np.int32))
