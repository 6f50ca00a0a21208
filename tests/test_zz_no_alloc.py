"""The allocation-free entry point, song_search_device_static: the same ids, distances and
traces as song_search_device on two calls in a row (the second with fewer queries), for the
plain, persistent and learned (entry seeds + cost-ordered queue + stop) searches, and its
size limit."""
import ctypes
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

learned = importlib.import_module("2020_song_ex_amd.learned")


@pytest.fixture(scope="module")
def setup(song, datasets):
    base, q = datasets.gmm_int(8000, 400, dim=64, n_centres=20, sigma=20.0, seed=21)
    adj = song.nsw_build(base, block=16, exact_order=False)
    idx = song.Index(base, adj)
    train = datasets.gmm_int(8000, 2400, dim=64, n_centres=20, sigma=20.0, seed=21)[1][400:]
    model = learned.fit(idx, base, train, k=10, ef=96, entry_epochs=6, stop_epochs=2, cost_epochs=6,
                        stop_min_hops=8, stop_interval=2)
    yield base, q, idx, model
    idx.close()


def _run(song, idx, dq, nq, k, ef, no_alloc, **kw):
    N = song._native
    d_ids = N.DeviceBuffer(nq * k * 4)
    d_d = N.DeviceBuffer(nq * k * 4)
    d_tr = N.DeviceBuffer(nq * 48)
    idx.search_device(dq, nq, k, ef, d_ids, d_d, d_tr, no_alloc=no_alloc, **kw)
    return (d_ids.download(np.empty((nq, k), np.int32)), d_d.download(np.empty((nq, k), np.float32)),
            d_tr.download(np.empty((nq, 12), np.int32)))


CASES = {
    "plain": dict(mode=0),
    "persistent": dict(mode=1, waves_per_cu=4),
    "learned": dict(mode=1, learned=7, stop_threshold=0.5, stop_min_hops=8, stop_interval=2,
                    stop_max_hops=200),
}


@pytest.mark.parametrize("case", list(CASES))
def test_static_equals_device_entry_point(song, setup, case):
    base, q, idx, model = setup
    kw = CASES[case]
    for nq in (400, 150):  # the second call is smaller: nothing stale from the first may leak in
        dq = idx.queries_to_device(q[:nq])
        want = _run(song, idx, dq, nq, 10, 96, False, **kw)
        got = _run(song, idx, dq, nq, 10, 96, True, **kw)
        for w, g in zip(want, got):
            np.testing.assert_array_equal(g, w)
        assert (got[0] >= 0).all()


def test_static_batches(song, setup):
    """Batches of the static entry point share its one scratch array in stream order."""
    base, q, idx, model = setup
    dq = idx.queries_to_device(q)
    kw = dict(CASES["learned"])
    want = _run(song, idx, dq, q.shape[0], 10, 96, False, **kw)
    got = _run(song, idx, dq, q.shape[0], 10, 96, True, batch=64, **kw)
    for w, g in zip(want, got):
        np.testing.assert_array_equal(g, w)


def test_static_size_limit(song, setup):
    """A planning batch beyond the scratch array is refused before any launch."""
    base, q, idx, model = setup
    N = song._native
    L = N.lib()
    limit = 66 * 16384 // (64 + 2)  # n_seeds = 64: 16384 queries per batch
    nq = limit + 1
    dq = N.DeviceBuffer(nq * idx.row_stride * 4)
    d_ids = N.DeviceBuffer(nq * 10 * 4)
    o = N.make_opts(n_seeds=64, learned=1, mode=1)
    rc = L.song_search_device_static(idx.handle, dq.ptr, nq, 10, 96, ctypes.byref(o), d_ids.ptr, None, None,
                                     None)
    assert rc == -1
    assert b"at most 16384 queries" in L.song_last_error()
    # without the planning pass only the queue head is needed: no limit on nq
    o = N.make_opts(mode=1)
    small = idx.queries_to_device(q[:50])
    assert L.song_search_device_static(idx.handle, small.ptr, 50, 10, 96, ctypes.byref(o), d_ids.ptr, None,
                                       None, None) == 0
