"""Argument checks that need a resident index: per-query budgets of the mixed-budget entry
point and non-finite queries are refused before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(song, datasets):
    base, q = datasets.gmm_int(3000, 64, dim=32, n_centres=10, sigma=20.0, seed=2)
    adj = song.nsw_build(base, block=16, exact_order=False)
    idx = song.Index(base, adj)
    yield base, q, adj, idx
    idx.close()


@pytest.mark.parametrize("ks_bad,efs_bad,msg", [
    ((5, 0), None, "k must be in"),      # k < 1 would write before the query's row
    ((5, -3), None, "k must be in"),
    ((5, 11), None, "k must be in"),     # k > k_max
    (None, (2, 65), "ef_max"),            # ef > ef_max
    ((9, 10), (9, 70), "ef_max"),         # max(ef, k) > ef_max
])
def test_mixed_budgets_are_checked(song, small, ks_bad, efs_bad, msg):
    base, q, adj, idx = small
    N = song._native
    nq = q.shape[0]
    ks = np.full(nq, 10, np.int32)
    efs = np.full(nq, 64, np.int32)
    if ks_bad:
        ks[ks_bad[0]] = ks_bad[1]
    if efs_bad:
        efs[efs_bad[0]] = efs_bad[1]
    dq = idx.queries_to_device(q)
    d_ks = N.DeviceBuffer(nq * 4).upload(ks)
    d_efs = N.DeviceBuffer(nq * 4).upload(efs)
    sentinel = np.full((nq, 10), 12345, np.int32)
    d_ids = N.DeviceBuffer(nq * 10 * 4).upload(sentinel)
    with pytest.raises(song.SongError, match=msg):
        idx.search_device_mixed(dq, nq, d_ks, d_efs, 10, 64, d_ids)
    # nothing ran: the output buffer is untouched
    np.testing.assert_array_equal(d_ids.download(np.empty((nq, 10), np.int32)), sentinel)
    # the same call with valid budgets runs and equals the uniform search
    ks[:] = 10
    efs[:] = 64
    d_ks.upload(ks)
    d_efs.upload(efs)
    idx.search_device_mixed(dq, nq, d_ks, d_efs, 10, 64, d_ids)
    want = idx.search(q, 10, 64)
    np.testing.assert_array_equal(d_ids.download(np.empty((nq, 10), np.int32)), want.ids)


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_queries_are_refused(song, small, bad):
    base, q, adj, idx = small
    qb = q.copy()
    qb[3, 5] = bad
    with pytest.raises(song.SongError, match="finite"):
        idx.search(qb, 10, 64)
    with pytest.raises(song.SongError, match="finite"):
        idx.queries_to_device(qb)
    with pytest.raises(song.SongError, match="finite"):
        idx.bruteforce(qb, 10)
    serving = __import__("importlib").import_module("2020_song_ex_amd.serving")
    sv = serving.Server(idx, max_batch=16, n_streams=1, k_max=10, ef_max=64)
    try:
        with pytest.raises(song.SongError, match="finite"):
            sv.submit(qb[3], 10, 64)
    finally:
        sv.close()
    # finite queries still search
    assert idx.search(q, 10, 64).ids.shape == (q.shape[0], 10)
