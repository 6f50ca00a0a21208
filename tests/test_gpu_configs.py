"""The benchmark's other workloads as parity cases, at their real dimensions and data
distributions, with the base scaled down so that the CPU graph build stays within seconds:

* 960-d Beta-perturbed mixture (1,000,000 x 960 in the benchmark; 100,000 here): every
  configuration the autotuner tries for the persistent kernel (queue, waves per CU, Bloom size)
  returns the oracle's search with that queue and Bloom size, and the persistent grid returns
  the plain launch.
* 256-bit Hamming codes of a 2,000-centre vMF mixture, at the full 1,000,000 codes: the oracle
  at two queue sizes, persistent = plain, and the learned entry + stop trained on held-out codes.
* Inner product on a 5,000-centre vMF mixture, dim 200 (1,200,000 vectors; 200,000 here), with
  requests of k in {1, 10, 100} and queue budgets in {64, 256}: each request of a mixed launch
  and of the serving loop is the oracle's / the uniform search with its own k and queue.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

autotune = importlib.import_module("2020_song_ex_amd.autotune")
learned = importlib.import_module("2020_song_ex_amd.learned")
serving = importlib.import_module("2020_song_ex_amd.serving")


def build(song, base, metric):
    return song.nsw_build(base, metric=metric, block=32, n_threads=16, batch_div=64, seq_prefix=1024,
                          exact_order=False)


def same(r, o, label):
    oi, od, ot = o
    bad = np.where((r.ids != oi).any(1))[0]
    assert bad.size == 0, f"{label}: ids differ for queries {bad[:10]}"
    np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32), err_msg=label)
    np.testing.assert_array_equal(r.trace, ot, err_msg=label)


def recall(ids, gt):
    return float(np.mean([len(set(a) & set(b)) / gt.shape[1] for a, b in zip(ids, gt)]))


# ---------------------------------------------------------------- 960-d, persistent, autotuned

@pytest.fixture(scope="module")
def c3(song, datasets):
    cfg = datasets.config3(n=100_000, nq=200)
    adj = build(song, cfg["base"], "l2")
    idx = song.Index(cfg["base"], adj)
    yield cfg["base"], cfg["queries"], adj, idx
    idx.close()


def test_config3_autotuned_persistent(song, oracle, c3):
    base, q, adj, idx = c3
    gt, _, _ = idx.bruteforce(q, 10)
    grid = {"ef": [64, 128], "mode": [1], "waves_per_cu": [8, 32], "bloom_bits": [1 << 13, 1 << 15]}
    res = autotune.tune(idx, q, gt, 10, grid, 0.0)
    assert len(res.tried) == 8 and res.best is not None
    sample = q[::10]
    for c in res.tried:
        p = dict(c.params)
        ef = p.pop("ef")
        r = idx.search(sample, 10, ef, **p)
        # 64 lanes per 960-float row, 4 chunks each
        o = oracle.search(base, adj, sample, 10, ef, block=32, bloom_bits=p["bloom_bits"], group=64)
        same(r, o, f"config 3 {c.params}")
        plain = idx.search(q, 10, ef, bloom_bits=p["bloom_bits"])
        pers = idx.search(q, 10, ef, **p)
        np.testing.assert_array_equal(pers.ids, plain.ids)
        np.testing.assert_array_equal(pers.trace, plain.trace)
        assert abs(c.recall - recall(plain.ids, gt)) < 1e-9
    r64 = [c.recall for c in res.tried if c.params["ef"] == 64]
    r128 = [c.recall for c in res.tried if c.params["ef"] == 128]
    assert min(r128) >= max(r64) - 0.02  # the longer queue finds at least as much


# ---------------------------------------------------------------- Hamming codes, learned

@pytest.fixture(scope="module")
def c4(song, datasets):
    cfg = datasets.config4(n=1_000_000, nq=1000, n_extra=4000)
    adj = build(song, cfg["base"], "hamming")
    idx = song.Index(cfg["base"], adj, metric="hamming")
    yield cfg["base"], cfg["queries"], cfg["extra"], adj, idx
    idx.close()


def test_config4_hamming_matches_oracle(song, oracle, c4):
    base, q, extra, adj, idx = c4
    sample = q[::10]
    for ef in (64, 128):
        r = idx.search(sample, 10, ef)
        o = oracle.search(base, adj, sample, 10, ef, block=32, metric="hamming", bloom_bits=1 << 14, group=2)
        same(r, o, f"config 4 queue {ef}")
    plain = idx.search(q, 10, 128)
    pers = idx.search(q, 10, 128, mode=1)
    np.testing.assert_array_equal(pers.ids, plain.ids)
    np.testing.assert_array_equal(pers.trace, plain.trace)


def test_config4_learned(song, c4):
    """Entry and stop networks trained on held-out codes: the stop that never fires is the plain
    search; the learned entry + stop expands fewer nodes than plain at the same queue, at a
    recall within 0.08 of it (Hamming distances tie often, so recall counts a tie either way)."""
    base, q, extra, adj, idx = c4
    learned.fit(idx, base, extra, k=10, ef=128, entry_epochs=5, stop_epochs=3, cost_epochs=5,
                stop_min_hops=8, stop_interval=2)
    gt, _, _ = idx.bruteforce(q, 10)
    plain = idx.search(q, 10, 128)
    never = idx.search(q, 10, 128, learned=2, stop_threshold=1e30, stop_min_hops=1)
    np.testing.assert_array_equal(never.ids, plain.ids)
    np.testing.assert_array_equal(never.trace, plain.trace)
    lr = idx.search(q, 10, 128, learned=3, stop_threshold=2.0, stop_min_hops=8, stop_interval=2, mode=1)
    assert lr.trace[:, 2].mean() < plain.trace[:, 2].mean()
    assert recall(lr.ids, gt) > recall(plain.ids, gt) - 0.08


# ---------------------------------------------------------------- inner product, mixed budgets

@pytest.fixture(scope="module")
def c5(song, datasets):
    cfg = datasets.config5(n=200_000, nq=1200)
    adj = build(song, cfg["base"], "ip")
    idx = song.Index(cfg["base"], adj, metric="ip")
    rng = np.random.default_rng(5)
    ks = rng.choice(cfg["ks"], size=cfg["queries"].shape[0]).astype(np.int32)
    efs = rng.choice(cfg["efs"], size=cfg["queries"].shape[0]).astype(np.int32)
    yield cfg["base"], cfg["queries"], adj, idx, ks, efs
    idx.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_config5_mixed_budgets_match_oracle(song, oracle, c5, mode):
    base, q, adj, idx, ks, efs = c5
    N = song._native
    nq = q.shape[0]
    dq = idx.queries_to_device(q)
    d_ks = N.DeviceBuffer(nq * 4).upload(ks)
    d_efs = N.DeviceBuffer(nq * 4).upload(efs)
    d_ids = N.DeviceBuffer(nq * 100 * 4)
    d_d = N.DeviceBuffer(nq * 100 * 4)
    d_tr = N.DeviceBuffer(nq * 48)
    idx.search_device_mixed(dq, nq, d_ks, d_efs, 100, 256, d_ids, d_d, d_tr, mode=mode)
    ids = d_ids.download(np.empty((nq, 100), np.int32))
    dd = d_d.download(np.empty((nq, 100), np.float32))
    tr = d_tr.download(np.empty((nq, 12), np.int32))
    for kk in (1, 10, 100):
        for ee in (64, 256):
            sel = np.where((ks == kk) & (efs == ee))[0]
            assert sel.size > 0
            r = idx.search(q[sel], kk, ee)
            np.testing.assert_array_equal(ids[sel, :kk], r.ids)
            np.testing.assert_array_equal(tr[sel], r.trace)
            assert (ids[sel, kk:] == -1).all()
            # 32 lanes per 200-float row
            sub = sel[:20]
            oi, od, ot = oracle.search(base, adj, q[sub], kk, ee, block=32, metric="ip", bloom_bits=1 << 14,
                                       group=32)
            np.testing.assert_array_equal(ids[sub, :kk], oi)
            np.testing.assert_array_equal(dd[sub, :kk].view(np.uint32), od.view(np.uint32))
            np.testing.assert_array_equal(tr[sub], ot)


def test_config5_serving(song, c5):
    """Poisson arrivals through the serving loop: every request equals the uniform search of its
    (k, queue) class."""
    base, q, adj, idx, ks, efs = c5
    sv = serving.Server(idx, max_batch=128, n_streams=4, k_max=100, ef_max=256)
    try:
        ids, d, lat, wall = sv.run_poisson(q, ks, efs, rate_qps=100000.0, seed=2)
        assert (lat > 0).all()
    finally:
        sv.close()
    for kk in (1, 10, 100):
        for ee in (64, 256):
            sel = np.where((ks == kk) & (efs == ee))[0]
            r = idx.search(q[sel], kk, ee)
            np.testing.assert_array_equal(ids[sel, :kk], r.ids)
            np.testing.assert_array_equal(d[sel, :kk].view(np.uint32), r.dists.view(np.uint32))
