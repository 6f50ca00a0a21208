"""The benchmark workload at full size (1,000,000 x 128 integer GMM, out-degree
32, 10,000 queries, the graph bench.py builds): properties that do not depend
on size, checked where the timed numbers come from.

* A sample of queries against the oracle, bit for bit, at the tuned plain
  queue (320) with its 2^15-bit visited filter (16 and 8 lanes per row), and
  at the reference's 2^14 bits: ids, distances and the reference's trace
  counters.
* The persistent kernel against the plain one on all 10,000 queries.
* The learned search's result cap with a stop that never fires, against the
  oracle's restatement of that option.
* The tuned learned search bench.py times (profiles/tuned_configs.json), stop
  network included, against the oracle's restatement on a sample.
* Recall@10 of both tuned methods against the GPU brute force: >= 0.90.
"""
import importlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

learned = importlib.import_module("2020_song_ex_amd.learned")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNED = json.load(open(os.path.join(ROOT, "profiles", "tuned_configs.json")))["1000000x128 nq=10000 k=10 recall>=0.9"]


@pytest.fixture(scope="module")
def full(song, datasets):
    cfg = datasets.config2(n_extra=25000)  # bench.py's draw: training queries first
    base, q = cfg["base"], cfg["queries"]
    adj = song.nsw_build(base, block=32, n_threads=16, batch_div=64, seq_prefix=1024, exact_order=False)
    idx = song.Index(base, adj)
    gt, _, _ = idx.bruteforce(q, 10)
    yield base, q, adj, idx, gt, cfg["extra"][:15000]
    idx.close()


def recall(ids, gt):
    return float(np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)]))


def test_sample_matches_oracle(full, oracle):
    base, q, adj, idx, gt, _ = full
    sample = q[::50]  # 200 queries
    for ef, bits, lanes in ((320, 1 << 15, 16), (320, 1 << 15, 8), (352, 1 << 14, 16)):
        r = idx.search(sample, 10, ef, bloom_bits=bits, lanes_per_row=lanes)
        # 16 (the default) or 8 lanes per 128-float row; integer data sums exactly in any order
        oi, od, ot = oracle.search(base, adj, sample, 10, ef, block=32, bloom_bits=bits, group=lanes)
        np.testing.assert_array_equal(r.ids, oi)
        np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32))
        np.testing.assert_array_equal(r.trace, ot)


def test_persistent_equals_plain(full):
    base, q, adj, idx, gt, _ = full
    a = idx.search(q, 10, 320, mode=0, waves_per_block=1, bloom_bits=1 << 15)
    b = idx.search(q, 10, 320, mode=1, bloom_bits=1 << 15)
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.trace, b.trace)
    assert recall(a.ids, gt) >= 0.90


def test_learned_search(full, oracle):
    base, q, adj, idx, gt, extra = full
    # the model bench.py trains (same queries, same settings; training is deterministic): the
    # stop network is trained at the interval it is consulted at, on searches from the reference seeds
    tuned = dict(TUNED["learned"])
    model = learned.fit(idx, base, extra, k=10, ef=384, stop_min_hops=16, stop_interval=tuned["stop_interval"],
                        max_evals=192, entry_epochs=20, stop_epochs=4, cost_threshold=3.0, trace_entry=False)
    # the tuned configuration bench.py times reaches the benchmark's recall
    ef = tuned.pop("ef")
    r = idx.search(q, 10, ef, **tuned)
    assert recall(r.ids, gt) >= 0.90
    assert (r.trace[:, 2] < 436).mean() > 0.5  # fewer expansions than the plain search's mean
    assert (r.trace[:, 11] == 2).mean() > 0.5  # most searches end by the stop network
    # ... and equals the oracle's restatement on a sample, stop network included
    sample = q[::100]
    oi, od, ot = oracle.search_learned(
        base, adj, sample, 10, ef, block=32, group=16, bloom_bits=tuned["bloom_bits"], stop_w=model.stop_w,
        dist_scale=model.dist_scale, threshold=tuned["stop_threshold"], min_hops=tuned["stop_min_hops"],
        interval=tuned["stop_interval"], confirm=tuned["stop_confirm"], max_hops=tuned["stop_max_hops"],
        result_cap=tuned["result_cap"], late_hops=tuned.get("stop_late_hops", 0),
        late_interval=tuned.get("stop_late_interval", 0))
    np.testing.assert_array_equal(r.ids[::100], oi)
    np.testing.assert_array_equal(r.dists[::100].view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(r.trace[::100], ot)
    # result cap with a stop that never fires: the oracle's capped search with that result set
    sample = q[::100]
    s = idx.search(sample, 10, 64, learned=2, stop_threshold=1e30, stop_min_hops=16, stop_interval=8,
                   stop_max_hops=150, result_cap=16, bloom_bits=1 << 15)
    oi, od, ot = oracle.search(base, adj, sample, 10, 64, block=32, bloom_bits=1 << 15, group=16, max_hops=150,
                               result_cap=16)
    np.testing.assert_array_equal(s.ids, oi)
    np.testing.assert_array_equal(s.dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(s.trace, ot)
