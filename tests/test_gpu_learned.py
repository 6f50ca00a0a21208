"""Learned entry / termination: HIP inference vs a plain FP32 MLP, HIP training,
the plain-kernel fallback, trace mode, and the effect on work and recall."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

learned = importlib.import_module("2020_song_ex_amd.learned")


@pytest.mark.parametrize("in_dim,out_dim", [(8, 1), (16, 1), (16, 64)])
def test_kernel_inference_matches_fp32_mlp(in_dim, out_dim):
    """The search kernel's fused MLP code against a float64-accumulated FP32 MLP (1e-5 relative)."""
    rng = np.random.default_rng(in_dim)
    w = learned.mlp_init(in_dim, out_dim, seed=3)
    w += rng.normal(0, 0.05, size=w.shape).astype(np.float32)
    X = rng.normal(size=(500, in_dim)).astype(np.float32)
    got = learned.mlp_forward(w, X, in_dim, out_dim)
    want = learned.mlp_forward_numpy(w, X, in_dim, out_dim)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_training_fits_a_simple_rule():
    """HIP training: logistic loss falls and the trained net separates a linear rule."""
    rng = np.random.default_rng(0)
    X = rng.normal(size=(20000, 8)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 3] > 0.2).astype(np.float32)
    w, loss = learned.mlp_train(X, y, 8, 1, epochs=5, lr=3e-3, seed=1)
    assert loss[-1] < loss[0] * 0.5, loss
    pred = learned.mlp_forward(w, X, 8, 1)[:, 0] > 0
    assert (pred == (y > 0)).mean() > 0.97
    # deterministic for a fixed seed
    w2, loss2 = learned.mlp_train(X, y, 8, 1, epochs=5, lr=3e-3, seed=1)
    np.testing.assert_array_equal(w, w2)


def test_training_regression():
    """Squared-error training (the cost network's loss) fits a smooth target."""
    rng = np.random.default_rng(4)
    X = rng.normal(size=(20000, 16)).astype(np.float32)
    y = (0.8 * X[:, 0] - 0.5 * X[:, 5] + 0.3 * np.abs(X[:, 2])).astype(np.float32)
    w, loss = learned.mlp_train(X, y, 16, 1, epochs=8, lr=3e-3, seed=3, loss=learned.LOSS_MSE)
    assert loss[-1] < loss[0] * 0.2, loss
    pred = learned.mlp_forward(w, X, 16, 1)[:, 0]
    assert np.corrcoef(pred, y)[0, 1] > 0.98


def test_training_softmax():
    rng = np.random.default_rng(1)
    centers = rng.normal(size=(64, 16)).astype(np.float32) * 3
    lab = rng.integers(0, 64, size=20000)
    X = (centers[lab] + rng.normal(size=(20000, 16))).astype(np.float32)
    w, loss = learned.mlp_train(X, lab.astype(np.float32), 16, 64, epochs=8, lr=3e-3, seed=2)
    assert loss[-1] < loss[0]
    acc = (learned.mlp_forward(w, X, 16, 64).argmax(1) == lab).mean()
    assert acc > 0.9, acc


@pytest.fixture(scope="module")
def setup(song, datasets):
    base, q = datasets.gmm_int(20000, 400, dim=64, n_centres=40, sigma=20.0, seed=5)
    adj = song.nsw_build(base, block=32, exact_order=False)
    idx = song.Index(base, adj)
    train = datasets.gmm_int(20000, 3000, dim=64, n_centres=40, sigma=20.0, seed=5)[1][200:]
    model = learned.fit(idx, base, train, k=10, ef=128, entry_epochs=10, stop_epochs=3, cost_epochs=10,
                        stop_min_hops=8, stop_interval=2)
    yield base, q, adj, idx, model
    idx.close()


def test_nan_model_falls_back_to_plain(song, setup):
    """A model whose outputs are NaN leaves the plain search's results untouched."""
    base, q, adj, idx, model = setup
    plain = idx.search(q, 10, 128)
    bad = learned.LearnedModel(model.proj_dim, model.mu, model.proj, model.proj_scale,
                               np.full_like(model.entry_w, np.nan), model.pivots,
                               np.full_like(model.stop_w, np.nan), model.dist_scale)
    bad.set_on(idx)
    try:
        r = idx.search(q, 10, 128, learned=3, stop_threshold=0.0, stop_min_hops=1)
        r2 = idx.search(q, 10, 128, learned=11, stop_threshold=0.0, stop_min_hops=1)
    finally:
        model.set_on(idx)
    for x in (r, r2):
        np.testing.assert_array_equal(x.ids, plain.ids)
        np.testing.assert_array_equal(x.dists, plain.dists)
        np.testing.assert_array_equal(x.trace, plain.trace)


def test_stop_never_fires_equals_plain(song, setup):
    base, q, adj, idx, model = setup
    plain = idx.search(q, 10, 128)
    r = idx.search(q, 10, 128, learned=2, stop_threshold=1e30, stop_min_hops=1)
    np.testing.assert_array_equal(r.ids, plain.ids)
    np.testing.assert_array_equal(r.trace, plain.trace)


@pytest.mark.parametrize("mode", [0, 1])
def test_stop_results_do_not_depend_on_workgroup_shape(song, setup, mode):
    """Every workgroup stages the stop network in its LDS; 1 and 8 waves per
    workgroup give the same results."""
    base, q, adj, idx, model = setup
    kw = dict(learned=2, stop_threshold=2.0, stop_min_hops=8, stop_interval=2, mode=mode)
    a = idx.search(q, 10, 128, waves_per_block=1, **kw)
    b = idx.search(q, 10, 128, waves_per_block=8, **kw)
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.dists, b.dists)
    np.testing.assert_array_equal(a.trace, b.trace)
    assert (a.trace[:, 11] == 2).any()


def test_learned_entry_matches_oracle_with_those_seeds(song, oracle, setup):
    """With learned entry only, the search equals the oracle run from the same seed list."""
    base, q, adj, idx, model = setup
    r = idx.search(q, 10, 128, learned=1)
    scores = learned.mlp_forward(model.entry_w, model.project(q, 0).astype(np.float32), 16, 64)
    order = np.argsort(-scores, axis=1, kind="stable")[:, :4]
    seeds = model.pivots[order]
    oi, od, ot = oracle.search(base, adj, q, 10, 128, block=32, bloom_bits=1 << 14, group=4, entries=seeds)
    agree = (r.ids == oi).all(1).mean()
    # the projection is computed in a different summation order than numpy's, so
    # a near-tie between two pivot scores can pick another seed list
    assert agree > 0.97, agree


@pytest.mark.parametrize("mode", [0, 1])
def test_cost_order_keeps_results(song, setup, mode):
    """Running the queue in predicted-cost order changes when a query runs, not its result."""
    base, q, adj, idx, model = setup
    assert model.cost_w is not None
    for base_flags in (1, 3):
        a = idx.search(q, 10, 128, learned=base_flags, stop_threshold=2.0, stop_min_hops=8, stop_interval=2,
                       mode=mode)
        b = idx.search(q, 10, 128, learned=base_flags | 4, stop_threshold=2.0, stop_min_hops=8,
                       stop_interval=2, mode=mode)
        np.testing.assert_array_equal(a.ids, b.ids)
        np.testing.assert_array_equal(a.dists, b.dists)
        np.testing.assert_array_equal(a.trace, b.trace)


def test_cost_network_ranks_hops(song, setup):
    """The cost network's prediction correlates with the hop counts it was trained on."""
    base, q, adj, idx, model = setup
    r = idx.search(q, 10, 128, learned=3, stop_threshold=2.5, stop_min_hops=8, stop_interval=2)
    pred = learned.mlp_forward(model.cost_w, model.project(q, 0).astype(np.float32), 16, 1)[:, 0]
    assert np.corrcoef(pred, np.log2(r.trace[:, 2]))[0, 1] > 0.3


def test_nan_cost_keeps_results(song, setup):
    base, q, adj, idx, model = setup
    ref = idx.search(q, 10, 128, learned=1)
    bad = learned.LearnedModel(model.proj_dim, model.mu, model.proj, model.proj_scale, model.entry_w,
                               model.pivots, model.stop_w, model.dist_scale, np.full_like(model.cost_w, np.nan))
    bad.set_on(idx)
    try:
        r = idx.search(q, 10, 128, learned=5, mode=1)
    finally:
        model.set_on(idx)
    np.testing.assert_array_equal(r.ids, ref.ids)
    np.testing.assert_array_equal(r.trace, ref.trace)


def test_order_needs_cost_network(song, setup):
    base, q, adj, idx, model = setup
    plain = learned.LearnedModel(model.proj_dim, model.mu, model.proj, model.proj_scale, model.entry_w,
                                 model.pivots, model.stop_w, model.dist_scale)
    plain.set_on(idx)
    try:
        with pytest.raises(song.SongError, match="cost network"):
            idx.search(q[:4], 10, 64, learned=4)
    finally:
        model.set_on(idx)


@pytest.mark.parametrize("cap", [1, 20, 60])
def test_hop_cap_matches_oracle(song, oracle, setup, cap):
    """The learned termination's hop cap, with the stop network never firing,
    equals the oracle's capped search (early_stopped = 3 where it cut)."""
    base, q, adj, idx, model = setup
    r = idx.search(q, 10, 128, learned=2, stop_threshold=1e30, stop_min_hops=1, stop_max_hops=cap)
    oi, od, ot = oracle.search(base, adj, q, 10, 128, block=32, bloom_bits=1 << 14, group=4, max_hops=cap)
    np.testing.assert_array_equal(r.ids, oi)
    np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(r.trace, ot)
    assert (r.trace[:, 11] == 3).any()


@pytest.mark.parametrize("rcap", [10, 32])
def test_result_cap_matches_oracle(song, oracle, setup, rcap):
    """A result cap below the queue size, with the stop network never firing:
    the search runs to the hop cap (or an empty frontier) without the reference's
    early stop and keeps a result set of rcap entries -- the oracle's search with
    the same result capacity, bit for bit."""
    base, q, adj, idx, model = setup
    r = idx.search(q, 10, 128, learned=2, stop_threshold=1e30, stop_min_hops=1, stop_max_hops=40, result_cap=rcap)
    oi, od, ot = oracle.search(base, adj, q, 10, 128, block=32, bloom_bits=1 << 14, group=4, max_hops=40,
                               result_cap=rcap)
    np.testing.assert_array_equal(r.ids, oi)
    np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(r.trace, ot)
    assert (r.trace[:, 9] <= rcap).all()


def test_result_cap_options(song, setup):
    """result_cap = queue size is the plain rule; below it the learned stop and a hop
    cap are required, and it must hold k."""
    base, q, adj, idx, model = setup
    kw = dict(learned=2, stop_threshold=2.0, stop_min_hops=8, stop_interval=2, stop_max_hops=200)
    a = idx.search(q, 10, 128, **kw)
    b = idx.search(q, 10, 128, result_cap=128, **kw)
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.trace, b.trace)
    c = idx.search(q, 10, 128, result_cap=64, **kw)
    assert c.trace[:, 7].sum() < a.trace[:, 7].sum()  # fewer result-set updates
    for bad in (dict(result_cap=64), dict(result_cap=64, learned=2, stop_threshold=2.0),
                dict(result_cap=5, **kw), dict(result_cap=-1, **kw)):
        with pytest.raises(song.SongError, match="result_cap"):
            idx.search(q[:4], 10, 128, **bad)


@pytest.mark.parametrize("mode", [0, 1])
def test_result_cap_with_mixed_budgets(song, setup, mode):
    """Per-query k and queue size in one launch with a result cap (the serving loop's path):
    each query equals its uniform search -- queues above the cap without the reference's early
    stop, queues at the cap with it."""
    base, q, adj, idx, model = setup
    N = song._native
    nq = q.shape[0]
    rng = np.random.default_rng(1)
    ks = rng.choice([1, 10], size=nq).astype(np.int32)
    efs = rng.choice([16, 64, 128], size=nq).astype(np.int32)
    kw = dict(learned=2, stop_threshold=2.0, stop_min_hops=8, stop_interval=2, stop_max_hops=200, result_cap=16)
    dq = idx.queries_to_device(q)
    d_ks = N.DeviceBuffer(nq * 4).upload(ks)
    d_efs = N.DeviceBuffer(nq * 4).upload(efs)
    d_ids = N.DeviceBuffer(nq * 10 * 4)
    d_tr = N.DeviceBuffer(nq * 48)
    idx.search_device_mixed(dq, nq, d_ks, d_efs, 10, 128, d_ids, None, d_tr, mode=mode, **kw)
    ids = d_ids.download(np.empty((nq, 10), np.int32))
    tr = d_tr.download(np.empty((nq, 12), np.int32))
    for kk in (1, 10):
        for ee in (16, 64, 128):
            sel = np.where((ks == kk) & (efs == ee))[0]
            if sel.size == 0:
                continue
            r = idx.search(q[sel], kk, ee, **kw)
            np.testing.assert_array_equal(ids[sel, :kk], r.ids)
            np.testing.assert_array_equal(tr[sel], r.trace)


@pytest.mark.parametrize("mode", [0, 1])
def test_learned_server_falls_back_to_plain(song, setup, mode):
    """A learned server with k_max above what the learned search serves (64): requests with k <= 64
    get the learned search, those with k > 64 the plain one, each equal to its uniform search."""
    serving = importlib.import_module("2020_song_ex_amd.serving")
    base, q, adj, idx, model = setup
    kw = dict(learned=2, stop_threshold=2.0, stop_min_hops=8, stop_interval=2)
    rng = np.random.default_rng(7)
    ks = rng.choice([1, 10, 100], size=q.shape[0]).astype(np.int32)
    efs = rng.choice([64, 128], size=q.shape[0]).astype(np.int32)
    sv = serving.Server(idx, max_batch=64, n_streams=4, k_max=100, ef_max=128, mode=mode, **kw)
    try:
        ids, d, lat, wall = sv.run_poisson(q, ks, efs, rate_qps=100000.0, seed=4)
    finally:
        sv.close()
    for kk in (1, 10, 100):
        for ee in (64, 128):
            sel = np.where((ks == kk) & (efs == ee))[0]
            assert sel.size > 0
            r = idx.search(q[sel], kk, ee, **(kw if kk <= 64 else {}))
            np.testing.assert_array_equal(ids[sel, :kk], r.ids)
            np.testing.assert_array_equal(d[sel, :kk].view(np.uint32), r.dists.view(np.uint32))
    with pytest.raises(song.SongError, match="result_cap"):
        serving.Server(idx, k_max=100, result_cap=16, stop_max_hops=100, **kw)  # result cap below the learned k


def test_confirmed_stop_runs_longer(song, setup):
    """Needing two consecutive stop votes never stops earlier than one vote,
    and one confirmation is the default."""
    base, q, adj, idx, model = setup
    kw = dict(learned=2, stop_threshold=1.0, stop_min_hops=8, stop_interval=2)
    one = idx.search(q, 10, 128, **kw)
    dflt = idx.search(q, 10, 128, stop_confirm=1, **kw)
    two = idx.search(q, 10, 128, stop_confirm=2, **kw)
    np.testing.assert_array_equal(one.trace, dflt.trace)
    assert (two.trace[:, 2] >= one.trace[:, 2]).all()
    assert two.trace[:, 2].mean() > one.trace[:, 2].mean()


def test_augmented_entry_matches_oracle_with_those_seeds(song, oracle, setup):
    """Learned bit 3: each seed row is the 4 reference seeds followed by the 4
    best pivots; the search equals the oracle run from that row."""
    base, q, adj, idx, model = setup
    r = idx.search(q, 10, 128, learned=9)
    scores = learned.mlp_forward(model.entry_w, model.project(q, 0).astype(np.float32), 16, 64)
    order = np.argsort(-scores, axis=1, kind="stable")[:, :4]
    ref = np.array([(p * base.shape[0]) // 4 for p in range(4)], np.int32)
    seeds = np.concatenate([np.tile(ref, (q.shape[0], 1)), model.pivots[order]], axis=1)
    oi, od, ot = oracle.search(base, adj, q, 10, 128, block=32, bloom_bits=1 << 14, group=4, entries=seeds)
    assert (r.ids == oi).all(1).mean() > 0.97
    with pytest.raises(song.SongError, match="bit 3"):
        idx.search(q[:2], 10, 128, learned=8)


def test_trace_labels(song, setup):
    base, q, adj, idx, model = setup
    N = song._native
    L = learned._bind()
    import ctypes
    nq, maxe = q.shape[0], 64
    dq = idx.queries_to_device(q)
    d_tr, d_c, d_f = N.DeviceBuffer(nq * maxe * 9 * 4), N.DeviceBuffer(nq * 4), N.DeviceBuffer(nq * 4)
    o = N.make_opts()
    o.stop_min_hops, o.stop_interval = 4, 2
    N.check(L.song_learn_trace(idx.handle, dq.ptr, nq, 10, 128, ctypes.byref(o), maxe, d_tr.ptr, d_c.ptr, d_f.ptr),
            "song_learn_trace")
    cnt = d_c.download(np.empty(nq, np.int32))
    fin = d_f.download(np.empty(nq, np.float32))
    tr = d_tr.download(np.empty((nq, maxe, 9), np.float32))
    plain = idx.search(q, 10, 128)
    # entry off: the trace run is the plain search, whose 10th distance is final
    np.testing.assert_array_equal(fin, plain.dists[:, 9])
    assert (cnt >= 1).all() and (cnt <= maxe).all()
    for i in range(nq):
        rows = tr[i, :cnt[i]]
        dk = rows[:, 8]
        assert (np.diff(dk) <= 0).all()  # the k-th best never gets worse
        assert dk[-1] >= fin[i]
        assert (np.abs(rows[:, :8]) <= 8).all()  # features are clamped to [-8, 8]


def test_learned_search_saves_work(song, setup):
    """Learned entry + stop: fewer expansions than plain at the same queue size, recall close."""
    base, q, adj, idx, model = setup
    gt, _, _ = idx.bruteforce(q, 10)

    def rec(ids):
        return np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)])

    plain = idx.search(q, 10, 128)
    lr = idx.search(q, 10, 128, learned=3, stop_threshold=2.0, stop_min_hops=8, stop_interval=2)
    assert lr.trace[:, 2].mean() < plain.trace[:, 2].mean()
    assert rec(lr.ids) > rec(plain.ids) - 0.05, (rec(lr.ids), rec(plain.ids))
    assert (lr.trace[:, 11] == 2).any()


@pytest.mark.parametrize("metric", ["hamming", "ip"])
def test_learned_search_other_metrics(song, datasets, metric):
    """The networks train and run on Hamming codes (256-bit, +-1 projection) and on
    inner-product data; the stop saves work at a small recall cost and never
    firing returns the plain search."""
    vec, qv = datasets.vmf_mix(12000, 1600, dim=64, n_centres=40, kappa=20.0, seed=4)
    if metric == "hamming":
        base, planes = datasets.hyperplane_codes(vec, seed=5)
        qs, _ = datasets.hyperplane_codes(qv, planes=planes)
    else:
        base, qs = vec, qv
    adj = song.nsw_build(base, metric=metric, block=32, exact_order=False)
    idx = song.Index(base, adj, metric=metric)
    try:
        q, train = qs[:300], qs[300:]
        model = learned.fit(idx, base, train, k=10, ef=128, entry_epochs=5, stop_epochs=3, cost_epochs=5,
                            stop_min_hops=8, stop_interval=2)
        gt, _, _ = idx.bruteforce(q, 10)

        def rec(ids):
            return np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)])

        plain = idx.search(q, 10, 128)
        never = idx.search(q, 10, 128, learned=2, stop_threshold=1e30, stop_min_hops=1)
        np.testing.assert_array_equal(never.ids, plain.ids)
        np.testing.assert_array_equal(never.trace, plain.trace)
        lr = idx.search(q, 10, 128, learned=6, stop_threshold=2.0, stop_min_hops=8, stop_interval=2, mode=1)
        assert lr.trace[:, 2].mean() < plain.trace[:, 2].mean()
        assert rec(lr.ids) > rec(plain.ids) - 0.08, (rec(lr.ids), rec(plain.ids))
    finally:
        idx.close()


STOP_VARIANTS = {
    "one_vote": dict(stop_min_hops=8, stop_interval=2),
    "two_votes_hop_cap_result_cap": dict(stop_min_hops=8, stop_interval=2, stop_confirm=2, stop_max_hops=60,
                                         result_cap=16),
    "late_first_consult": dict(stop_min_hops=21, stop_interval=5, stop_max_hops=200),
    "late_interval": dict(stop_min_hops=8, stop_interval=2, stop_late_hops=30, stop_late_interval=9,
                          stop_max_hops=200, result_cap=16),
}


@pytest.mark.parametrize("variant", list(STOP_VARIANTS))
@pytest.mark.parametrize("mode", [0, 1])
def test_stop_network_that_fires_matches_oracle(song, oracle, setup, variant, mode):
    """The learned stop as it ends searches -- features, consult schedule, network, votes --
    against the oracle's restatement: ids, distances and all 12 counters bit for bit, at
    thresholds where the network stops some queries and not others."""
    base, q, adj, idx, model = setup
    kw = STOP_VARIANTS[variant]
    fired_some = False
    for thr in (-1.0, 0.5, 2.0, 4.0):
        r = idx.search(q, 10, 128, learned=2, stop_threshold=thr, mode=mode, **kw)
        oi, od, ot = oracle.search_learned(
            base, adj, q, 10, 128, block=32, group=8, bloom_bits=1 << 14, stop_w=model.stop_w,
            dist_scale=model.dist_scale, threshold=thr, min_hops=kw["stop_min_hops"], interval=kw["stop_interval"],
            confirm=kw.get("stop_confirm", 1), max_hops=kw.get("stop_max_hops", 0), result_cap=kw.get("result_cap", 0),
            late_hops=kw.get("stop_late_hops", 0), late_interval=kw.get("stop_late_interval", 0))
        np.testing.assert_array_equal(r.ids, oi)
        np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32))
        np.testing.assert_array_equal(r.trace, ot)
        stopped = (r.trace[:, 11] == 2).mean()
        fired_some = fired_some or 0.05 < stopped < 0.95
    assert fired_some, "no threshold stopped a part of the queries"


@pytest.mark.parametrize("hop", [6, 20])
@pytest.mark.parametrize("mode", [0, 1])
def test_late_entry_matches_oracle(song, oracle, setup, hop, mode):
    """entry_late_hops: the search starts at the reference seeds and the entry network's 4 best
    pivots join at that hop; equal to the oracle run with those pivots (the planning pass's
    projection sums in another order than numpy's, so a near-tie of two pivot scores can pick
    another list for a few queries)."""
    base, q, adj, idx, model = setup
    kw = dict(stop_threshold=2.0, stop_min_hops=8, stop_interval=4, stop_max_hops=200, result_cap=16)
    r = idx.search(q, 10, 48, learned=3, entry_late_hops=hop, mode=mode, **kw)
    scores = learned.mlp_forward(model.entry_w, model.project(q, 0).astype(np.float32), 16, 64)
    seeds = model.pivots[np.argsort(-scores, axis=1, kind="stable")[:, :4]]
    oi, od, ot = oracle.search_learned(
        base, adj, q, 10, 48, block=32, group=8, bloom_bits=1 << 14, stop_w=model.stop_w,
        dist_scale=model.dist_scale, threshold=2.0, min_hops=8, interval=4, max_hops=200, result_cap=16,
        entries=seeds, entry_late_hops=hop)
    same = (r.ids == oi).all(1) & (r.trace == ot).all(1) & (r.dists.view(np.uint32) == od.view(np.uint32)).all(1)
    assert same.mean() > 0.97, same.mean()
    # the join changed some searches: without it they run from the reference seeds alone
    plain_seeds = idx.search(q, 10, 48, learned=2, mode=mode, **kw)
    assert (plain_seeds.ids != r.ids).any()
    assert (r.trace[:, 1] >= 0).all()


def test_late_entry_needs_the_entry_network_and_the_stop(song, setup):
    base, q, adj, idx, model = setup
    for bits in (2, 1, 11):
        with pytest.raises(song.SongError, match="entry_late_hops"):
            idx.search(q[:4], 10, 48, learned=bits, entry_late_hops=10, stop_max_hops=100)
    with pytest.raises(song.SongError, match="entry_late_hops"):
        idx.search(q[:4], 10, 128, learned=3, entry_late_hops=10, stop_max_hops=100)  # queue over 64
    # trace mode records searches without it
    N = song._native
    L = learned._bind()
    import ctypes
    dq = idx.queries_to_device(q[:4])
    bufs = N.DeviceBuffer(4 * 8 * 9 * 4), N.DeviceBuffer(16), N.DeviceBuffer(16)
    o = N.make_opts(learned=1, entry_late_hops=10)
    with pytest.raises(song.SongError, match="entry_late_hops"):
        N.check(L.song_learn_trace(idx.handle, dq.ptr, 4, 10, 48, ctypes.byref(o), 8, *(b.ptr for b in bufs)),
                "song_learn_trace")


@pytest.mark.parametrize("late", [(0, 0), (23, 7)])
def test_trace_features_match_oracle(song, oracle, setup, late):
    """The features the stop network sees (trace mode records them) equal the oracle's, bit for bit,
    with and without a late consult interval."""
    base, q, adj, idx, model = setup
    N = song._native
    L = learned._bind()
    import ctypes
    nq, maxe = q.shape[0], 96
    dq = idx.queries_to_device(q)
    d_tr, d_c, d_f = N.DeviceBuffer(nq * maxe * 9 * 4), N.DeviceBuffer(nq * 4), N.DeviceBuffer(nq * 4)
    o = N.make_opts()
    o.stop_min_hops, o.stop_interval = 5, 3
    o.stop_late_hops, o.stop_late_interval = late
    N.check(L.song_learn_trace(idx.handle, dq.ptr, nq, 10, 128, ctypes.byref(o), maxe, d_tr.ptr, d_c.ptr, d_f.ptr),
            "song_learn_trace")
    cnt = d_c.download(np.empty(nq, np.int32))
    fin = d_f.download(np.empty(nq, np.float32))
    tr = d_tr.download(np.empty((nq, maxe, 9), np.float32))
    rows, ocnt, ofin = oracle.learned_trace(base, adj, q, 10, 128, block=32, group=8, bloom_bits=1 << 14,
                                            dist_scale=model.dist_scale, min_hops=5, interval=3, late_hops=late[0],
                                            late_interval=late[1], max_evals=maxe)
    np.testing.assert_array_equal(cnt, ocnt)
    np.testing.assert_array_equal(fin.view(np.uint32), ofin.view(np.uint32))
    for i in range(nq):
        np.testing.assert_array_equal(tr[i, :cnt[i]].view(np.uint32), rows[i, :cnt[i]].view(np.uint32))
