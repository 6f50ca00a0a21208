"""GPU search vs the oracle, through the C ABI (libsong.so).

Bar: ids, distances and all 12 QueryTrace counters identical to the oracle's
re-execution of the reference GPU algorithm (kernalGraph.h:184-299) on the
same graph, seeds, queue size and Bloom size.  The oracle sums float
distances in the kernel's order, so the comparison is exact (tolerance 0)
for float data too; the 1e-5-relative allowance of the spec is not needed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def lanes_for(dim):
    """The lane-group size the library picks for a dim: the smallest power of
    two (2..64) whose lanes hold the row's 4-float chunks two per lane."""
    d4 = (dim + 3) // 4
    g = 2
    while g < 64 and g * 2 < d4:
        g *= 2
    return g


def run_both(song, oracle, base, adj, queries, k, ef, metric="l2", bloom_bits=1 << 14, n_seeds=4,
             lanes=0, waves_per_block=0, batch=0):
    idx = song.Index(base, adj, metric=metric)
    try:
        r = idx.search(queries, k, ef, bloom_bits=bloom_bits, n_seeds=n_seeds, lanes_per_row=lanes,
                       waves_per_block=waves_per_block, batch=batch)
    finally:
        idx.close()
    g = lanes or (lanes_for(base.shape[1]) if metric != "hamming" else 2)
    oi, od, ot = oracle.search(base, adj, queries, k, ef, block=adj.shape[1], metric=metric,
                               bloom_bits=bloom_bits, n_seeds=n_seeds, group=g)
    return r, (oi, od, ot)


def assert_same(r, o, label=""):
    oi, od, ot = o
    bad = np.where((r.ids != oi).any(1))[0]
    assert bad.size == 0, f"{label}: ids differ for queries {bad[:10]}: gpu {r.ids[bad[0]]} oracle {oi[bad[0]]}"
    np.testing.assert_array_equal(r.dists.view(np.uint32), od.view(np.uint32), err_msg=f"{label}: dists")
    badt = np.where((r.trace != ot).any(1))[0]
    assert badt.size == 0, f"{label}: trace differs for {badt[:10]}: gpu {r.trace[badt[0]]} oracle {ot[badt[0]]}"


@pytest.fixture(scope="module")
def c1(song, datasets):
    cfg = datasets.config1()
    adj = song.nsw_build(cfg["base"], block=16, exact_order=False)
    return cfg, adj


def test_config1_exact(song, oracle, c1):
    """Config 1: 10k x 128 integer GMM, out-degree 16, k=10, queue 64, Bloom 2^15."""
    cfg, adj = c1
    r, o = run_both(song, oracle, cfg["base"], adj, cfg["queries"], 10, 64, bloom_bits=1 << 15)
    assert_same(r, o, "config1")
    assert (r.trace[:, 0] == 4).all()


@pytest.mark.parametrize("k,ef", [(1, 1), (10, 10), (10, 32), (10, 50), (10, 100), (10, 150), (10, 200),
                                  (100, 100), (100, 256), (10, 257), (10, 300), (10, 320), (10, 321), (100, 384), (10, 448),
                                  (10, 512), (64, 64), (65, 65)])
def test_queue_sizes(song, oracle, c1, k, ef):
    cfg, adj = c1
    r, o = run_both(song, oracle, cfg["base"], adj, cfg["queries"][:60], k, ef)
    assert_same(r, o, f"k={k} ef={ef}")


@pytest.mark.parametrize("bits", [512, 4096, 1 << 14, 1 << 16])
def test_bloom_sizes(song, oracle, c1, bits):
    """Small filters force false positives and in-row collisions."""
    cfg, adj = c1
    r, o = run_both(song, oracle, cfg["base"], adj, cfg["queries"][:60], 10, 64, bloom_bits=bits)
    assert_same(r, o, f"bloom={bits}")


@pytest.mark.parametrize("bits", [512, 1 << 15])
def test_visit_table_off(song, oracle, c1, bits):
    """Every other test searches with the index's visit table (the filter positions hashed once at
    index creation); without it the search hashes each neighbour id itself: the same results."""
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    try:
        on = idx.search(cfg["queries"], 10, 64, bloom_bits=bits)
        idx.set_visit_table(False)
        off = idx.search(cfg["queries"], 10, 64, bloom_bits=bits)
        idx.set_visit_table(True)
        again = idx.search(cfg["queries"], 10, 64, bloom_bits=bits)
    finally:
        idx.close()
    o = oracle.search(cfg["base"], adj, cfg["queries"], 10, 64, block=16, bloom_bits=bits,
                      group=lanes_for(cfg["base"].shape[1]))
    for r, label in ((on, "table"), (off, "hashed"), (again, "table rebuilt")):
        assert_same(r, o, f"{label}, bloom={bits}")


@pytest.mark.parametrize("dim", [128, 100])
def test_wide_offsets(song, oracle, datasets, dim):
    """Indexes above 2 GiB address rows with 64-bit offsets; forced on a small index it gives the
    oracle's results too, for rows the lane groups cover exactly (128) and with a ragged end (100)."""
    base, q = datasets.gmm_int(3000, 40, dim=dim, n_centres=20, sigma=20.0, seed=3)
    adj = song.nsw_build(base, block=16, exact_order=False)
    idx = song.Index(base, adj)
    try:
        idx.set_wide_offsets(True)
        r = idx.search(q, 10, 64, bloom_bits=1 << 15)
    finally:
        idx.close()
    o = oracle.search(base, adj, q, 10, 64, block=16, bloom_bits=1 << 15, group=lanes_for(dim))
    assert_same(r, o, f"wide offsets, dim {dim}")


@pytest.mark.parametrize("lanes,dim", [(2, 30), (4, 30), (8, 30), (16, 30), (32, 30), (64, 30),
                                       (8, 100), (16, 100), (64, 100)])
def test_lane_groups(song, oracle, datasets, lanes, dim):
    """Float data: every lane-group size gives the oracle's distances bit for bit
    (a group of G lanes holds 4 chunks each, so G * 16 >= dim or G = 64)."""
    base, q = datasets.beta_mix(3000, 40, dim=dim, n_centres=20, seed=3)
    adj = song.nsw_build(base, block=32, exact_order=False)
    r, o = run_both(song, oracle, base, adj, q, 10, 48, lanes=lanes)
    assert_same(r, o, f"lanes={lanes}")


@pytest.mark.parametrize("dim", [1, 3, 5, 31, 200, 960, 1500])
def test_dims(song, oracle, dim):
    rng = np.random.default_rng(dim)
    base = rng.random((1500, dim), dtype=np.float32)
    q = rng.random((30, dim), dtype=np.float32)
    adj = song.nsw_build(base, block=32, exact_order=False)
    r, o = run_both(song, oracle, base, adj, q, 10, 40)
    assert_same(r, o, f"dim={dim}")


def test_beta_mix_960(song, oracle, datasets):
    """960-d Beta-perturbed mixture (float data, config-3 shape): exact through the
    fixed summation order, queue 50 as the reference's default."""
    base, q = datasets.beta_mix(3000, 30, dim=960, n_centres=30, seed=3)
    adj = song.nsw_build(base, block=32, exact_order=False)
    r, o = run_both(song, oracle, base, adj, q, 10, 50)
    assert_same(r, o, "beta960")


def test_inner_product(song, oracle, datasets):
    base, q = datasets.vmf_mix(4000, 50, dim=200, n_centres=50, kappa=30.0, seed=1)
    adj = song.nsw_build(base, metric="ip", block=32, exact_order=False)
    r, o = run_both(song, oracle, base, adj, q, 10, 64, metric="ip")
    assert_same(r, o, "ip")


def test_hamming(song, oracle, datasets):
    """Hamming distances tie everywhere, so this checks the positional tie rules."""
    vec, qv = datasets.vmf_mix(4000, 80, dim=64, n_centres=40, kappa=20.0, seed=2)
    codes, planes = datasets.hyperplane_codes(vec, seed=5)
    qcodes, _ = datasets.hyperplane_codes(qv, planes=planes)
    adj = song.nsw_build(codes, metric="hamming", block=32)
    for k, ef in [(10, 10), (10, 64), (10, 200)]:
        r, o = run_both(song, oracle, codes, adj, qcodes, k, ef, metric="hamming")
        assert_same(r, o, f"hamming k={k} ef={ef}")


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_tiny_graphs(song, oracle, n):
    """n < 4: seeds repeat and are skipped by the Bloom filter; fewer than k results -> -1."""
    rng = np.random.default_rng(n)
    base = rng.integers(0, 9, size=(n, 8)).astype(np.float32)
    adj = song.nsw_build(base, block=8)
    q = rng.integers(0, 9, size=(5, 8)).astype(np.float32)
    r, o = run_both(song, oracle, base, adj, q, 10, 16)
    assert_same(r, o, f"n={n}")
    assert (r.ids[:, n:] == -1).all()


def test_isolated_nodes_and_seeds(song, oracle, c1):
    """Rows emptied to -1 and 1..64 seeds."""
    cfg, adj = c1
    adj2 = adj.copy()
    adj2[::7] = -1
    for seeds in (1, 4, 17, 64):
        r, o = run_both(song, oracle, cfg["base"], adj2, cfg["queries"][:40], 10, 32, n_seeds=seeds)
        assert_same(r, o, f"seeds={seeds}")


def test_batches_and_blocks(song, oracle, c1):
    """Results do not depend on batch size or queries per workgroup."""
    cfg, adj = c1
    for wpb, batch in [(1, 0), (2, 7), (4, 33), (4, 512), (8, 0), (8, 9)]:
        r, o = run_both(song, oracle, cfg["base"], adj, cfg["queries"], 10, 64,
                        waves_per_block=wpb, batch=batch)
        assert_same(r, o, f"wpb={wpb} batch={batch}")


def test_duplicate_ids_in_row(song, oracle, c1):
    """A row naming the same node twice: the second is a Bloom hit (in-row order)."""
    cfg, adj = c1
    adj2 = adj.copy()
    adj2[:, 8:] = adj2[:, :8]
    r, o = run_both(song, oracle, cfg["base"], adj2, cfg["queries"][:50], 10, 64)
    assert_same(r, o, "dup rows")


@pytest.mark.parametrize("k,ef", [(10, 64), (10, 300), (64, 200), (100, 100)])
def test_tied_distances(song, oracle, datasets, k, ef):
    """Every vector three times over: each distance repeats, so the final
    trim and drain resolve ties by the reference's positional rules."""
    base, q = datasets.gmm_int(1500, 40, dim=24, n_centres=20, sigma=12.0, seed=11)
    base = np.repeat(base, 3, axis=0)
    adj = song.nsw_build(base, block=16, exact_order=False)
    r, o = run_both(song, oracle, base, adj, q, k, ef)
    assert_same(r, o, f"ties k={k} ef={ef}")
    # the test is only meaningful if the k best really do tie
    assert (np.diff(r.dists[:, :min(k, 3)], axis=1) == 0).any()


def test_lane_group_too_small_is_an_error(song, c1):
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    with pytest.raises(song.SongError, match="lanes_per_row"):
        idx.search(cfg["queries"][:4], 10, 32, lanes_per_row=2)
    idx.close()


def test_persistent_equals_plain(song, oracle, c1):
    """The persistent work-queue kernel returns the plain kernel's results for every grid size."""
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    plain = idx.search(cfg["queries"], 10, 64)
    for wpc in (1, 3, 8, 32):
        for wpb in (1, 4, 8):
            r = idx.search(cfg["queries"], 10, 64, mode=1, waves_per_cu=wpc, waves_per_block=wpb)
            np.testing.assert_array_equal(r.ids, plain.ids)
            np.testing.assert_array_equal(r.dists, plain.dists)
            np.testing.assert_array_equal(r.trace, plain.trace)
    idx.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_mixed_budgets_match_uniform_runs(song, c1, mode):
    """Per-query k and queue size in one launch == separate uniform searches."""
    cfg, adj = c1
    q = cfg["queries"]
    nq = q.shape[0]
    rng = np.random.default_rng(0)
    ks = rng.choice([1, 10, 100], size=nq).astype(np.int32)
    efs = rng.choice([16, 64, 256], size=nq).astype(np.int32)
    idx = song.Index(cfg["base"], adj)
    N = song._native
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
        for ee in (16, 64, 256):
            sel = np.where((ks == kk) & (efs == ee))[0]
            if sel.size == 0:
                continue
            r = idx.search(q[sel], kk, ee)
            np.testing.assert_array_equal(ids[sel, :kk], r.ids)
            np.testing.assert_array_equal(dd[sel, :kk], r.dists)
            assert (ids[sel, kk:] == -1).all()
            np.testing.assert_array_equal(tr[sel], r.trace)
    idx.close()


def test_empty_query_set(song, c1):
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    r = idx.search(cfg["queries"][:0], 10, 64)
    assert r.ids.shape == (0, 10)
    idx.close()


def test_device_resident_search_matches_host(song, c1):
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    q = cfg["queries"]
    host = idx.search(q, 10, 64)
    dq = idx.queries_to_device(q)
    N = song._native
    d_ids = N.DeviceBuffer(q.shape[0] * 10 * 4)
    d_tr = N.DeviceBuffer(q.shape[0] * 48)
    tm = idx.search_device(dq, q.shape[0], 10, 64, d_ids, None, d_tr)
    ids = d_ids.download(np.empty((q.shape[0], 10), np.int32))
    tr = d_tr.download(np.empty((q.shape[0], 12), np.int32))
    np.testing.assert_array_equal(ids, host.ids)
    np.testing.assert_array_equal(tr, host.trace)
    assert tm["kernel_ms"] > 0
    idx.close()


def test_recall_against_bruteforce(song, oracle, c1):
    """Config 1 (out-degree 16): recall@10 rises with the queue; the values are the
    oracle's on the same graph (0.631 at 64, 0.967 at 512 on the CPU)."""
    cfg, adj = c1
    idx = song.Index(cfg["base"], adj)
    gt, _, _ = idx.bruteforce(cfg["queries"], 10)
    recs = []
    for ef in (64, 128, 256, 512):
        r = idx.search(cfg["queries"], 10, ef, bloom_bits=1 << 15)
        recs.append(np.mean([len(set(a) & set(b)) / 10 for a, b in zip(r.ids, gt)]))
    idx.close()
    assert recs == sorted(recs), recs
    assert recs[-1] >= 0.95, recs


@pytest.mark.parametrize("metric,k", [("l2", 10), ("l2", 100), ("ip", 10), ("hamming", 10), ("l2", 128)])
def test_gpu_bruteforce(song, oracle, datasets, metric, k):
    """Ground truth on the device equals the oracle's brute force (ties -> smaller id)."""
    if metric == "l2":
        base, q = datasets.gmm_int(20000, 70, dim=128, n_centres=30, sigma=30.0, seed=k)
    elif metric == "ip":
        base, q = datasets.vmf_mix(8000, 40, dim=200, n_centres=20, kappa=20.0, seed=1)
    else:
        v, qv = datasets.vmf_mix(8000, 40, dim=64, n_centres=20, kappa=10.0, seed=2)
        base, planes = datasets.hyperplane_codes(v, seed=3)
        q, _ = datasets.hyperplane_codes(qv, planes=planes)
    adj = np.full((base.shape[0], 8), -1, np.int32)
    idx = song.Index(base, adj, metric=metric)
    ids, d, ms = idx.bruteforce(q, k)
    idx.close()
    oi, od = oracle.bruteforce(base, q, k, metric=metric)
    if metric == "ip":
        # float data: the fp32 sums differ from the float64 oracle in the last bits
        np.testing.assert_allclose(d, od, rtol=1e-5, atol=1e-5)
        agree = np.mean([len(set(a) & set(b)) / k for a, b in zip(ids, oi)])
        assert agree > 0.99
    else:
        np.testing.assert_array_equal(ids, oi)
        np.testing.assert_array_equal(d, od.astype(np.float32))


@pytest.mark.parametrize("dim", [1, 7, 64, 200, 256, 960])
def test_bithash_on_gpu(song, dim):
    """The GPU projection gives the CPU projection's codes bit for bit (the CPU one equals the
    reference's own BitHash, tests/test_build.py), for any dim and a partial last wave, with a
    zero vector (every bit 0), a vector and its negation, tiny values, and a vector made
    orthogonal to row 0 of R so that its first sum sits at rounding level."""
    rng = np.random.default_rng(dim)
    R = song.bithash_matrix(dim, 42)
    v = rng.normal(size=(1003, dim)).astype(np.float32)
    v[5] = 0.0
    v[7] = -v[8]
    v[6] *= np.float32(1e-30)
    r0 = R[0].astype(np.float64)
    x = v[9].astype(np.float64)
    v[9] = (x - (x @ r0) / (r0 @ r0) * r0).astype(np.float32)
    cpu = song.bithash_project(R, v)
    gpu, ms = song.bithash_project_device(R, v)
    np.testing.assert_array_equal(gpu, cpu)
    assert (gpu[5] == 0).all() and ms > 0
    with pytest.raises(song.SongError, match="dim"):
        song.bithash_project_device(np.zeros((256, 4096), np.float32), np.zeros((2, 4096), np.float32))
