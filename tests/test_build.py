"""CPU NSW build and BitHash of libsong.so against the reference's own code.

The reference's build_graph (graph.h:190-199, textually the same algorithm as
build_kernel_graph, kernalGraph.h:140-149) and BitHash (bitHash.h:17-56) are
compiled from its sources into oracle/_ref; the product must reproduce them
exactly.  Invariants (main.cpp:84-93 and the reference's degree statistics)
cover the parameterised and batched builds the reference cannot run.
"""
import numpy as np
import pytest

REF_ONLY = pytest.mark.skipif("not __import__('oracle').ref_available()",
                              reason="oracle/_ref not built (needs the reference sources)")


@REF_ONLY
@pytest.mark.parametrize("n,dim,exact_order", [(2500, 16, True), (2000, 128, False), (1500, 3, False)])
def test_exact_build_matches_reference_integer(song, oracle, n, dim, exact_order):
    rng = np.random.default_rng(n + dim)
    data = rng.integers(0, 256, size=(n, dim)).astype(np.float32)
    np.testing.assert_array_equal(song.nsw_build(data, block=32, exact_order=exact_order),
                                  oracle.ref_build_graph(data))


@REF_ONLY
def test_exact_build_matches_reference_float(song, oracle):
    """Float data: the distance.h summation order reproduces the reference graph."""
    rng = np.random.default_rng(7)
    data = rng.standard_normal((2000, 20)).astype(np.float32)
    np.testing.assert_array_equal(song.nsw_build(data, block=32, exact_order=True),
                                  oracle.ref_build_graph(data))


@REF_ONLY
def test_exact_build_with_duplicates(song, oracle):
    """Repeated points (zero distances, ties everywhere)."""
    rng = np.random.default_rng(8)
    pts = rng.integers(0, 3, size=(60, 4)).astype(np.float32)
    data = pts[rng.integers(0, 60, size=1200)]
    np.testing.assert_array_equal(song.nsw_build(data, block=32), oracle.ref_build_graph(data))


def check_invariants(adj, fixed, search):
    n, block = adj.shape
    deg = (adj >= 0).sum(1)
    assert (adj[:, fixed:] == -1).all()  # slots beyond FIXED_DEGREE stay empty (sentinel)
    # filled slots are a prefix of the row
    for u in range(n):
        row = adj[u]
        k = deg[u]
        assert (row[:k] >= 0).all() and (row[k:] == -1).all()
        assert len(set(row[:k].tolist())) == k  # no duplicate edges
        assert u not in row[:k]  # no self loops
        if u >= 1:
            assert deg[u] >= min(search, u)  # main.cpp:84-93: every inserted node has edges
    assert adj.max() < n


@pytest.mark.parametrize("block", [8, 16, 32, 64])
def test_build_invariants(song, datasets, block):
    base, _ = datasets.gmm_int(3000, 1, dim=32, n_centres=20, sigma=20.0, seed=block)
    b, f, s = song.degrees_for_block(block)
    adj = song.nsw_build(base, block=block, exact_order=False)
    check_invariants(adj, f, s)


def test_batched_build_deterministic(song, datasets):
    cfg = datasets.config1()
    a1 = song.nsw_build(cfg["base"], block=16, n_threads=1, batch_div=32, seq_prefix=512, exact_order=False)
    a4 = song.nsw_build(cfg["base"], block=16, n_threads=4, batch_div=32, seq_prefix=512, exact_order=False)
    np.testing.assert_array_equal(a1, a4)
    check_invariants(a1, 15, 7)
    # with the sequential prefix covering every node the schedule is the exact build
    ex = song.nsw_build(cfg["base"][:3000], block=16, exact_order=False)
    bx = song.nsw_build(cfg["base"][:3000], block=16, n_threads=4, batch_div=8, seq_prefix=3000,
                        exact_order=False)
    np.testing.assert_array_equal(ex, bx)


def test_batched_build_quality(song, oracle, datasets):
    """Recall of the exact search oracle on the batched graph is close to the sequential graph's."""
    cfg = datasets.config1()
    seq = song.nsw_build(cfg["base"], block=16, exact_order=False)
    bat = song.nsw_build(cfg["base"], block=16, n_threads=4, batch_div=64, seq_prefix=1024, exact_order=False)
    gt, _ = oracle.bruteforce(cfg["base"], cfg["queries"], 10)

    def rec(adj):
        ids, _, _ = oracle.search(cfg["base"], adj, cfg["queries"], 10, 64, block=16, bloom_bits=1 << 15, group=8)
        return np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt)])

    assert rec(bat) >= rec(seq) - 0.03


def test_hamming_build_invariants(song, datasets):
    vec, _ = datasets.vmf_mix(2000, 1, dim=32, n_centres=10, kappa=10.0, seed=1)
    codes, _ = datasets.hyperplane_codes(vec, seed=2)
    adj = song.nsw_build(codes, metric="hamming", block=32)
    check_invariants(adj, 31, 15)


def test_graph_file_roundtrip(song, tmp_path):
    rng = np.random.default_rng(0)
    adj = rng.integers(-1, 100, size=(100, 32)).astype(np.int32)
    p = tmp_path / "g.bin"
    song.graph_save(str(p), adj, metric="l2", dim=128)
    back, meta = song.graph_load(str(p))
    np.testing.assert_array_equal(back, adj)
    assert meta == {"metric": 0, "dim": 128}


@REF_ONLY
@pytest.mark.parametrize("dim", [1, 16, 128])
def test_bithash_matches_reference(song, oracle, dim):
    rng = np.random.default_rng(dim)
    R = song.bithash_matrix(dim, 42)
    np.testing.assert_array_equal(R, oracle.ref_bithash_matrix(dim, 42))
    v = rng.standard_normal((300, dim)).astype(np.float32)
    codes = song.bithash_project(R, v, n_threads=3)
    np.testing.assert_array_equal(codes.view(np.uint32).reshape(300, 8), oracle.ref_bithash(v, 42))


def test_build_rejects_bad_arguments(song):
    data = np.zeros((10, 4), np.float32)
    with pytest.raises(song.SongError):
        song.nsw_build(data, block=1)
    with pytest.raises(song.SongError):
        song.nsw_build(data, block=8, fixed_deg=8)
    with pytest.raises(song.SongError):
        song.nsw_build(data, block=8, fixed_deg=7, search_deg=8)
    with pytest.raises(song.SongError):
        song.nsw_build(data, block=65)


def test_workload_generators(datasets):
    """The benchmark workloads' generators, at small n: shapes, value ranges and dtypes as
    described (960-d values in [0, 1]; 256-bit codes as 4 x uint64 whose bits follow the sign
    of the hyperplane projections; unit vectors for the inner-product workload)."""
    c3 = datasets.config3(n=500, nq=20)
    assert c3["base"].shape == (500, 960) and c3["queries"].shape == (20, 960)
    assert c3["base"].min() >= 0.0 and c3["base"].max() <= 1.0
    c4 = datasets.config4(n=500, nq=20, n_extra=30)
    assert c4["base"].shape == (500, 4) and c4["base"].dtype == np.uint64
    assert c4["queries"].shape == (20, 4) and c4["extra"].shape == (30, 4)
    bits = np.unpackbits(c4["base"].view(np.uint8), axis=1)
    assert 0.3 < bits.mean() < 0.7  # random hyperplanes split the sphere about evenly
    c5 = datasets.config5(n=500, nq=20, n_extra=10)
    assert c5["base"].shape == (500, 200) and c5["extra"].shape == (10, 200)
    np.testing.assert_allclose(np.linalg.norm(c5["base"], axis=1), 1.0, rtol=1e-5)
    assert set(c5["ks"]) == {1, 10, 100} and set(c5["efs"]) == {64, 256}
