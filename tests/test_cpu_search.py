"""song_cpu_search, the product's restatement of the reference's CPU search baseline
(Graph::search, graph.h:33-115): pinned to the reference's own outputs (the golden vectors
written from its compiled sources, and oracle/_ref where it is built), to brute force where the
search must be exact, and to its argument checks.  Runs on the host only."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.mark.parametrize("tag", ["int", "float"])
@pytest.mark.parametrize("ef", [10, 50])
@pytest.mark.parametrize("threads", [1, 3])
def test_equals_reference_golden_vectors(song, tag, ef, threads):
    """Ids (worst first, as Graph::search returns them) and all 12 counters equal the
    reference's own Graph::search on integer data (distance ties) and float data."""
    z = np.load(GOLDEN)
    ids, d, tr, per, total = song.cpu_search(z[f"{tag}_base"], z[f"{tag}_adj"], z[f"{tag}_queries"], 10, ef,
                                             n_threads=threads, best_first=False)
    np.testing.assert_array_equal(ids, z[f"{tag}_ids_ef{ef}"])
    np.testing.assert_array_equal(tr, z[f"{tag}_trace_ef{ef}"])
    assert (per >= 0).all() and total >= 0


def test_equals_compiled_reference(song, oracle, datasets):
    """Against the reference's Graph::search compiled from its sources (oracle/_ref), on a
    graph of 32-slot rows (BLOCK_SIZE) with many distance ties, at several queue sizes."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built (the reference sources are absent)")
    base, q = datasets.gmm_int(3000, 60, dim=16, n_centres=8, sigma=6.0, seed=9)
    adj = song.nsw_build(base, block=32, exact_order=False)
    for ef in (10, 33, 100):
        rid, rtr, _, _ = oracle.ref_graph_search(base, adj, q, 10, ef)
        ids, d, tr, _, _ = song.cpu_search(base, adj, q, 10, ef, best_first=False)
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(tr, rtr)


def test_exact_when_the_queue_holds_every_node(song, datasets):
    """With ef >= n on a connected graph the search visits every node and keeps the k best:
    the ids are brute force's, best first after reversal, ties broken by the smaller id."""
    base, q = datasets.gmm_int(400, 20, dim=8, n_centres=4, sigma=10.0, seed=3)
    adj = song.nsw_build(base, block=16, exact_order=True)
    ids, d, tr, _, _ = song.cpu_search(base, adj, q, 10, 400)
    assert (tr[:, 1] == 400).all()  # every node visited
    diff = q[:, None, :].astype(np.float64) - base[None, :, :].astype(np.float64)
    dist = (diff * diff).sum(-1)  # exact for these integers
    for i in range(q.shape[0]):
        order = np.lexsort((np.arange(base.shape[0]), dist[i]))[:10]
        np.testing.assert_array_equal(ids[i], order)
        np.testing.assert_array_equal(d[i], dist[i][order].astype(np.float32))


def test_trace_relations_and_padding(song, datasets):
    """The reference's trace relations, and fewer reachable nodes than k pad with -1."""
    base, q = datasets.gmm_int(2000, 30, dim=12, n_centres=5, sigma=8.0, seed=4)
    adj = song.nsw_build(base, block=16, exact_order=False)
    _, _, tr, _, _ = song.cpu_search(base, adj, q, 10, 40)
    assert (tr[:, 1] == tr[:, 4] + tr[:, 0]).all()
    assert (tr[:, 5] == tr[:, 1]).all()
    assert (tr[:, 3] == 16 * (tr[:, 2] - tr[:, 11])).all()  # the stopping pop scans no row
    # 3 nodes, no edges: only the seeds are reachable
    small = np.arange(12, dtype=np.float32).reshape(3, 4)
    ids, d, tr, _, _ = song.cpu_search(small, np.full((3, 4), -1, np.int32), small[:1], 5, 5)
    assert tr[0, 10] == 3 and tr[0, 0] == 3
    np.testing.assert_array_equal(ids[0], [0, 1, 2, -1, -1])
    assert np.isinf(d[0, 3:]).all()


def test_argument_checks(song):
    base = np.zeros((10, 4), np.float32)
    adj = np.full((10, 8), -1, np.int32)
    q = np.zeros((2, 4), np.float32)
    bad_q = q.copy()
    bad_q[1, 2] = np.nan
    with pytest.raises(song.SongError, match="finite"):
        song.cpu_search(base, adj, bad_q, 3, 5)
    bad_adj = adj.copy()
    bad_adj[4, 1] = 10
    with pytest.raises(song.SongError, match="adjacency"):
        song.cpu_search(base, bad_adj, q, 3, 5)
    with pytest.raises(song.SongError, match="k must"):
        song.cpu_search(base, adj, q, 0, 5)
