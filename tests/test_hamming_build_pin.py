"""The Hamming NSW build (build_hashed_graph, hashKernelGraph.h:132-207) against an independent
restatement, on codes where Hamming distances tie everywhere.

The reference's frontier is a std::priority_queue whose comparator looks at the distance only
(hashKernelGraph.h:135), so which of several equally distant nodes it expands first is decided by
the heap's own sift order.  The reference's sources for this build include CUDA headers and do not
build here, so the pin restates the algorithm in Python together with the heap algorithm of the
GNU C++ library the image's g++ uses (std::push_heap / std::pop_heap: sift up from the new leaf;
pop moves the last element's hole down along the larger child to a leaf, then sifts it up).  The
library's build must give the same adjacency, slot for slot.  A planted mistake -- the float
metrics' full (distance, id) frontier order in place of the distance-only one -- must change the
graph on this data, so the test sees the tie order.
"""
import numpy as np
import pytest


class StlHeap:
    """std::priority_queue<T, std::vector<T>, Comp> with libstdc++'s heap algorithm."""

    def __init__(self, comp):
        self.c = []
        self.comp = comp

    def top(self):
        return self.c[0]

    def __len__(self):
        return len(self.c)

    def _sift_up(self, hole, top, value):
        c, comp = self.c, self.comp
        parent = (hole - 1) // 2
        while hole > top and comp(c[parent], value):
            c[hole] = c[parent]
            hole = parent
            parent = (hole - 1) // 2
        c[hole] = value

    def push(self, x):
        self.c.append(x)
        self._sift_up(len(self.c) - 1, 0, x)

    def pop(self):
        c, comp = self.c, self.comp
        n = len(c)
        if n > 1:
            value = c[-1]
            c[-1] = c[0]
            length = n - 1
            hole = second = 0
            while second < (length - 1) // 2:
                second = 2 * (second + 1)
                if comp(c[second], c[second - 1]):
                    second -= 1
                c[hole] = c[second]
                hole = second
            if length % 2 == 0 and second == (length - 2) // 2:
                second = 2 * (second + 1)
                c[hole] = c[second - 1]
                hole = second - 1
            self._sift_up(hole, 0, value)
        c.pop()


def distance_only(a, b):  # hashKernelGraph.h:135, a min-heap on the distance alone
    return a[0] > b[0]


def full_pair(a, b):  # std::greater<pair>: the float metrics' frontier order
    return a > b


def build_hashed(D, block, fixed, search, ef, frontier=distance_only):
    """hkg_search_layer / hkg_select_neighbors / add_vertex / hkg_add_reverse_edge
    (hashKernelGraph.h:35-38, 132-207) over a precomputed distance matrix D."""
    n = D.shape[0]
    adj = np.full((n, block), -1, np.int64)
    for u in range(1, n):
        cands, res = StlHeap(frontier), StlHeap(lambda a, b: a < b)
        visited = {0}
        cands.push((int(D[u, 0]), 0))
        res.push((int(D[u, 0]), 0))
        while len(cands):
            cd, cu = cands.top()
            cands.pop()
            if len(res) >= ef and cd > res.top()[0]:
                break
            for j in range(block):
                v = int(adj[cu, j])
                if v < 0 or v >= u or v in visited:
                    continue
                visited.add(v)
                d = int(D[u, v])
                cands.push((d, v))
                res.push((d, v))
                if len(res) > ef:
                    res.pop()
        ret = []
        while len(res):
            ret.append(res.top())
            res.pop()
        ret.reverse()
        nb = [e for _, e in ret[:search]]
        for j, e in enumerate(nb[:fixed]):
            adj[u, j] = e
        for v in nb:
            row = adj[v]
            free = np.nonzero(row[:fixed] < 0)[0]
            if free.size:
                row[free[0]] = u
                continue
            du = D[v, u]
            fj, fd = 0, D[v, row[0]]
            for j in range(1, fixed):
                if D[v, row[j]] > fd:
                    fd, fj = D[v, row[j]], j
            if du < fd:
                row[fj] = u
    return adj


def tie_heavy_codes(n, n_base, flips, seed, dup_every=7):
    """Codes near a few random bases (0..flips bits flipped), every dup_every-th a repeat of an
    earlier code: distances are small integers shared by many pairs, and zero for repeats."""
    rng = np.random.default_rng(seed)
    bases = rng.integers(0, np.iinfo(np.uint64).max, size=(n_base, 4), dtype=np.uint64, endpoint=True)
    codes = bases[rng.integers(0, n_base, size=n)].copy()
    for i in range(n):
        for b in rng.choice(256, size=rng.integers(0, flips + 1), replace=False):
            codes[i, b // 64] ^= np.uint64(1) << np.uint64(b % 64)
        if i and i % dup_every == 0:
            codes[i] = codes[rng.integers(0, i)]
    return codes


def hamming_matrix(codes):
    bits = np.unpackbits(codes.view(np.uint8), axis=1).astype(np.int32)
    return (bits[:, None, :] != bits[None, :, :]).sum(2)


@pytest.mark.parametrize("n,block,fixed,search,ef,n_base,flips,seed", [
    (600, 16, 15, 8, 24, 5, 4, 1),
    (400, 16, 15, 8, 40, 3, 2, 2),
    (260, 32, 31, 15, 150, 6, 5, 3),  # the reference's own degrees and EF_CONSTRUCTION
])
def test_hamming_build_equals_restatement_with_ties(song, n, block, fixed, search, ef, n_base, flips, seed):
    codes = tie_heavy_codes(n, n_base, flips, seed)
    D = hamming_matrix(codes)
    want = build_hashed(D, block, fixed, search, ef)
    got = song.nsw_build(codes, metric="hamming", block=block, fixed_deg=fixed, search_deg=search,
                         ef_construction=ef)
    np.testing.assert_array_equal(got, want)


def test_planted_frontier_order_changes_the_graph():
    """The tie order matters on this data: a frontier that also orders by id builds another graph."""
    codes = tie_heavy_codes(600, 5, 4, 1)
    D = hamming_matrix(codes)
    a = build_hashed(D, 16, 15, 8, 24)
    b = build_hashed(D, 16, 15, 8, 24, frontier=full_pair)
    assert (a != b).any()


def test_stl_heap_is_a_priority_queue():
    """Whatever its internal order, the restated heap pops in priority order (distance-only
    comparator: distances non-decreasing; full pair order: sorted)."""
    rng = np.random.default_rng(4)
    xs = [(int(d), int(i)) for d, i in zip(rng.integers(0, 5, size=300), rng.permutation(300))]
    h = StlHeap(full_pair)
    for x in xs:
        h.push(x)
    out = []
    while len(h):
        out.append(h.top())
        h.pop()
    assert out == sorted(xs)
    h = StlHeap(distance_only)
    for x in xs:
        h.push(x)
    ds = []
    while len(h):
        ds.append(h.top()[0])
        h.pop()
    assert ds == sorted(d for d, _ in xs)
