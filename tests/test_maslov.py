"""The Maslov random networks (bna_maslov, host code, no device): the rewiring
of W/Cp_GPU/CUCP/Maslov.cpp:49-289 draws from the clock there, so the check is
on what every network it can produce satisfies: the degree of every vertex,
symmetry, no self-loops or repeated edges, ascending rows, the multiset of
weights (each weight on both copies of its edge), and, for graphs that admit no
swap (complete graphs, stars), the input itself."""
from collections import Counter

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
from test_fw_oracle import csr_w, random_wgraph
from test_paths_oracle import csr_from_edges


def graph(R, C, V=None):
    return bna.Graph.from_arrays(np.asarray(R, np.int32), np.asarray(C, np.int32),
                                 None if V is None else np.asarray(V, np.float32))


def edges_of(g):
    """{(a, b): weight} over a < b, checking symmetry, order and self-loops."""
    R, C = g.R.astype(np.int64), g.C
    e = {}
    for i in range(len(R) - 1):
        row = C[R[i]:R[i + 1]]
        assert np.all(np.diff(row) > 0), "columns ascend strictly"
        assert not np.any(row == i), "no self-loop"
        for k in range(R[i], R[i + 1]):
            j = int(C[k])
            w = None if g.V is None else float(g.V[k])
            if i < j:
                e[(i, j)] = w
            else:
                assert e.get((j, i), "missing") == w, "symmetric, same weight in both rows"
    assert 2 * len(e) == len(C)
    return e


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_degrees_weights_and_structure_kept(seed):
    rng = np.random.default_rng(seed)
    N = 300
    R, C, V = random_wgraph(N, 0.04, rng, lambda r, n: r.uniform(0.2, 1.0, n).astype(np.float32))
    g = graph(R, C, V)
    e0 = edges_of(g)
    for permute in (False, True):
        h = bna.maslov(g, seed, permute)
        assert np.array_equal(h.R, g.R)  # every degree kept
        e1 = edges_of(h)
        assert Counter(e1.values()) == Counter(e0.values())  # the same weights
        moved = len(set(e1) - set(e0))
        assert moved > len(e0) // 2, moved  # 2M attempts rewire most edges
        assert bna.check_csr(h) == 0


def test_weights_travel_with_edges():
    # two disjoint edges with different weights: the swap a-b, c-d -> a-c, b-d
    # keeps one weight per new edge; without permutation each weight stays on
    # the edge that took the place of its own
    R, C, V = csr_w(4, [(0, 1, 0.25), (2, 3, 0.75)])
    g = graph(R, C, V)
    seen = set()
    for seed in range(40):
        e = edges_of(bna.maslov(g, seed))
        assert sorted(e.values()) == [0.25, 0.75]
        seen.add(tuple(sorted(e)))
    assert len(seen) >= 2  # both rewired layouts appear, and the input's too


def test_same_seed_same_network():
    rng = np.random.default_rng(7)
    R, C = csr_from_edges(200, [(a, b) for a in range(200) for b in range(a + 1, 200) if rng.random() < 0.05])
    g = graph(R, C)
    a, b, c = bna.maslov(g, 11), bna.maslov(g, 11), bna.maslov(g, 12)
    assert a.V is None
    assert np.array_equal(a.C, b.C) and not np.array_equal(a.C, c.C)


def test_graphs_without_swaps_are_unchanged():
    n = 12
    for edges in ([(i, j) for i in range(n) for j in range(i + 1, n)],  # complete: every new edge exists
                  [(0, j) for j in range(1, n)],                        # star: every pair shares the centre
                  [(0, 1)], []):
        R, C = csr_from_edges(n, edges)
        g = graph(R, C)
        h = bna.maslov(g, 5)
        assert np.array_equal(h.R, g.R) and np.array_equal(h.C, g.C)


def test_bad_inputs_rejected():
    with pytest.raises(bna.BnaError):  # asymmetric: row 1 holds 0, row 0 does not hold 1
        bna.maslov(graph([0, 0, 2, 3], [0, 2, 1]), 1)
    with pytest.raises(bna.BnaError):  # a self-loop
        bna.maslov(graph([0, 1, 1, 1], [0]), 1)
