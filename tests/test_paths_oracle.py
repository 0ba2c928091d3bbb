"""Pins of the oracle's shortest-path measures (orc_lp, orc_bc).

Lp follows CUBFS_Lp (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP.cpp:33-46); BC follows
CUBC's Brandes dependencies (B/src/BC/CUBC/BC_GPU_node_pred.cu).  Each is checked
against closed forms (complete graphs, stars, paths, disconnected cliques) and
against brute force that does not share the oracle's method: all-pairs hop
distances by repeated boolean matrix products, and betweenness as
sum over (s, t) of sigma_sv * sigma_vt / sigma_st over v on a shortest s-t path.
"""
import numpy as np
import pytest

import orc


def csr_from_edges(N, edges):
    adj = [set() for _ in range(N)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    R = np.zeros(N + 1, np.int64)
    cols = []
    for i in range(N):
        row = sorted(adj[i])
        cols.extend(row)
        R[i + 1] = len(cols)
    return R, np.array(cols, np.int32)


def lp_from_counts(N, counts):
    # the reference's arithmetic: float reciprocals, double sum, / N / (N - 1), inverted
    s = 0.0
    for d, c in counts.items():
        s += c * float(np.float32(1.0) / np.float32(d))
    return 1.0 / (s / N / (N - 1))


def dist_matrix(N, R, C):
    """Hop distances by repeated boolean products of the adjacency (-1: no path)."""
    A = np.zeros((N, N), bool)
    for i in range(N):
        A[i, C[R[i]:R[i + 1]]] = True
    D = np.full((N, N), -1, np.int64)
    np.fill_diagonal(D, 0)
    reach = np.eye(N, dtype=bool)
    frontier = np.eye(N, dtype=bool)
    d = 0
    while frontier.any():
        d += 1
        nxt = (frontier.astype(np.int64) @ A.astype(np.int64)) > 0
        nxt &= ~reach
        D[nxt] = d
        reach |= nxt
        frontier = nxt
    return D


def brute_bc(N, R, C):
    D = dist_matrix(N, R, C)
    A = np.zeros((N, N), np.int64)
    for i in range(N):
        A[i, C[R[i]:R[i + 1]]] = 1
    # path counts: S[s, t] = number of shortest s-t paths, level by level
    S = np.zeros((N, N))
    np.fill_diagonal(S, 1.0)
    for d in range(1, int(D.max()) + 1):
        for s in range(N):
            for t in np.nonzero(D[s] == d)[0]:
                prev = np.nonzero((D[s] == d - 1) & (A[:, t] > 0))[0]
                S[s, t] = S[s, prev].sum()
    bc = np.zeros(N)
    for v in range(N):
        for s in range(N):
            if s == v or D[s, v] < 0:
                continue
            for t in range(N):
                if t == v or t == s or D[s, t] < 0 or D[v, t] < 0:
                    continue
                if D[s, v] + D[v, t] == D[s, t]:
                    bc[v] += S[s, v] * S[v, t] / S[s, t]
    return bc


def test_lp_closed_forms():
    n = 9
    R, C = csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
    v, h = orc.lp(R, C, n)
    assert v == 1.0 and h[1] == n * (n - 1) and h[0] == 0
    # star: centre 0 and n leaves
    N = n + 1
    R, C = csr_from_edges(N, [(0, k) for k in range(1, N)])
    v, h = orc.lp(R, C, N)
    assert h[1] == 2 * n and h[2] == n * (n - 1)
    assert v == lp_from_counts(N, {1: 2 * n, 2: n * (n - 1)})
    # path of n vertices: 2 (n - d) ordered pairs at distance d
    R, C = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    v, h = orc.lp(R, C, n)
    for d in range(1, n):
        assert h[d] == 2 * (n - d)
    assert v == pytest.approx(lp_from_counts(n, {d: 2 * (n - d) for d in range(1, n)}), rel=1e-15)
    # two disjoint 5-cliques: unreachable pairs add nothing, N(N-1) still divides
    R, C = csr_from_edges(10, [(a + o, b + o) for o in (0, 5) for a in range(5) for b in range(a + 1, 5)])
    v, h = orc.lp(R, C, 10)
    assert h[1] == 40 and h[0] == 50
    assert v == lp_from_counts(10, {1: 40})
    # no edge: the sum is 0 and Lp is +inf
    R, C = csr_from_edges(4, [])
    v, h = orc.lp(R, C, 4)
    assert np.isinf(v) and h[0] == 12


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lp_against_matrix_distances(seed):
    rng = np.random.default_rng(seed)
    N = 45
    edges = [(a, b) for a in range(N) for b in range(a + 1, N) if rng.random() < 0.06]
    R, C = csr_from_edges(N, edges)
    D = dist_matrix(N, R, C)
    v, h = orc.lp(R, C, N)
    for d in range(1, 64):
        assert h[d] == np.count_nonzero(D == d)
    assert h[0] == np.count_nonzero(D < 0)
    # the reference adds the N^2 reciprocals in row-major order
    s = 0.0
    for i in range(N):
        for j in range(N):
            if D[i, j] > 0:
                s += float(np.float32(1.0) / np.float32(D[i, j]))
    assert v == 1.0 / (s / N / (N - 1))


def test_bc_closed_forms():
    n = 7
    R, C = csr_from_edges(n + 1, [(0, k) for k in range(1, n + 1)])
    b = orc.bc(R, C, n + 1)
    assert b[0] == n * (n - 1) and np.all(b[1:] == 0)
    R, C = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    b = orc.bc(R, C, n)
    assert np.allclose(b, [2 * k * (n - 1 - k) for k in range(n)], rtol=0, atol=1e-12)
    R, C = csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
    assert np.all(orc.bc(R, C, n) == 0)


@pytest.mark.parametrize("seed", [4, 5])
def test_bc_against_path_counts(seed):
    rng = np.random.default_rng(seed)
    N = 24
    edges = [(a, b) for a in range(N) for b in range(a + 1, N) if rng.random() < 0.15]
    R, C = csr_from_edges(N, edges)
    assert np.allclose(orc.bc(R, C, N), brute_bc(N, R, C), rtol=1e-12, atol=1e-12)


def test_bc_counts_ordered_pairs():
    # both (s, t) and (t, s) are counted, as CUBC's sum over every source is:
    # the values are twice the unordered-pair betweenness
    N = 12
    R, C = csr_from_edges(N, [(i, i + 1) for i in range(N - 1)] + [(0, 5), (3, 9)])
    good = orc.bc(R, C, N)
    assert np.allclose(good, brute_bc(N, R, C), rtol=1e-12, atol=1e-12)
    assert not np.allclose(good / 2, brute_bc(N, R, C))
