"""The oracle's Newman modularity partition (orc_modularity, the restatement
of Partition_GPU / Sub_Partition / Lead_Vector of the reference's
CUmodularity tool) pinned on the CPU: closed forms of Q for two disjoint
cliques and a ring of cliques, Q against the dense definition for the
partition found, and the first round's vector and eigenvalue against
numpy's eigendecomposition of the modularity matrix, including a graph whose
most negative eigenvalue dominates (the shifted second power iteration)."""
import numpy as np
import pytest

import orc


def csr_from_adj(A):
    A = np.asarray(A, bool)
    R = np.concatenate([[0], np.cumsum(A.sum(1))]).astype(np.int64)
    C = np.concatenate([np.nonzero(A[i])[0] for i in range(len(A))]).astype(np.int32)
    return R, C


def cliques(r, s, ring):
    N = r * s
    A = np.zeros((N, N), bool)
    for c in range(r):
        A[c * s:(c + 1) * s, c * s:(c + 1) * s] = True
        if ring and r > 1:
            a, b = c * s + s - 1, ((c + 1) % r) * s
            A[a, b] = A[b, a] = True
    np.fill_diagonal(A, False)
    return A


def q_dense(A, lab):
    k = A.sum(1).astype(float)
    m2 = k.sum()
    same = lab[:, None] == lab[None, :]
    return float(((A - np.outer(k, k) / m2) * same).sum() / m2)


def planted(N, groups, p_in, p_out, seed):
    rng = np.random.default_rng(seed)
    g = rng.integers(0, groups, N)
    P = np.where(g[:, None] == g[None, :], p_in, p_out)
    U = rng.random((N, N)) < P
    A = np.triu(U, 1)
    return A | A.T, g


def test_two_disjoint_cliques():
    A = cliques(2, 10, ring=False)
    R, C = csr_from_adj(A)
    lab, Q, mods, rounds, its = orc.modularity(R, C, 20)
    assert mods == 2
    assert Q == 0.5  # each clique: 1/2 of the edges, (1/2)^2 of the degree
    assert len(set(lab[:10])) == 1 and len(set(lab[10:])) == 1 and lab[0] != lab[10]


def test_ring_of_cliques_closed_form():
    r, s = 4, 12
    A = cliques(r, s, ring=True)
    R, C = csr_from_adj(A)
    lab, Q, mods, rounds, its = orc.modularity(R, C, r * s)
    # the partition found is the cliques, whose Q is r (e/m - (d/2m)^2) with
    # e = s(s-1)/2 edges inside a clique, m = r e + r, d = s(s-1) + 2
    assert mods == r
    for c in range(r):
        assert len(set(lab[c * s:(c + 1) * s])) == 1
    e, m, d = s * (s - 1) / 2, r * s * (s - 1) / 2 + r, s * (s - 1) + 2
    assert Q == pytest.approx(r * (e / m - (d / (2 * m)) ** 2), abs=1e-14)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_q_is_the_definition_of_the_partition_found(seed):
    A, _ = planted(300, 5, 0.3, 0.02, seed)
    R, C = csr_from_adj(A)
    lab, Q, mods, rounds, its = orc.modularity(R, C, 300)
    assert mods >= 2 and lab.min() == 1 and lab.max() == mods
    assert Q == pytest.approx(q_dense(A, lab), abs=1e-12)


@pytest.mark.parametrize("kind", ["planted", "bipartite"])
def test_first_round_is_the_leading_eigenvector(kind):
    if kind == "planted":
        A, _ = planted(60, 3, 0.5, 0.05, 11)
    else:
        # complete bipartite graph plus a clique on 8 nodes of one side: the
        # most negative eigenvalue of B dominates, so the shifted iteration
        # decides the split
        A = np.zeros((40, 40), bool)
        A[:20, 20:] = True
        A[20:, :20] = True
        A[:8, :8] = True
        np.fill_diagonal(A, False)
    R, C = csr_from_adj(A)
    N = len(A)
    k = A.sum(1).astype(float)
    B = A.astype(float) - np.outer(k, k) / k.sum()
    w, U = np.linalg.eigh(B)
    v, lam = orc.modularity_first(R, C, N)
    if kind == "bipartite":
        assert abs(w[0]) > abs(w[-1])  # the dominant eigenvalue is negative
    assert lam == pytest.approx(w[-1], rel=1e-6)
    cos = abs(float(v @ U[:, -1])) / np.linalg.norm(v)
    assert cos > 1 - 1e-9
