"""The clustering-coefficient oracle (orc_cp, the CUCP tool's Cp_2,
W/Cp_GPU/CUCP/Cp.cu:111-471) pinned by closed forms and by a dense matrix cube.

- complete graphs: every vertex has Cp = 1, weighted or not (w = 1 after the
  scaling by the largest |V|);
- a star: no triangles, Cp = 0; degree-one vertices have t < 2 and give 0;
- a weighted triangle with a pendant: Cp = cube root of the weight product over
  the largest weight cubed, with the 2 / (t (t - 1)) factor of each vertex;
- random weighted graphs: diag(W @ W @ W) from numpy's dense product (another
  algorithm than the oracle's scatter loop) times the float degree factor;
- cliques of 100 past 2^24 entries: the reference's float degree term
  t = (float)R[i+1] - (float)R[i] (cal_k_kernel, Cp.cu:15-31) differs from the
  row length there, and Cp = 99 * 98 / (t (t - 1)) shows it.
"""
import numpy as np
import pytest

import orc


def csr_from_dense(A):
    """Symmetric dense weights (0 = no edge) as R (int64), C, V."""
    N = A.shape[0]
    rows, cols = np.nonzero(A)
    R = np.zeros(N + 1, np.int64)
    np.add.at(R, rows + 1, 1)
    return np.cumsum(R), cols.astype(np.int32), A[rows, cols].astype(np.float32)


def inv_factor(R):
    t = np.float32(R[1:]).astype(np.float32) - np.float32(R[:-1]).astype(np.float32)
    with np.errstate(divide="ignore"):
        inv = np.where(t < 2, np.float32(0), (1.0 / (t * (t - np.float32(1))).astype(np.float64)).astype(np.float32))
    return inv.astype(np.float32)


@pytest.mark.parametrize("n", [3, 4, 17])
@pytest.mark.parametrize("weight", [None, 0.6])
def test_complete_graph_is_one(n, weight):
    A = np.ones((n, n), np.float32) - np.eye(n, dtype=np.float32)
    R, Cc, V = csr_from_dense(A * (weight or 1.0))
    cp, mean = orc.cp(R, Cc, None if weight is None else V, n)
    assert np.allclose(cp, 1.0, rtol=2e-7, atol=0)
    assert mean == pytest.approx(1.0, rel=2e-7)


def test_star_is_zero():
    n = 9
    A = np.zeros((n, n), np.float32)
    A[0, 1:] = A[1:, 0] = np.linspace(0.3, 0.9, n - 1, dtype=np.float32)
    R, Cc, V = csr_from_dense(A)
    cp, mean = orc.cp(R, Cc, V, n)
    assert np.all(cp == 0.0) and mean == 0.0


def test_weighted_triangle_with_pendant():
    a, b, c, d = 0.5, 0.8, 0.4, 0.9  # (0,1) (1,2) (0,2) (0,3); largest |V| = d
    A = np.zeros((4, 4), np.float32)
    A[0, 1] = A[1, 0] = a
    A[1, 2] = A[2, 1] = b
    A[0, 2] = A[2, 0] = c
    A[0, 3] = A[3, 0] = d
    R, Cc, V = csr_from_dense(A)
    cp, _ = orc.cp(R, Cc, V, 4)
    prod = np.cbrt(a * b * c / d ** 3)
    assert cp[0] == pytest.approx(2 * prod / 6, rel=1e-6)  # degree 3
    assert cp[1] == pytest.approx(prod, rel=1e-6) and cp[2] == pytest.approx(prod, rel=1e-6)
    assert cp[3] == 0.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_weighted_matches_dense_cube(seed):
    rng = np.random.default_rng(seed)
    n = 120
    M = np.triu(rng.random((n, n)) < 0.15, 1)
    A = np.where(M, rng.uniform(0.2, 0.95, (n, n)), 0.0).astype(np.float32)
    A = A + A.T
    if seed == 3:
        A[5, :] = A[:, 5] = 0.0  # an isolated vertex
    R, Cc, V = csr_from_dense(A)
    cp, mean = orc.cp(R, Cc, V, n)
    alpha = np.float32(1.0 / np.float64(V[np.argmax(np.abs(V))]))
    Wt = np.zeros((n, n), np.float64)
    rows = np.repeat(np.arange(n), np.diff(R))
    Wt[rows, Cc] = np.power((V * alpha).astype(np.float64), 1.0 / 3).astype(np.float32)
    cci = np.diag(Wt @ Wt @ Wt)
    exp = cci.astype(np.float32) * inv_factor(R)
    assert np.allclose(cp, exp, rtol=1e-6, atol=1e-30)
    assert mean == pytest.approx(float(np.mean(np.abs(exp))), rel=1e-6)
    if seed == 3:
        assert cp[5] == 0.0


def clique_graph(n_cliques, m):
    """n_cliques disjoint cliques of m vertices, unit weights."""
    N = n_cliques * m
    R = np.arange(N + 1, dtype=np.int64) * (m - 1)
    base = (np.arange(N) // m) * m
    own = np.arange(N) % m
    offs = np.arange(m - 1)
    # row i: the clique's other members, ascending
    Cc = base[:, None] + offs[None, :] + (offs[None, :] >= own[:, None])
    return N, R, Cc.astype(np.int32).ravel()


def test_float_degree_term_past_2_24():
    N, R, Cc = clique_graph(2000, 100)  # 19.8 M entries
    assert R[-1] > 2 ** 24
    cp, _ = orc.cp(R, Cc, None, N)
    inv = inv_factor(R)
    exp = np.float32(99 * 98) * inv
    assert np.array_equal(cp, exp)
    quirk = (np.float32(R[1:]) - np.float32(R[:-1])) != 99
    assert quirk.sum() > 1000 and np.all(cp[~quirk] == np.float32(9702) * np.float32(1.0 / 9702))
