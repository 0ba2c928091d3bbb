"""Pins of the oracle's weighted characteristic path length (orc_lpw), the
CUBFW_Lp tool (W/src/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:13-162).

orc_lpw runs blocked Floyd-Warshall over the float costs 1.0f / w.  Its
distances are checked against Dijkstra with double path sums (orc_sssp64, a
different method): equal bit for bit when every cost is a power of two (all
path sums are then exact), and within float rounding of the path sums
otherwise; the same holds across block sizes, one block being the plain
recurrence.  The Lp and nodal-efficiency sums are checked against sums formed
here from the Dijkstra distances, against closed forms (complete graphs, paths,
isolated vertices) and, for unit weights, against the binary tool's Lp.
"""
import numpy as np
import pytest

import orc
from test_paths_oracle import csr_from_edges


def csr_w(N, wedges):
    """Undirected weighted CSR: each edge in both rows with the same weight,
    columns ascending."""
    adj = [dict() for _ in range(N)]
    for a, b, w in wedges:
        if a != b:
            adj[a][b] = w
            adj[b][a] = w
    R = np.zeros(N + 1, np.int64)
    cols, vals = [], []
    for i in range(N):
        for j in sorted(adj[i]):
            cols.append(j)
            vals.append(adj[i][j])
        R[i + 1] = len(cols)
    return R, np.array(cols, np.int32), np.array(vals, np.float32)


def random_wgraph(N, p, rng, weights):
    iu = np.triu_indices(N, 1)
    keep = rng.random(iu[0].size) < p
    a, b = iu[0][keep], iu[1][keep]
    w = weights(rng, a.size)
    return csr_w(N, zip(a.tolist(), b.tolist(), w.tolist()))


def dyadic(rng, n):
    return rng.choice(np.array([1.0, 0.5, 0.25, 0.125], np.float32), n)


def sums_from_dist(D):
    """The reference's sums over a distance matrix, formed here sequentially
    with numpy: per row, float and double sums of 1.0f / d over v != u with d
    not below 0.1 (CUBFW_Lp.cpp:70-113)."""
    N = D.shape[0]
    eff = np.empty(N, np.float32)
    total = 0.0
    iso = 0
    for u in range(N):
        d = D[u].astype(np.float32)
        m = np.ones(N, bool)
        m[u] = False
        m &= ~(d.astype(np.float64) < 0.1)
        t = np.float32(1.0) / d[m]
        li = np.cumsum(t, dtype=np.float32)[-1] if t.size else np.float32(0)
        total += float(np.cumsum(t.astype(np.float64))[-1]) if t.size else 0.0
        e = np.float32(li / np.float32(N - 1))
        if e < np.float32(1.0) / np.float32(N):
            e = np.float32(0)
            iso += 1
        eff[u] = e
    return 1.0 / (total / N / (N - 1)), eff, iso


@pytest.mark.parametrize("N,p,B", [(37, 0.15, 8), (130, 0.05, 16), (200, 0.02, 64), (150, 0.04, 128)])
def test_dyadic_costs_exact(N, p, B):
    # costs 1, 2, 4, 8: every path sum is an exact float, so the blocked
    # recurrence and Dijkstra agree bit for bit; disconnected parts included
    rng = np.random.default_rng(N * 7 + B)
    R, C, V = random_wgraph(N, p, rng, dyadic)
    lp, eff, iso, D = orc.lpw(R, C, V, N, B=B, with_dist=True)
    D64 = orc.sssp64(R, C, V, N)
    assert np.array_equal(D.astype(np.float64), D64)
    lp2, eff2, iso2 = sums_from_dist(D64)
    assert np.array_equal(eff, eff2) and iso == iso2
    assert lp == lp2


@pytest.mark.parametrize("N,p,B", [(90, 0.08, 16), (260, 0.03, 64), (300, 0.02, 128)])
def test_real_weights_close_to_dijkstra(N, p, B):
    rng = np.random.default_rng(N + B)
    R, C, V = random_wgraph(N, p, rng, lambda r, n: r.uniform(0.2, 1.0, n).astype(np.float32))
    lp, eff, iso, D = orc.lpw(R, C, V, N, B=B, with_dist=True)
    D64 = orc.sssp64(R, C, V, N)
    fin = np.isfinite(D64)
    assert np.array_equal(fin, np.isfinite(D))
    # float sums of at most N costs of at most 5: a few units in the last place
    rel = np.abs(D[fin] - D64[fin]) / np.maximum(D64[fin], 1e-30)
    assert rel.max() <= 4e-6, rel.max()
    # one block (the plain recurrence) vs small blocks: same paths, rounding only
    _, _, _, Dp = orc.lpw(R, C, V, N, B=(N + 15) // 16 * 16, with_dist=True)
    assert np.array_equal(np.isfinite(Dp), fin)
    assert np.abs(Dp[fin] - D[fin]).max() <= 4e-6 * D[fin].max()
    lp2, eff2, iso2 = sums_from_dist(D)
    assert np.array_equal(eff, eff2) and iso == iso2 and lp == lp2


def test_closed_forms():
    # complete graph, weight 0.5: every distance 2, efficiency 0.5, Lp 2
    n = 25
    R, C, V = csr_w(n, [(i, j, 0.5) for i in range(n) for j in range(i + 1, n)])
    lp, eff, iso = orc.lpw(R, C, V, n, B=8)
    assert lp == 2.0 and np.all(eff == np.float32(0.5)) and iso == 0
    # complete graph, weight w: one hop (1/w) beats two (2/w)
    R, C, V = csr_w(n, [(i, j, 0.3) for i in range(n) for j in range(i + 1, n)])
    lp, eff, iso, D = orc.lpw(R, C, V, n, B=8, with_dist=True)
    c = np.float32(1.0) / np.float32(0.3)
    assert np.all(D[~np.eye(n, dtype=bool)] == c) and np.all(np.diag(D) == 0)
    # path of unit costs: d(i, j) = |i - j|, efficiency from harmonic sums
    n = 60
    R, C, V = csr_w(n, [(i, i + 1, 1.0) for i in range(n - 1)])
    lp, eff, iso, D = orc.lpw(R, C, V, n, B=16, with_dist=True)
    idx = np.arange(n)
    assert np.array_equal(D, np.abs(idx[:, None] - idx[None, :]).astype(np.float32))
    H = np.concatenate([[0.0], np.cumsum(1.0 / np.arange(1, n))])
    for u in range(n):
        assert eff[u] == pytest.approx((H[u] + H[n - 1 - u]) / (n - 1), rel=1e-6)
    # the reference's terms are float reciprocals, added in double
    pairs = 2 * sum((n - d) * float(np.float32(1.0) / np.float32(d)) for d in range(1, n))
    assert lp == pytest.approx(n * (n - 1) / pairs, rel=1e-13)


def test_isolated_and_weak_vertices():
    # 40 vertices: a 30-clique of weight 1, a pair of weight 0.5, 8 isolated.
    # The pair's vertices reach one vertex at cost 2: efficiency 0.5 / 39 is
    # below 1 / 40, so they count as isolated too (CUBFW_Lp.cpp:101-107).
    N = 40
    e = [(i, j, 1.0) for i in range(30) for j in range(i + 1, 30)] + [(30, 31, 0.5)]
    R, C, V = csr_w(N, e)
    lp, eff, iso = orc.lpw(R, C, V, N, B=16)
    assert iso == 10
    assert np.all(eff[30:] == 0) and np.all(eff[:30] == np.float32(np.float32(29.0) / np.float32(39)))
    tot = 30 * 29 * 1.0 + 2 * 0.5
    assert lp == pytest.approx(N * (N - 1) / tot, rel=1e-15)
    # a zero weight is an infinite cost: no edge
    R, C, V = csr_w(3, [(0, 1, 0.0), (1, 2, 1.0)])
    lp, eff, iso, D = orc.lpw(R, C, V, 3, B=8, with_dist=True)
    assert np.isinf(D[0, 1]) and np.isinf(D[0, 2]) and D[1, 2] == 1


@pytest.mark.parametrize("seed", [4, 5])
def test_unit_weights_match_binary_lp(seed):
    # without weights every cost is 1, so d is the hop distance of CUBFS_Lp
    rng = np.random.default_rng(seed)
    N = 120
    edges = [(a, b) for a in range(N) for b in range(a + 1, N) if rng.random() < 0.03]
    R, C = csr_from_edges(N, edges)
    lpw, eff, iso = orc.lpw(R, C, None, N, B=32)
    lpb, _ = orc.lp(R, C, N)
    assert lpw == pytest.approx(lpb, rel=1e-13)
