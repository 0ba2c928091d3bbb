"""GPU parity of the characteristic path length (bna_lp) and betweenness
(bna_bc) against the oracle (test_paths_oracle.py pins it).

Lp: the count of ordered pairs at every distance must equal the oracle's
exactly; Lp itself is formed from those counts, so it differs from the
reference's one-by-one sum of N^2 reciprocals only by that sum's rounding
(bar: 1e-13 relative).  BC: path counts and dependencies are double on both
sides, added in source order; the device sums each vertex's successors in a
different order and rounds to float at the end (bar: 2e-6 relative, 1e-6
absolute for values near 0).
"""
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc
from test_paths_oracle import csr_from_edges

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


def graph(N, R, C):
    return bna.Graph.from_arrays(np.asarray(R, np.int32), np.asarray(C, np.int32), None)


def check(ctx, N, R, C, bc=True):
    g = graph(N, R, C)
    lp, hist = ctx.lp(g, hist_len=N + 1)
    olp, ohist = orc.lp(R, C, N, hist_len=N + 1)
    assert np.array_equal(hist, ohist)
    if np.isinf(olp):
        assert np.isinf(lp)
    else:
        assert abs(lp - olp) <= 1e-13 * olp, (lp, olp)
    if bc:
        b = ctx.bc(g)
        ob = orc.bc(R, C, N)
        err = np.abs(b.astype(np.float64) - ob)
        assert np.all(err <= 2e-6 * np.abs(ob) + 1e-6), float(err.max())
    return lp


def test_closed_forms(ctx):
    n = 40
    say("paths: clique, star, path, two cliques, edgeless")
    R, C = csr_from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
    assert check(ctx, n, R, C) == 1.0
    R, C = csr_from_edges(n + 1, [(0, k) for k in range(1, n + 1)])
    check(ctx, n + 1, R, C)
    b = ctx.bc(graph(n + 1, R, C))
    assert b[0] == n * (n - 1) and np.all(b[1:] == 0)
    R, C = csr_from_edges(300, [(i, i + 1) for i in range(299)])  # 299 levels
    check(ctx, 300, R, C)
    R, C = csr_from_edges(20, [(a + o, b + o) for o in (0, 10) for a in range(10) for b in range(a + 1, 10)])
    check(ctx, 20, R, C)
    R, C = csr_from_edges(5, [])
    check(ctx, 5, R, C)


@pytest.mark.parametrize("N,p", [(2, 1.0), (257, 0.03), (700, 0.01), (1300, 0.006)])
def test_random_graphs(ctx, N, p):
    # more than one batch of 256 sources, ragged last batch, isolated vertices
    say(f"paths random N={N} p={p}")
    rng = np.random.default_rng(N)
    iu = np.triu_indices(N, 1)
    keep = rng.random(iu[0].size) < p
    R, C = csr_from_edges(N, list(zip(iu[0][keep].tolist(), iu[1][keep].tolist())))
    check(ctx, N, R, C, bc=N <= 700)


def test_built_graph_device_resident(ctx):
    # a threshold graph of the factor model, straight from the device CSR
    N, L = 1500, 160
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A9000)
    with ctx.series(x) as s:
        s.normalize(bna.BF16X3)
        gd = s.build_device(thr=0.45, precision="bf16x3", weighted=False)
        try:
            h = s.to_host(gd)
            lp_dev = ctx.lp(gd)
            bc_dev = ctx.bc(gd)
        finally:
            s.free_device(gd)
    R = np.asarray(h.R, np.int64)
    olp, _ = orc.lp(R, h.C, N)
    assert abs(lp_dev - olp) <= 1e-13 * olp
    ob = orc.bc(R, h.C, N)
    assert np.all(np.abs(bc_dev - ob) <= 2e-6 * np.abs(ob) + 1e-6)
    say(f"  built graph: nnz {h.nnz}, Lp {lp_dev:.6g}, max BC {bc_dev.max():.6g}")


def test_errors(ctx):
    R, C = csr_from_edges(1, [])
    with pytest.raises(bna.BnaError):
        ctx.lp(graph(1, R, C))
