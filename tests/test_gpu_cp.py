"""GPU parity of the clustering coefficient (bna_cp, the CUCP tool's Cp_2,
W/Cp_GPU/CUCP/Cp.cu:111-471) against the oracle (test_cp_oracle.py pins it).

Both sum cci in double, in different orders, and round it to float once, so a
weighted Cp may differ from the oracle's by a float ulp or two of cci: the bar
is 1e-6 relative.  Binary graphs count triangles exactly, and the degree factor
is the same float computation on both sides, so binary Cp is bit-exact; so are
cliques past 2^24 entries, where the reference's float degree term shows.
"""
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc
from test_cp_oracle import clique_graph

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


def graph(N, R, Cc, V=None):
    return SimpleNamespace(N=N, nnz=int(R[-1]), R=R, C=Cc, V=V)


def built(ctx, N, L, density, weighted, seed):
    x = bna.synth_factor_model(N, L, 8, 0.7, seed)
    with ctx.series(x) as s:
        return s.build(density=density, weighted=weighted)


@pytest.mark.parametrize("N,density", [(2000, 2.0), (5000, 1.0)])
def test_weighted_built_graph(ctx, N, density):
    say(f"cp weighted N={N} density={density}")
    G = built(ctx, N, 150, density, True, 0xC0C0 + N)
    cp, mean = ctx.cp(G)
    ocp, omean = orc.cp(G.R, G.C, G.V, N)
    rel = np.abs(cp.astype(np.float64) - ocp) / np.maximum(np.abs(ocp), 1e-30)
    say(f"  nnz {G.nnz}: max rel {rel.max():.3g}, equal {np.mean(cp == ocp):.4f}, mean {mean:.6g}")
    assert np.all(np.isfinite(cp)) and rel.max() <= RTOL
    assert np.array_equal(cp == 0, ocp == 0)
    assert mean == pytest.approx(omean, rel=RTOL)
    assert ctx.timings()["cp_ms"] > 0


def test_binary_exact_and_device_resident(ctx):
    N = 3000
    say(f"cp binary N={N}")
    x = bna.synth_factor_model(N, 120, 8, 0.7, 0xC0C1)
    with ctx.series(x) as s:
        G = s.build(density=1.5, weighted=False)
        ocp, omean = orc.cp(G.R, G.C, None, N)
        cp, mean = ctx.cp(G)
        assert np.array_equal(cp, ocp) and mean == omean
        Gw = s.build(density=1.5, weighted=True)
        D = s.build_device(density=1.5, weighted=True)
        try:
            dcp, dmean = ctx.cp_device(D)
        finally:
            s.free_device(D)
    hcp, hmean = ctx.cp(Gw)
    assert np.array_equal(dcp, hcp) and dmean == hmean


def test_cliques_float_degree_term(ctx):
    N, R, Cc = clique_graph(2000, 100)
    say(f"cp cliques N={N} nnz={R[-1]}")
    g = graph(N, R, Cc)
    cp, _ = ctx.cp(g)
    ocp, _ = orc.cp(R, Cc, None, N)
    assert np.array_equal(cp, ocp)
    # weighted with equal weights: w = 1 after scaling, the same values
    wcp, _ = ctx.cp(graph(N, R, Cc, np.full(R[-1], 0.7, np.float32)))
    assert np.array_equal(wcp, ocp)


def test_small_and_edge_cases(ctx):
    # one vertex, no edges; a triangle with a pendant and an isolated vertex;
    # a negative largest weight (alpha < 0 turns every positive weight negative
    # and the cube root of a negative double is NaN, as pow gives it)
    cp, mean = ctx.cp(graph(1, np.zeros(2, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32)))
    assert cp.tolist() == [0.0] and mean == 0.0
    A = np.zeros((5, 5), np.float32)
    for (i, j, w) in [(0, 1, 0.5), (1, 2, 0.8), (0, 2, 0.4), (0, 3, 0.9)]:
        A[i, j] = A[j, i] = w
    rows, cols = np.nonzero(A)
    R = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=5))]).astype(np.int64)
    V = A[rows, cols]
    for vals in (V, -V, np.where(V == np.float32(0.9), np.float32(-0.95), V)):
        g = graph(5, R, cols.astype(np.int32), vals.astype(np.float32))
        cp, _ = ctx.cp(g)
        ocp, _ = orc.cp(R, cols, vals, 5)
        assert np.array_equal(np.isnan(cp), np.isnan(ocp))
        ok = ~np.isnan(ocp)
        assert np.allclose(cp[ok], ocp[ok], rtol=RTOL, atol=0)


def test_too_large_rejected(ctx):
    N = 524289
    g = graph(N, np.zeros(N + 1, np.int64), np.zeros(0, np.int32))
    with pytest.raises(bna.BnaError):
        ctx.cp(g)


def test_unsorted_host_csr_is_refused(ctx):
    g = bna.Graph.__new__(bna.Graph)
    g.N, g.nnz = 4, 6
    g.R = np.array([0, 3, 4, 5, 6])
    g.C = np.array([2, 1, 3, 0, 0, 0], np.int32)
    g.V = None
    with pytest.raises(bna.BnaError) as ei:
        ctx.cp(g)
    assert ei.value.status == bna.ERR_FORMAT
    g.C = np.array([1, 2, 3, 0, 0, 0], np.int32)
    cp, mean = ctx.cp(g)
    assert cp.shape == (4,)
