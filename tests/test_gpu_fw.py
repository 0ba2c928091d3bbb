"""GPU parity of the weighted characteristic path length (bna_lp_weighted)
against the oracle (test_fw_oracle.py pins it).

Both run the same blocked Floyd-Warshall recurrence (128-vertex blocks) over
the float costs 1.0f / w, and the same sums in the same order, so the nodal
efficiencies must be equal bit for bit and Lp exactly equal: any distance off
by one unit in the last place would change a row's float sum.
"""
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc
from test_fw_oracle import csr_w, dyadic, random_wgraph
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


def check(ctx, N, R, C, V):
    g = bna.Graph.from_arrays(np.asarray(R, np.int32), np.asarray(C, np.int32),
                              None if V is None else np.asarray(V, np.float32))
    lp, eff, iso = ctx.lp_weighted(g)
    olp, oeff, oiso = orc.lpw(R, C, V, N, B=128)
    assert np.array_equal(eff, oeff), int(np.count_nonzero(eff != oeff))
    assert iso == oiso
    assert lp == olp or (np.isinf(lp) and np.isinf(olp)), (lp, olp)
    return lp


@pytest.mark.parametrize("N,p", [(2, 1.0), (50, 0.2), (128, 0.05), (129, 0.05), (333, 0.02), (700, 0.01)])
def test_random_weights(ctx, N, p):
    # one block, exactly one, one plus a vertex, several with a ragged last one
    say(f"lpw random N={N} p={p}")
    rng = np.random.default_rng(N)
    R, C, V = random_wgraph(N, p, rng, lambda r, n: r.uniform(0.2, 1.0, n).astype(np.float32))
    check(ctx, N, R, C, V)


@pytest.mark.parametrize("N,p", [(300, 0.015), (520, 0.006)])
def test_dyadic_and_disconnected(ctx, N, p):
    say(f"lpw dyadic N={N}")
    rng = np.random.default_rng(N + 1)
    R, C, V = random_wgraph(N, p, rng, dyadic)
    check(ctx, N, R, C, V)


def test_closed_forms_and_unit_weights(ctx):
    say("lpw closed forms")
    n = 25
    R, C, V = csr_w(n, [(i, j, 0.5) for i in range(n) for j in range(i + 1, n)])
    assert check(ctx, n, R, C, V) == 2.0
    n = 300  # a path: distances up to 299 across three blocks
    R, C, V = csr_w(n, [(i, i + 1, 1.0) for i in range(n - 1)])
    check(ctx, n, R, C, V)
    R, C, V = csr_w(40, [(i, j, 1.0) for i in range(30) for j in range(i + 1, 30)] + [(30, 31, 0.5)])
    check(ctx, 40, R, C, V)
    rng = np.random.default_rng(9)
    N = 260
    edges = [(a, b) for a in range(N) for b in range(a + 1, N) if rng.random() < 0.02]
    R, C = csr_from_edges(N, edges)
    lp = check(ctx, N, R, C, None)
    assert lp == pytest.approx(orc.lp(R, C, N)[0], rel=1e-13)
    R, C = csr_from_edges(6, [])
    assert np.isinf(check(ctx, 6, R, C, None))


def test_device_csr_from_build(ctx):
    # a weighted CSR built on the device from series, left there
    say("lpw device CSR N=600")
    x = bna.synth_factor_model(600, 200, 8, 0.7, 0xF00D)
    with ctx.series(x) as s:
        s.normalize(bna.BF16X3)
        d = s.build_device(density=3.0, weighted=True)
        try:
            lp, eff, iso = ctx.lp_weighted(d)
            h = s.to_host(d)
        finally:
            s.free_device(d)
    olp, oeff, oiso = orc.lpw(h.R, h.C, h.V, 600, B=128)
    assert np.array_equal(eff, oeff) and iso == oiso and lp == olp


def test_bad_weights_rejected(ctx):
    say("lpw bad weights")
    R, C, V = csr_w(4, [(0, 1, 0.5), (1, 2, 0.5), (2, 3, 0.5)])
    for bad in (-0.5, np.nan, -0.0):
        W = V.copy()
        W[2] = bad
        g = bna.Graph.from_arrays(R.astype(np.int32), C, W)
        with pytest.raises(bna.BnaError):
            ctx.lp_weighted(g)
    # the context still works afterwards
    check(ctx, 4, R, C, V)
    R1, C1 = csr_from_edges(1, [])
    with pytest.raises(bna.BnaError):
        ctx.lp_weighted(bna.Graph.from_arrays(R1.astype(np.int32), C1, None))
