"""Batched construction: several subjects of different N and L through one
histogram launch and one emission launch of the tile kernel.  Each subject's
CSR must equal its solo build bit for bit (and the oracle), and one subject's
failure must not disturb the others.
"""
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


SUBJECTS = [  # N, L, seed, options
    (700, 150, 1, dict(density=2.0)),
    (2049, 61, 2, dict(thr=0.45)),
    (3000, 240, 3, dict(density=1.0, weighted=False)),
    (1280, 96, 4, dict(thr=-0.5)),  # nearly every pair: its lists overflow and it runs again
    (517, 1200, 5, dict(density=5.0)),
]


def same(a, b):
    assert a.nnz == b.nnz
    assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C)
    assert (a.V is None) == (b.V is None)
    if a.V is not None:
        assert np.array_equal(a.V, b.V)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_batch_equals_solo_and_oracle(ctx, prec):
    say(f"batch of {len(SUBJECTS)} prec={prec}")
    xs = [bna.synth_factor_model(N, L, 8, 0.7, seed) for N, L, seed, _ in SUBJECTS]
    series = [ctx.series(x) for x in xs]
    try:
        opts = [bna.make_opts(precision=prec, **kw) for *_, kw in SUBJECTS]
        res = bna.build_batch(ctx, series, opts)
        t = ctx.timings()
        say(f"  corr launches {t['corr_launches']} corr_ms {t['corr_ms']:.2f} total_ms {t['total_ms']:.2f}")
        for (N, L, seed, kw), x, s, g in zip(SUBJECTS, xs, series, res):
            assert not isinstance(g, Exception), g
            solo = s.build(precision=prec, **kw)
            same(g, solo)
            z64, _, _, _, valid = orc.znorm(x, 0)
            if "density" in kw:
                og = orc.build_density(z64, valid, orc.density_k(N, kw["density"]))
            else:
                og = orc.build_threshold(z64, valid, kw["thr"])
            assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C), (N, L)
            if g.V is not None:
                assert np.abs(g.V.astype(np.float64) - og.V64).max(initial=0) <= 1e-6
            say(f"  N={N} L={L} {kw}: nnz={g.nnz} ok")
    finally:
        for s in series:
            s.close()


def test_batch_isolates_a_failing_subject(ctx):
    xs = [bna.synth_factor_model(900, 80, 6, 0.6, 7), bna.synth_factor_model(600, 50, 6, 0.6, 8)]
    xs[1][:, :590] = 3.0  # 10 valid voxels: 45 valid pairs
    series = [ctx.series(x) for x in xs]
    try:
        opts = [bna.make_opts(density=1.0), bna.make_opts(density=50.0)]
        res = bna.build_batch(ctx, series, opts)
        assert isinstance(res[1], bna.BnaError) and res[1].status == -11
        same(res[0], series[0].build(density=1.0))
    finally:
        for s in series:
            s.close()
    assert ctx._L.bna_trim(ctx._h) == 0
    assert ctx._L.bna_workspace_bytes(ctx._h) == 0


def test_batch_graph_replay_equals_direct_launches(ctx):
    # The batched build's emission launch and every subject's tails are one
    # captured HIP graph, replayed while the launch parameters stay the same:
    # each replay gives the bits of the direct launches.
    xs = [bna.synth_factor_model(N, L, 8, 0.7, seed) for N, L, seed, _ in SUBJECTS]
    series = [ctx.series(x) for x in xs]
    try:
        opts = [bna.make_opts(precision="bf16x3", **kw) for *_, kw in SUBJECTS]
        ctx.set_graphs(False)
        ref = bna.build_batch(ctx, series, opts)
        c0, r0 = ctx.batch_graph_stats()
        ctx.set_graphs(True)
        for rep in range(3):
            got = bna.build_batch(ctx, series, opts)
            for g, h in zip(got, ref):
                same(g, h)
        c1, r1 = ctx.batch_graph_stats()
        say(f"  batch graph: {c1 - c0} captures, {r1 - r0} replays in 3 builds")
        assert c1 > c0 and r1 - r0 >= 1
    finally:
        ctx.set_graphs(True)
        for s in series:
            s.close()
