"""bna_modularity (Newman leading-eigenvector modularity, the CUmodularity
tool) against the oracle's restatement on the same binary graphs: the same
labels for every voxel, the same module count and Q within 1e-12.  Graphs:
planted partitions, a ring of cliques (whose Q has a closed form), one group
above 25000 nodes (where the reference switches to the GPU sub-partition and
its shift rule), and a device-resident CSR from a threshold build."""
import sys
import time

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


def graph_from_pairs(N, a, b):
    key = np.unique(np.minimum(a, b).astype(np.int64) * N + np.maximum(a, b))
    key = key[key // N != key % N]
    a, b = key // N, key % N
    rows = np.concatenate([a, b])
    cols = np.concatenate([b, a]).astype(np.int32)
    order = np.lexsort((cols, rows))
    g = bna.Graph.__new__(bna.Graph)
    g.N, g.nnz = N, len(rows)
    g.R = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=N))]).astype(np.int64)
    g.C = cols[order]
    g.V = None
    return g


def planted(N, groups, deg_in, deg_out, seed):
    rng = np.random.default_rng(seed)
    g = rng.integers(0, groups, N)
    members = [np.nonzero(g == c)[0] for c in range(groups)]
    m_in = N * deg_in // 2
    m_out = N * deg_out // 2
    a1 = rng.integers(0, N, m_in)
    b1 = np.array([rng.choice(members[g[x]]) for x in a1]) if N <= 4000 else \
        np.concatenate([members[c][rng.integers(0, len(members[c]), np.sum(g[a1] == c))] for c in range(groups)])
    if N > 4000:  # same multiset of in-group partners, drawn per group
        a1 = np.concatenate([a1[g[a1] == c] for c in range(groups)])
    a2 = rng.integers(0, N, m_out)
    b2 = rng.integers(0, N, m_out)
    return graph_from_pairs(N, np.concatenate([a1, a2]), np.concatenate([b1, b2]))


def check(ctx, g, what):
    t0 = time.time()
    lab, Q, mods = ctx.modularity(g)
    t_gpu = time.time() - t0
    t0 = time.time()
    olab, oQ, omods, rounds, its = orc.modularity(g.R, g.C, g.N)
    say(f"  {what}: N={g.N} nnz={g.nnz} modules {mods} (oracle {omods}) Q={Q:.12f} (oracle {oQ:.12f}) "
        f"rounds {rounds} oracle iterations {its}; gpu {t_gpu:.2f} s, oracle {time.time() - t0:.2f} s")
    assert mods == omods
    assert np.array_equal(lab, olab)
    assert Q == pytest.approx(oQ, abs=1e-12)
    return lab, Q


@pytest.mark.parametrize("N,groups,seed", [(2000, 5, 1), (3000, 8, 2), (1500, 3, 3)])
def test_planted_partitions(ctx, N, groups, seed):
    say(f"modularity planted N={N} groups={groups}")
    check(ctx, planted(N, groups, 16, 2, seed), "planted")


def test_ring_of_cliques(ctx):
    r, s = 4, 12
    a, b = [], []
    for c in range(r):
        for i in range(s):
            for j in range(i + 1, s):
                a.append(c * s + i)
                b.append(c * s + j)
        a.append(c * s + s - 1)
        b.append(((c + 1) % r) * s)
    g = graph_from_pairs(r * s, np.array(a), np.array(b))
    lab, Q = check(ctx, g, "ring of cliques")
    e, m, d = s * (s - 1) / 2, r * s * (s - 1) / 2 + r, s * (s - 1) + 2
    assert Q == pytest.approx(r * (e / m - (d / (2 * m)) ** 2), abs=1e-12)


def test_group_above_25000_nodes(ctx):
    say("modularity N=26000")
    check(ctx, planted(26000, 4, 10, 1, 5), "large")


def test_device_csr_from_a_threshold_build(ctx):
    say("modularity on a device CSR (N=5000, L=200, r > 0.5)")
    x = bna.synth_factor_model(5000, 200, 8, 0.7, 77, bna.TIME_MAJOR)
    with ctx.series(x) as s:
        gd = s.build_device(thr=0.5, weighted=False)
        try:
            host = s.to_host(gd)
            lab, Q, mods = ctx.modularity(gd)
        finally:
            s.free_device(gd)
    olab, oQ, omods, rounds, its = orc.modularity(host.R, host.C, host.N)
    say(f"  device CSR: nnz={host.nnz} modules {mods} (oracle {omods}) Q={Q:.12f} (oracle {oQ:.12f})")
    assert mods == omods and np.array_equal(lab, olab)
    assert Q == pytest.approx(oQ, abs=1e-12)


def test_errors(ctx):
    g = bna.Graph.__new__(bna.Graph)
    g.N, g.nnz = 3, 0
    g.R = np.zeros(4, np.int64)
    g.C = np.zeros(0, np.int32)
    g.V = None
    with pytest.raises(bna.BnaError):
        ctx.modularity(g)
