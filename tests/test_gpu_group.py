"""GPU parity of group averaging (avg = y / b) against the oracle.

The oracle (test_group_oracle.py pins it) adds each subject's r64, taken as
float, or its Fisher z when it lies in [-1, 1], divides by the subject count in
float and maps Fisher sums back (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:36-50,
548-558, 824-887); the average is thresholded by FormFullGraph / FormFullGraph_s
(:683-822) and listed by FormCSRGraph (:889-918).

The device sums add the tile kernel's r, so before a graph is built each
average differs from the oracle's by the MFMA error of its terms (bars AVG_TOL
for plain means, AVG_TOL_Z for Fisher means, whose slope 1 / (1 - r^2)
magnifies it).  An exact group (the default) then recomputes, from FP64 r in the
reference's float arithmetic, every pair the error bound leaves on either side
of the threshold and every kept edge: edge sets, tie rule included, equal the
oracle's, and the values are the oracle's bit for bit for plain means.  Fisher
means take logf / expf on the device and in the C library, which may differ by
an ulp, so their values are compared at FISHER_ULP_TOL and an edge may differ
only for a pair that close to the threshold.
"""
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

pytestmark = pytest.mark.gpu

AVG_TOL = 4e-6
AVG_TOL_Z = 2e-5
FISHER_ULP_TOL = 4 * 2.0 ** -24  # a few ulp of a value below 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


def subjects(N, Ls, seed, dead=None):
    xs = []
    for q, L in enumerate(Ls):
        x = bna.synth_factor_model(N, L, 8, 0.7, seed + q)
        if dead is not None and dead.get(q) is not None:
            x[:, dead[q]] = 3.0  # constant series: invalid voxels, NaN r
        xs.append(x)
    return xs


def oracle_packed(x):
    z64, _, _, _, valid = orc.znorm(x, 0)
    return orc.all_pairs(z64, valid)


def gpu_group(ctx, xs, N, fisher, prec):
    g = ctx.group(N, fisher)
    for x in xs:
        with ctx.series(x) as s:
            g.add(s, prec)
    return g


def edges(G):
    rows = np.repeat(np.arange(G.N), np.diff(G.R))
    keep = rows < G.C
    return set(zip(rows[keep].tolist(), G.C[keep].tolist()))


def check_graph(G, og, avg_ref, N, thr, tol):
    """Edge sets equal except pairs within tol of thr (by the oracle's average)."""
    assert np.all(np.diff(G.R) >= 0) and G.R[-1] == G.nnz
    e_gpu, e_orc = edges(G), edges(og)
    diff = e_gpu ^ e_orc
    for (i, j) in diff:
        p = i * N - (i + 1) * i // 2 + j - i - 1
        assert abs(float(avg_ref[p]) - thr) <= tol, f"pair ({i},{j}) avg {avg_ref[p]} thr {thr}"
    return len(diff)


def kept_pairs(G, N):
    rows = np.repeat(np.arange(G.N), np.diff(G.R))
    lo, hi = np.minimum(rows, G.C), np.maximum(rows, G.C)
    return lo * N - (lo + 1) * lo // 2 + hi - lo - 1


def exact_graph_check(G, og, ref, N, thr, fisher):
    """Edges equal the oracle's (Fisher: except pairs within FISHER_ULP_TOL of
    thr); values equal the oracle's averages (Fisher: within FISHER_ULP_TOL)."""
    e_gpu, e_orc = edges(G), edges(og)
    diff = e_gpu ^ e_orc
    if fisher:
        for (i, j) in diff:
            p = i * N - (i + 1) * i // 2 + j - i - 1
            assert abs(float(ref[p]) - thr) <= FISHER_ULP_TOL, f"pair ({i},{j}) avg {ref[p]} thr {thr}"
    else:
        assert not diff, f"{len(diff)} edges differ from the oracle's"
    if G.V is not None and G.nnz:
        p = kept_pairs(G, N)
        if fisher:
            assert np.max(np.abs(G.V.astype(np.float64) - ref[p])) <= FISHER_ULP_TOL
        else:
            assert np.array_equal(G.V, ref[p])
    return len(diff)


@pytest.mark.parametrize("fisher", [False, True])
@pytest.mark.parametrize("prec", ["bf16x3", "fp32", "bf16"])
def test_group_average_and_threshold(ctx, fisher, prec):
    N, Ls = 700, [120, 200, 333]
    say(f"group N={N} L={Ls} fisher={fisher} prec={prec}")
    xs = subjects(N, Ls, 0xB17A7000, dead={1: [5, 300]})
    ref = orc.group_average([oracle_packed(x) for x in xs], N, fisher)
    g = gpu_group(ctx, xs, N, fisher, prec)
    try:
        assert g.count == 3
        g.finish()
        avg = g.packed()
        tol = (0.1 if fisher else 0.02) if prec == "bf16" else (AVG_TOL_Z if fisher else AVG_TOL)
        err = np.abs(avg.astype(np.float64) - ref)
        say(f"  device sums: max |avg - oracle| = {err.max():.3g}")
        assert err.max() <= tol
        # voxel 5 is invalid in subject 1 only: its pairs average the other two
        assert np.all(np.isfinite(avg))
        thr = 0.25
        G = g.build(thr)
        og = orc.threshold_packed(ref, N, thr)
        nd = exact_graph_check(G, og, ref, N, thr, fisher)
        say(f"  r > {thr}: nnz {G.nnz} oracle {og.nnz}, recomputed at the threshold {g.band_pairs}, "
            f"differences {nd}")
        assert G.thr_used == pytest.approx(thr)
        # density: exactly k edges, the oracle's r* and tie rule
        spa = 3.0
        k = orc.density_k(N, spa)
        Gd = g.build(density=spa, weighted=False)
        assert Gd.nnz == 2 * k and Gd.V is None
        okd = orc.select_packed(ref.astype(np.float64), N, k, literal=False)
        if fisher:
            assert abs(Gd.thr_used - okd.r_star) <= FISHER_ULP_TOL
        else:
            assert Gd.thr_used == okd.r_star
            assert np.array_equal(Gd.R, okd.R) and np.array_equal(Gd.C, okd.C)
        nd = exact_graph_check(Gd, okd, ref, N, okd.r_star, fisher)
        say(f"  density {spa}%: k={k} r*={Gd.thr_used:.9g} oracle {okd.r_star:.9g}, "
            f"recomputed {g.band_pairs}, differences {nd}")
        # the weighted density graph: every kept value is the oracle's average
        Gw = g.build(density=spa)
        exact_graph_check(Gw, okd, ref, N, okd.r_star, fisher)
    finally:
        g.close()


def test_group_inexact_option(ctx):
    # exact = off: the MFMA sums thresholded as they are, within their bound
    N, Ls = 500, [100, 160]
    xs = subjects(N, Ls, 0xB17A7050)
    ref = orc.group_average([oracle_packed(x) for x in xs], N, False)
    g = ctx.group(N, False, exact=False)
    try:
        for x in xs:
            with ctx.series(x) as s:
                g.add(s, "bf16x3")
        g.finish()
        G = g.build(0.25)
        assert g.band_pairs == 0
        og = orc.threshold_packed(ref, N, 0.25)
        assert check_graph(G, og, ref, N, 0.25, AVG_TOL) <= 4
    finally:
        g.close()


def test_group_ties_exact(ctx):
    # Half the voxels are constant in every subject, so their pairs sum nothing
    # and average exactly 0; the rest share one strong factor (r ~ 0.8).  A 40 %
    # density takes every valid pair and the first m zero pairs in row-major
    # order (FormFullGraph_s's tie rule), which the oracle fixes exactly.
    N, L, S = 600, 150, 2
    rng = np.random.default_rng(11)
    dead = rng.choice(N, N // 2, replace=False)
    xs = []
    for q in range(S):
        f = rng.normal(size=(L, 1)).astype(np.float32)
        x = (f + 0.5 * rng.normal(size=(L, N))).astype(np.float32)
        x[:, dead] = 2.0
        xs.append(np.ascontiguousarray(x))
    ref = orc.group_average([oracle_packed(x) for x in xs], N, False)
    spa = 40.0
    k = orc.density_k(N, spa)
    g = gpu_group(ctx, xs, N, False, "bf16x3")
    try:
        g.finish()
        avg = g.packed()
        assert np.array_equal(avg == 0, ref == 0)
        okd = orc.select_packed(ref.astype(np.float64), N, k, literal=False)
        assert okd.r_star == 0.0 and okd.n_ties > 0
        G = g.build(density=spa)
        say(f"  ties: k={k} nnz={G.nnz} oracle {okd.nnz}, ties {okd.n_ties}")
        assert G.nnz == okd.nnz == 2 * k
        assert np.array_equal(G.R, okd.R)
        assert np.array_equal(G.C, okd.C)
        assert G.thr_used == 0.0
    finally:
        g.close()


def test_single_subject_cormat_is_r_or_zero(ctx):
    # one subject, no Fisher: the packed matrix post_block leaves for the
    # individual .cormat (CorMat.cpp:301-310): r where valid, 0 for invalid pairs
    N = 300
    x = subjects(N, [90], 0xB17A7100, dead={0: [7]})[0]
    ref = oracle_packed(x)
    g = gpu_group(ctx, [x], N, False, "bf16x3")
    try:
        g.finish()
        avg = g.packed()
        exp = np.where(np.isnan(ref), 0.0, ref)
        assert np.max(np.abs(avg - exp)) <= AVG_TOL
        assert np.all(avg[np.isnan(ref)] == 0.0)
    finally:
        g.close()


def test_group_errors(ctx):
    x = subjects(200, [64], 0xB17A7200)[0]
    g = ctx.group(200)
    try:
        with pytest.raises(bna.BnaError):
            g.finish()  # nothing added
        with ctx.series(subjects(150, [64], 1)[0]) as s:
            with pytest.raises(bna.BnaError):
                g.add(s)  # wrong N
        with ctx.series(x) as s:
            g.add(s)
            g.finish()
            with pytest.raises(bna.BnaError):
                g.add(s)  # already averaged
        with pytest.raises(bna.BnaError):
            g.build(density=0.0)
    finally:
        g.close()


def test_corr_packed_and_select_density(ctx):
    # bna_corr_packed: CorMat_gpu + post_block's packed matrix; bna_select_density:
    # CorMat_spa2rth's r*, here exact (the k-th largest r64)
    N = 900
    x = subjects(N, [140], 0xB17A7300, dead={0: [11]})[0]
    ref = oracle_packed(x)
    with ctx.series(x) as s:
        p = s.corr_packed("bf16x3")
        exp = np.where(np.isnan(ref), 0.0, ref)
        assert np.max(np.abs(p - exp)) <= AVG_TOL
        r_star, k = s.select_density(2.5)
    z64, _, _, _, valid = orc.znorm(x, 0)
    kk = orc.density_k(N, np.float32(2.5))
    og = orc.build_density(z64, valid, kk)
    assert k == kk
    assert r_star == og.r_star


@pytest.mark.parametrize("N", [2, 50, 257])
def test_group_small_and_ragged(ctx, N):
    # one diagonal tile, one partial keep block, and a tile edge just past 256
    say(f"group small N={N}")
    xs = subjects(N, [40, 33], 0xB17A7400 + N)
    ref = orc.group_average([oracle_packed(x) for x in xs], N, True)
    g = gpu_group(ctx, xs, N, True, "bf16x3")
    try:
        g.finish()
        avg = g.packed()
        assert avg.shape == ref.shape and np.max(np.abs(avg - ref), initial=0.0) <= AVG_TOL_Z
        G = g.build(-2.0, weighted=False)  # every pair
        assert G.nnz == N * (N - 1) and G.V is None
        assert np.array_equal(G.R, np.arange(N + 1) * (N - 1))
        cols = np.concatenate([np.delete(np.arange(N), i) for i in range(N)]) if N > 1 else np.zeros(0)
        assert np.array_equal(G.C, cols)
    finally:
        g.close()
