"""GPU parity: libbna on the MI355X against the CPU oracle, through the C ABI.

Bar: edge sets and R/C bit-exact against the FP64 oracle; z64, mean and norm
bit-exact; weighted values within 1e-6 absolute of r64 (the library recomputes
each kept value in double once the edge set is fixed; without that refinement
they stay within the error band it reports); degree bit-exact; eigenvector
centrality within 1e-5 relative L2.
"""
import os
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

pytestmark = pytest.mark.gpu

# Weighted values: |V - r64| <= 1e-6 absolute against the FP64 oracle.
V_TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


def check_csr_invariants(g):
    N = g.N
    assert g.R[0] == 0 and g.R[-1] == g.nnz and np.all(np.diff(g.R) >= 0)
    rows = np.repeat(np.arange(N), np.diff(g.R))
    assert np.all(g.C != rows), "diagonal entry"
    for i in range(N):
        seg = g.C[g.R[i]:g.R[i + 1]]
        assert np.all(np.diff(seg) > 0), f"row {i} not strictly ascending"
    # symmetric: the (row, col) set equals its transpose
    a = rows.astype(np.int64) * N + g.C
    b = g.C.astype(np.int64) * N + rows
    assert np.array_equal(np.sort(a), np.sort(b))


def compare(g, og, weighted=True):
    assert g.nnz == og.nnz, f"nnz {g.nnz} != oracle {og.nnz}"
    assert np.array_equal(g.R, og.R)
    assert np.array_equal(g.C, og.C)
    if weighted:
        d = np.abs(g.V.astype(np.float64) - og.V64)
        assert d.max(initial=0.0) <= V_TOL, f"max |V - r64| = {d.max()}"
        return float(d.max(initial=0.0))
    assert g.V is None
    return 0.0


def series_and_oracle(N, L, seed, K=8, phi=0.7, layout=bna.TIME_MAJOR):
    x = bna.synth_factor_model(N, L, K, phi, seed, layout)
    z64, z32, mu, sd, valid = orc.znorm(x, layout)
    return x, z64, valid


@pytest.mark.parametrize("layout", [bna.TIME_MAJOR, bna.VOXEL_MAJOR])
@pytest.mark.parametrize("prec", [bna.FP32, bna.BF16, bna.BF16X3])
def test_znorm_bit_exact(ctx, layout, prec):
    say(f"znorm layout={layout} prec={prec}")
    x = bna.synth_factor_model(700, 203, 8, 0.7, 21, layout)
    x[:, 5 if layout else 0] = 0  # one odd column in either layout
    z64, z32, mu, sd, valid = orc.znorm(x, layout)
    with ctx.series(x, layout) as s:
        s.normalize(prec)
        gz, gmu, gsd = s.download()
    assert np.array_equal(gmu, mu)
    assert np.array_equal(gsd, sd)
    m = valid.astype(bool)
    assert np.array_equal(gz[m], z64[m])


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16x3"])
def test_threshold_c1_parity(ctx, prec):
    # configuration C1: N = 5000, L = 200, K = 8, phi = 0.7, r > 0.5
    say(f"C1 threshold prec={prec}")
    x, z64, valid = series_and_oracle(5000, 200, 0xB17A0000 + 1000)
    og = orc.build_threshold(z64, valid, 0.5)
    g = ctx.build_csr(x, thr=0.5, precision=prec)
    check_csr_invariants(g)
    err = compare(g, og)
    say(f"  nnz={g.nnz} cand={g.n_candidates} flips={g.n_flips} band={g.band:.3g} "
        f"max|V-r64|={err:.3g} max|r_mfma-r64| over band pairs={g.max_err:.3g}")
    assert g.max_err <= g.band


@pytest.mark.parametrize("N,L", [(2, 2), (3, 5), (255, 17), (256, 32), (257, 33), (600, 64), (1100, 97)])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_threshold_ragged_shapes(ctx, N, L, prec):
    say(f"ragged N={N} L={L} prec={prec}")
    rng = np.random.default_rng(N * 1000 + L)
    x = bna.synth_factor_model(N, L, 4, 0.5, N + L)
    thr = 0.2 if L > 8 else 0.5
    z64, _, _, _, valid = orc.znorm(x, 0)
    og = orc.build_threshold(z64, valid, thr)
    g = ctx.build_csr(x, thr=thr, precision=prec)
    check_csr_invariants(g)
    compare(g, og)


def test_invalid_voxels_and_negative_threshold(ctx):
    say("invalid voxels")
    x = bna.synth_factor_model(400, 50, 4, 0.5, 77)
    x[:, 10] = 3.0        # constant
    x[:, 200] = 0.0       # constant zero
    x[7, 300] = np.nan    # non-finite
    z64, _, _, _, valid = orc.znorm(x, 0)
    assert valid.sum() == 397
    for thr in (-0.5, 0.1):
        og = orc.build_threshold(z64, valid, thr)
        for prec in ("fp32", "bf16", "bf16x3"):
            g = ctx.build_csr(x, thr=thr, precision=prec)
            compare(g, og)
            for v in (10, 200, 300):
                assert g.R[v + 1] == g.R[v]


def test_empty_graph_and_binary(ctx):
    say("empty + binary")
    x = bna.synth_factor_model(1000, 100, 8, 0.7, 5)
    g = ctx.build_csr(x, thr=0.999)
    assert g.nnz == 0 and np.all(g.R == 0)
    z64, _, _, _, valid = orc.znorm(x, 0)
    og = orc.build_threshold(z64, valid, 0.3)
    gb = ctx.build_csr(x, thr=0.3, weighted=False)
    compare(gb, og, weighted=False)


def test_layouts_agree_and_compat_eps(ctx):
    say("layouts + eps")
    xt = bna.synth_factor_model(900, 120, 8, 0.7, 9, bna.TIME_MAJOR)
    xv = np.ascontiguousarray(xt.T)
    a = ctx.build_csr(xt, bna.TIME_MAJOR, thr=0.35)
    b = ctx.build_csr(xv, bna.VOXEL_MAJOR, thr=0.35)
    assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C) and np.array_equal(a.V, b.V)
    z64, _, _, _, valid = orc.znorm(xt, 0)
    og = orc.build_threshold(z64, valid, 0.35 - 1e-6)
    c = ctx.build_csr(xt, thr=0.35, ref_compat_eps=True)
    compare(c, og)


def test_recheck_off_reports_flips(ctx):
    say("recheck off")
    x = bna.synth_factor_model(3000, 150, 8, 0.7, 13)
    z64, _, _, _, valid = orc.znorm(x, 0)
    og = orc.build_threshold(z64, valid, 0.4)
    g = ctx.build_csr(x, thr=0.4, precision="bf16", recheck=False)
    assert g.n_candidates == 0 and g.band == 0.0
    # without the recheck, plain BF16 decides by its own value: edge sets may
    # differ from the oracle only for pairs within the BF16 error band
    A = set(zip(np.repeat(np.arange(3000), np.diff(og.R)).tolist(), og.C.tolist()))
    B = set(zip(np.repeat(np.arange(3000), np.diff(g.R)).tolist(), g.C.tolist()))
    diff = A ^ B
    for (i, j) in diff:
        r = orc.dot64(z64[i], z64[j])
        assert abs(r - 0.4) < 2e-2


def test_refine_off_values_within_band(ctx):
    say("refine off")
    x, z64, valid = series_and_oracle(2500, 300, 17)
    og = orc.build_threshold(z64, valid, 0.42)
    for prec in ("bf16x3", "fp32"):
        g = ctx.build_csr(x, thr=0.42, precision=prec, refine=False)
        assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C)
        d = np.abs(g.V.astype(np.float64) - og.V64)
        say(f"  {prec}: MFMA values, max |V - r64| = {d.max():.3g}, band {g.band:.3g}")
        assert d.max() <= g.band


def test_degree_and_ec_parity(ctx):
    say("degree + EC")
    x, z64, valid = series_and_oracle(3000, 200, 31)
    og = orc.build_threshold(z64, valid, 0.45)
    g = ctx.build_csr(x, thr=0.45)
    compare(g, og)
    deg = ctx.degree(g)
    assert np.array_equal(deg, orc.degree(og.R, g.V, 3000))
    gb = ctx.build_csr(x, thr=0.45, weighted=False)
    assert np.array_equal(ctx.degree(gb), orc.degree(og.R, None, 3000))
    for out_mode, binary in [(bna.EC_WEIGHTED_REF, False), (bna.EC_BINARY_REF, True)]:
        ec, lam, it = ctx.eigcent(g, output=out_mode, binary=binary)
        oec, olam, oit = orc.eigcent(og.R, og.C, None if binary else g.V, 3000, out_mode=out_mode)
        rel = np.linalg.norm(ec - oec) / np.linalg.norm(oec)
        say(f"  EC mode={out_mode} iters={it}/{oit} lambda={lam:.9g}/{olam:.9g} relL2={rel:.3g}")
        assert rel < 1e-5
        assert lam == pytest.approx(olam, rel=1e-9)
        assert abs(it - oit) <= 1


def ragged_graph(N, m, seed, hub_deg):
    """Symmetric graph with positive weights: m random edges, one hub joined to
    hub_deg voxels spread over every column range, a run of empty rows, and
    the columns of every row shuffled (a host CSR may list them in any order)."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, N, m)
    j = rng.integers(0, N, m)
    hub = N // 3
    hj = rng.choice(N, hub_deg, replace=False)
    i = np.concatenate([i, np.full(hub_deg, hub)])
    j = np.concatenate([j, hj])
    empty = (i >= 100) & (i < 140) | (j >= 100) & (j < 140)
    keep = (i != j) & ~empty
    a, b = np.minimum(i[keep], j[keep]), np.maximum(i[keep], j[keep])
    key = np.unique(a.astype(np.int64) * N + b)
    a, b = key // N, key % N
    w = rng.uniform(0.05, 1.0, len(a)).astype(np.float32)
    rows = np.concatenate([a, b])
    cols = np.concatenate([b, a]).astype(np.int32)
    vals = np.concatenate([w, w])
    order = np.lexsort((rng.random(len(rows)), rows))  # by row, columns shuffled
    g = bna.Graph.__new__(bna.Graph)
    g.N, g.nnz = N, len(rows)
    g.R = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=N))]).astype(np.int64)
    g.C = cols[order]
    g.V = vals[order]
    return g


@pytest.mark.parametrize("N,m", [(6000, 60000), (24000, 260000)])
def test_ec_blocked_layout_ragged(ctx, N, m):
    # 20 or more entries per row on average: bna_eigcent copies the graph
    # into its segment-blocked layout (one column segment at N = 6000, two at
    # N = 24000); empty rows, a hub row across all segments and unsorted
    # columns must give the oracle's vector in both output modes
    g = ragged_graph(N, m, 7 + N, hub_deg=N // 2)
    assert g.nnz >= 16 * N
    say(f"EC blocked layout: N={N} nnz={g.nnz} max row {np.diff(g.R).max()}")
    for out_mode, binary in [(bna.EC_WEIGHTED_REF, False), (bna.EC_BINARY_REF, True)]:
        ec, lam, it = ctx.eigcent(g, output=out_mode, binary=binary)
        oec, olam, oit = orc.eigcent(g.R, g.C, None if binary else g.V, N, out_mode=out_mode)
        rel = np.linalg.norm(ec - oec) / np.linalg.norm(oec)
        say(f"  mode={out_mode} iters={it}/{oit} lambda={lam:.12g}/{olam:.12g} relL2={rel:.3g}")
        assert rel < 1e-9
        assert lam == pytest.approx(olam, rel=1e-12)
        assert it == oit


def test_ec_graph_replay_matches_launches(ctx, monkeypatch):
    """The power iteration replays captured 16-iteration graphs; with graphs
    off (BNA_GRAPHS=0 at context creation) every kernel is launched directly.
    Both must give the same bits, call after call."""
    say("EC graph replay vs direct launches")
    x, z64, valid = series_and_oracle(3000, 200, 31)
    g = ctx.build_csr(x, thr=0.45)
    runs = [ctx.eigcent(g, tol=1e-12) for _ in range(2)]
    monkeypatch.setenv("BNA_GRAPHS", "0")
    with bna.Context(0) as c2:
        direct = c2.eigcent(g, tol=1e-12)
    assert direct[2] > 32  # several whole batches were replayed
    for ec, lam, it in runs:
        assert it == direct[2]
        assert lam == direct[1]
        assert np.array_equal(ec, direct[0])


def test_ec_known_answer_4x4(ctx):
    # B/src/cuda_EC/cuda_EC/ecGPU.cu:217-233
    g = bna.Graph.__new__(bna.Graph)
    g.N, g.nnz = 4, 8
    g.R = np.array([0, 2, 3, 6, 8])
    g.C = np.array([2, 3, 2, 0, 1, 3, 0, 2], np.int32)
    g.V = None
    y, lam, it = ctx.eigcent(g, output=bna.EC_BINARY_REF, binary=True)
    np.testing.assert_allclose(y, [0.5227207, 0.2818452, 0.6116285, 0.5227207], atol=2e-6)
    assert lam == pytest.approx(2.170086486626, abs=1e-5)
    assert it == 32


@pytest.mark.parametrize("N,L,layout", [(900, 97, bna.VOXEL_MAJOR), (600, 2100, bna.TIME_MAJOR)])
def test_refine_general_forms(ctx, N, L, layout):
    # rows that are not 16-byte aligned (voxel-major input, L % 4 != 0) and rows
    # longer than the register-resident form holds (L > 2048) take the general
    # refinement kernel; its values and mirror entries meet the same bar
    say(f"refine general form N={N} L={L} layout={layout}")
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A0000 + 7000 + L, layout)
    z64, _, _, _, valid = orc.znorm(x, layout)
    og = orc.build_threshold(z64, valid, 0.3)
    g = ctx.build_csr(x, layout, thr=0.3)
    err = compare(g, og)
    # refined values are the double sums rounded to float: half a float ulp of r
    assert err <= 6e-8, f"max |V - r64| = {err}"


def test_build_into_caller_buffers(ctx):
    # the caller-allocated form: a size query, a too-small buffer, then the copy
    N, L = 1500, 120
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A0000 + 7100)
    with ctx.series(x) as s:
        ref = s.build(thr=0.4)
        nnz = s.build_into(None, None, None, thr=0.4)
        assert nnz == ref.nnz
        R = np.empty(N + 1, np.int32)
        small = np.empty(max(nnz - 1, 0), np.int32)
        with pytest.raises(bna.BnaError):
            s.build_into(R, small, np.empty(small.size, np.float32), thr=0.4)
        Cc = np.empty(nnz, np.int32)
        V = np.empty(nnz, np.float32)
        assert s.build_into(R, Cc, V, thr=0.4) == nnz
        assert np.array_equal(R.astype(np.int64), ref.R) and np.array_equal(Cc, ref.C) and np.array_equal(V, ref.V)
        k = bna.density_to_k(N, 2.0)
        Cb = np.empty(2 * k, np.int32)
        assert s.build_into(R, Cb, None, density=2.0, weighted=False) == 2 * k
