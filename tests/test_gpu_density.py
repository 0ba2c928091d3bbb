"""GPU parity of the density threshold (exactly k edges) against the FP64 oracle.

The oracle keeps every pair with r64 > r*, r* the k-th largest r64, and of the
pairs tied at r* the first ones in row-major upper-triangle order
(FormFullGraph_s, W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:743-822; the oracle's
literal stack loop and its closed form are pinned against each other in
test_oracle.py).  Bar: R and C bit-exact, r* equal to the oracle's, V within the
reported band of r64 (candidates carry (float) r64 exactly).
"""
import hashlib
import json
import sys
from pathlib import Path

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


def compare(g, og):
    assert g.nnz == og.nnz, f"nnz {g.nnz} != oracle {og.nnz}"
    assert np.array_equal(g.R, og.R)
    assert np.array_equal(g.C, og.C)
    if g.V is not None:
        d = np.abs(g.V.astype(np.float64) - og.V64)
        assert d.max(initial=0.0) <= V_TOL, f"max |V - r64| = {d.max()}"


@pytest.mark.parametrize("prec", [bna.FP32, bna.BF16, bna.BF16X3])
def test_density_c1_shape(ctx, prec):
    # C1 shape with the density threshold: N = 5000, L = 200, K = 8, phi = 0.7
    N, L, spa = 5000, 200, 1.0
    say(f"density C1 prec={prec}")
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A0000 + 1000)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, spa)
    og = orc.build_density(z64, valid, k)
    g = ctx.build_csr(x, density=spa, precision=prec)
    say(f"  k={k} nnz={g.nnz} cand={g.n_candidates} r*={g.thr_used:.9g} oracle r*={og.r_star:.9g}")
    compare(g, og)
    assert g.nnz == 2 * k
    assert g.thr_used == og.r_star
    assert abs(g.density_pct - 100.0 * k / (N * (N - 1) / 2)) < 1e-9


@pytest.mark.parametrize("N,L,spa", [(1000, 61, 5.0), (777, 150, 0.37), (2049, 96, 2.0), (300, 40, 60.0)])
def test_density_ragged(ctx, N, L, spa):
    say(f"density ragged N={N} L={L} spa={spa}")
    x = bna.synth_factor_model(N, L, 6, 0.5, 77 + N)
    x[:, 3] = 1.5  # a constant voxel: never an edge, not counted as a pair
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, spa)
    og = orc.build_density(z64, valid, k)
    g = ctx.build_csr(x, density=spa, precision="bf16x3")
    compare(g, og)
    assert g.thr_used == og.r_star
    assert g.R[4] == g.R[3]


def test_density_ties(ctx):
    # Exact copies of a voxel give pairs whose r64 is the same double, so the
    # density cut lands inside groups of ties and the row-major tie rule decides.
    rng = np.random.default_rng(5)
    base = bna.synth_factor_model(120, 80, 4, 0.6, 9)
    cols = rng.integers(0, 120, size=900)
    x = np.ascontiguousarray(base[:, cols])
    N = x.shape[1]
    z64, _, _, _, valid = orc.znorm(x, 0)
    packed = orc.all_pairs(z64, valid)
    ones = packed[np.isfinite(packed)]
    top = np.sort(ones)[::-1]
    say(f"density ties: N={N}, distinct top values {len(np.unique(top[:3000]))} of 3000")
    for spa in (0.3, 0.71, 1.3, 4.0):
        k = orc.density_k(N, spa)
        og = orc.build_density(z64, valid, k)
        say(f"  spa={spa} k={k} oracle ties at r*={og.n_ties}")
        for prec in ("bf16x3", "fp32"):
            g = ctx.build_csr(x, density=spa, precision=prec)
            compare(g, og)
            assert g.thr_used == og.r_star


def test_density_binary_and_no_recheck(ctx):
    N, L, spa = 1500, 120, 1.0
    x = bna.synth_factor_model(N, L, 8, 0.7, 31)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, spa)
    og = orc.build_density(z64, valid, k)
    g = ctx.build_csr(x, density=spa, precision="bf16x3", weighted=False)
    assert g.V is None
    assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C)
    # without the recheck the selection runs on the MFMA r: still exactly k
    # edges, and every chosen pair is within the band of the oracle's choice
    g2 = ctx.build_csr(x, density=spa, precision="bf16x3", recheck=False)
    assert g2.nnz == 2 * k
    rows = np.repeat(np.arange(N), np.diff(g2.R))
    r64 = np.array([orc.dot64(z64[i], z64[j]) for i, j in zip(rows[:2000], g2.C[:2000])])
    assert np.all(r64 >= og.r_star - 5e-4)  # a few times the split-BF16 band at L = 120


def test_density_errors(ctx):
    N, L = 400, 50
    x = bna.synth_factor_model(N, L, 4, 0.5, 3)
    with pytest.raises(bna.BnaError):
        ctx.build_csr(x, density=0.0)  # k = 0
    x[:, :390] = 2.0  # 10 valid voxels: 45 valid pairs
    with pytest.raises(bna.BnaError):
        ctx.build_csr(x, density=50.0)


def test_density_c2_full_size(ctx):
    # C2 shape: N = 45000, L = 240, K = 20, phi = 0.6, 1 % density: exactly
    # k = 10,124,775 edges, split-BF16.  The CSR must be the FP64 oracle's bit
    # for bit: its r* and the digests of its R and C are in
    # golden/c2_density.json (make_c2_golden.py).
    gold = json.loads((Path(__file__).resolve().parent / "golden" / "c2_density.json").read_text())
    N, L, spa = gold["N"], gold["L"], gold["density_pct"]
    say("density C2 full size")
    x = bna.synth_factor_model(N, L, gold["K"], gold["phi"], gold["seed"])
    k = orc.density_k(N, spa)
    assert k == gold["k"] == 10124775
    with ctx.series(x) as s:
        g = s.build(density=spa, precision="bf16x3")
        t = ctx.timings()
    say(f"  nnz={g.nnz} r*={g.thr_used:.9g} cand={g.n_candidates} corr_ms={t['corr_ms']:.2f} "
        f"select_ms={t['select_ms']:.2f} total_ms={t['total_ms']:.2f}")
    assert g.nnz == 2 * k
    assert g.thr_used == gold["r_star"]
    assert hashlib.sha256(g.R.astype(np.int64).tobytes()).hexdigest() == gold["sha256_R_int64"]
    assert hashlib.sha256(g.C.astype(np.int32).tobytes()).hexdigest() == gold["sha256_C_int32"]
    assert np.all(g.V.astype(np.float64) >= g.thr_used - 1e-6)
