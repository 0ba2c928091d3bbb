"""Autotuner of the tile traversal (band height of the tile table) and its cache
file; the CSR must not depend on the traversal."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import cuda_newest20160331_amd as bna  # noqa: E402
import orc  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    with bna.Context(0) as c:
        yield c


def test_autotune_picks_a_candidate_and_persists(ctx, tmp_path):
    path = str(tmp_path / "bands.txt")
    say("autotune bf16x3 L=240 and fp32 L=1200")
    g1, ms1 = ctx.autotune("bf16x3", 240, N=4096, cache=path)
    g2, ms2 = ctx.autotune("fp32", 1200, N=4096, cache=path)
    assert g1 in (4, 8, 16, 32) and g2 in (4, 8, 16, 32)
    assert ms1 > 0 and ms2 > 0
    assert ctx.band_rows("bf16x3", 250) == g1  # same bucket as 240
    lines = open(path).read().split("\n")
    assert len([ln for ln in lines if ln.strip()]) == 2
    with bna.Context(0) as c2:
        assert c2.band_rows("bf16x3", 240) == 8  # default before loading
        c2.autotune_load(path)
        assert c2.band_rows("bf16x3", 240) == g1
        assert c2.band_rows("fp32", 1200) == g2


def test_csr_independent_of_band_height(ctx):
    N, L, thr = 2600, 200, 0.45
    x = bna.synth_factor_model(N, L, 8, 0.7, 77)
    z64, _, _, _, valid = orc.znorm(x, 0)
    og = orc.build_threshold(z64, valid, thr)
    ref = None
    for rows in (1, 3, 8, 32):
        say(f"band height {rows}")
        ctx.set_band_rows("bf16x3", L, rows)
        g = ctx.build_csr(x, thr=thr, precision="bf16x3")
        assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C)
        if ref is None:
            ref = g
        else:
            assert np.array_equal(g.V, ref.V)
    ctx.set_band_rows("bf16x3", L, 8)


def test_bad_cache_file_is_an_error(ctx, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("2 9 8 1.0\n")
    with pytest.raises(bna.BnaError):
        ctx.autotune_load(str(p))
    with pytest.raises(bna.BnaError):
        ctx.autotune_load(str(tmp_path / "missing.txt"))
