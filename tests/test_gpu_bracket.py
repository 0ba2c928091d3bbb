"""GPU parity of the sampled density bracket.

A density build of a subject with at least 2^26 pairs (N >= 11586) brackets r*
from a histogram of an eighth of its tiles instead of all of them.  After the
emission pass, the FP64 recheck and the exact selection, the build checks
s_lo + band <= r* < s_hi - band (every dropped pair then lies below r* and every
sure pair above it, so the selection is the exact top k) and runs both passes
again in full when the sample missed.  These tests take the sampled path on
small subjects (BNA_BRACKET_SAMPLE_MIN=0), force misses on both sides
(BNA_BRACKET_SAMPLE_SHIFT moves the sampled bracket by that many histogram
bins), and require the CSR to equal the full-pass build bit for bit and the FP64
oracle's edge set (FormFullGraph_s, W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:743-822).
"""
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc
from test_gpu_density import compare

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build(monkeypatch, env, x, **kw):
    """A build in a context created under env (the knobs are read at creation)."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    try:
        with bna.Context(0) as c:
            g = c.build_csr(x, **kw)
            t = c.timings()
    finally:
        for k in env:
            monkeypatch.delenv(k)
    return g, t


@pytest.fixture(scope="module")
def subject():
    N, L, spa = 4000, 200, 1.0
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB4AC)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, spa)
    return x, spa, k, orc.build_density(z64, valid, k)


def same(a, b, thr=True):
    assert a.nnz == b.nnz
    assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C)
    assert (a.V is None) == (b.V is None)
    if a.V is not None:
        assert np.array_equal(a.V, b.V)
    if thr:
        assert a.thr_used == b.thr_used


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_sampled_bracket_equals_full(monkeypatch, subject, prec):
    x, spa, k, og = subject
    say(f"sampled bracket prec={prec}")
    full, tf = build(monkeypatch, {"BNA_BRACKET_SAMPLE": "0"}, x, density=spa, precision=prec)
    samp, ts = build(monkeypatch, {"BNA_BRACKET_SAMPLE_MIN": "0"}, x, density=spa, precision=prec)
    say(f"  k={k} launches full {tf['corr_launches']} sampled {ts['corr_launches']} "
        f"cand full {full.n_candidates} sampled {samp.n_candidates}")
    assert ts["corr_launches"] == 2  # a sampled histogram and one emission: the bracket held
    compare(samp, og)
    assert samp.nnz == 2 * k and samp.thr_used == og.r_star
    same(samp, full)


@pytest.mark.parametrize("shift", [300, -300])
def test_sampled_bracket_miss_falls_back(monkeypatch, subject, shift):
    # +300 bins puts s_lo above r* (some edges never emitted: the selection
    # rank exceeds the candidates); -300 puts s_hi below it (too many sure
    # edges).  Either way the build must notice and redo both passes in full.
    x, spa, k, og = subject
    say(f"sampled bracket forced miss shift={shift}")
    full, _ = build(monkeypatch, {"BNA_BRACKET_SAMPLE": "0"}, x, density=spa, precision="bf16x3")
    g, t = build(monkeypatch, {"BNA_BRACKET_SAMPLE_MIN": "0", "BNA_BRACKET_SAMPLE_SHIFT": str(shift)}, x,
                 density=spa, precision="bf16x3")
    say(f"  launches {t['corr_launches']}")
    assert t["corr_launches"] >= 4  # sampled pass + emission, then the full pair
    compare(g, og)
    same(g, full)


def test_sampled_bracket_binary_and_batch(monkeypatch, subject):
    # a binary build, and a batch of three subjects (the batch histograms the
    # sampled tiles of every subject in one launch)
    x, spa, k, og = subject
    say("sampled bracket binary")
    full, _ = build(monkeypatch, {"BNA_BRACKET_SAMPLE": "0"}, x, density=spa, weighted=False)
    samp, _ = build(monkeypatch, {"BNA_BRACKET_SAMPLE_MIN": "0"}, x, density=spa, weighted=False)
    same(samp, full)
    say("sampled bracket batch")
    xs = [x, bna.synth_factor_model(3000, 150, 8, 0.7, 0xB4AD), bna.synth_factor_model(2500, 240, 8, 0.7, 0xB4AE)]
    monkeypatch.setenv("BNA_BRACKET_SAMPLE_MIN", "0")
    try:
        with bna.Context(0) as c:
            ser = [c.series(v) for v in xs]
            opts = [bna.make_opts(density=spa, precision="bf16x3") for _ in ser]
            for s_ in ser:
                s_.normalize(bna.BF16X3)
            gs, st = bna.build_batch_device(c, ser, opts)
            hosts = [s_.to_host(g) for s_, g in zip(ser, gs)]
            for g in gs:
                c._L.bna_csr_release(c._h, bna.C.byref(g))
            for s_ in ser:
                s_.close()
    finally:
        monkeypatch.delenv("BNA_BRACKET_SAMPLE_MIN")
    assert not any(st)
    for q, (v, h) in enumerate(zip(xs, hosts)):
        solo, _ = build(monkeypatch, {"BNA_BRACKET_SAMPLE": "0"}, v, density=spa, precision="bf16x3")
        say(f"  subject {q}: nnz {h.nnz}")
        same(h, solo, thr=False)  # a copied-out device CSR carries no r*


def test_small_subject_keeps_the_full_pass(monkeypatch):
    # N = 300 has 3 tiles, too few to sample: the build takes the full
    # histogram even with the sampling threshold at zero
    N, L, spa = 300, 40, 5.0
    say("sampled bracket small subject")
    x = bna.synth_factor_model(N, L, 6, 0.5, 0xB4AF)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, spa)
    og = orc.build_density(z64, valid, k)
    g, t = build(monkeypatch, {"BNA_BRACKET_SAMPLE_MIN": "0"}, x, density=spa, precision="bf16x3")
    assert t["corr_launches"] == 2
    compare(g, og)
    assert g.thr_used == og.r_star
