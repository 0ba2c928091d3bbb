"""The benchmark configurations at full size on the GPU, each against the FP64
oracle, through the C ABI.

  C2  N = 45000, L = 240, 1 % density, plain FP32: weighted and binary CSR
      equal to the oracle's bit for bit (digests of its R and C and its r* in
      tests/golden/c2_density.json, written by make_c2_golden.py), every value
      within 1e-6 of r64, degree bit-exact and eigenvector centrality within
      1e-5 relative L2 of the oracle's on the same CSR.
  C3  N = 45000, L = 1200, r > 0.4: split-BF16, plain BF16 and FP32 give the
      same CSR; one oracle row per 256-row tile-row (its first and last row)
      and row N - 1 are compared column for column; every value within 1e-6.
  C4  the 64-subject batch (N uniform in [8000, 60000], L in {150, 240, 480,
      1200}, 1 % density each) in one batched build: every subject equals its
      solo build bit for bit, has exactly 2k entries, and two oracle rows per
      subject match (pairs above the subject's r*, plus ties at r*).
  C5  16 subjects of the serving distribution (N log-uniform in [5000,
      150000], two of them above 100000, L in [150, 1200], binary and weighted
      mixed, 1 % density) through a server with three workers while they
      overlap on the device: every result equals its solo build bit for bit.

The weighted-value bar is 1e-6 absolute against the oracle's r64 (its fixed
FP64 order), checked on every entry: upper entries directly, lower entries by
equality with their mirrors.  Oracle rows come from a float64 product of the
oracle's z, with every pair within 1e-9 of the threshold recomputed in the
oracle's order.
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

pytestmark = pytest.mark.gpu

V_TOL = 1e-6
GOLDEN = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = bna.Context(0)
    yield c
    c.close()


def tile_rows(N, tile=256):
    """Each tile-row's first and last row, and row N - 1."""
    rows = set()
    for t0 in range(0, N, tile):
        rows.add(t0)
        rows.add(min(t0 + tile - 1, N - 1))
    rows.add(N - 1)
    return np.array(sorted(rows), np.int64)


def check_symmetric_sorted(g):
    N = g.N
    assert g.R[0] == 0 and g.R[-1] == g.nnz and np.all(np.diff(g.R) >= 0)
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(g.R))
    assert np.all(g.C != rows), "diagonal entry"
    # strictly ascending within each row: a descent may only sit at a row start
    d = np.diff(g.C.astype(np.int64))
    starts = np.zeros(g.nnz, bool)
    starts[g.R[1:-1][g.R[1:-1] < g.nnz]] = True
    assert np.all((d > 0) | starts[1:]), "a row is not strictly ascending"
    a = rows * N + g.C
    b = g.C.astype(np.int64) * N + rows
    assert np.array_equal(np.sort(a), np.sort(b)), "not symmetric"


def upper_r64(g, z64):
    """The oracle's r64 of every upper-triangle entry (i < j) of g, in its
    fixed evaluation order, and the mask of those entries."""
    rows = np.repeat(np.arange(g.N, dtype=np.int32), np.diff(g.R))
    up = g.C > rows
    ii = np.ascontiguousarray(rows[up])
    jj = np.ascontiguousarray(g.C[up])
    r = np.empty(len(ii))
    orc.olib().orc_corr_pairs(z64.ctypes.data, z64.shape[1], ii.ctypes.data, jj.ctypes.data, len(ii), r.ctypes.data)
    return up, r


def max_value_error(g, up, r64):
    """max |V - r64| over the upper entries; each lower entry (j, i) must hold
    the same value as its upper entry (i, j)."""
    N = g.N
    rows = np.repeat(np.arange(N, dtype=np.int64), np.diff(g.R))
    ku = rows[up] * N + g.C[up]
    kl = g.C[~up].astype(np.int64) * N + rows[~up]
    vu = g.V[up][np.argsort(ku)]
    vl = g.V[~up][np.argsort(kl)]
    assert np.array_equal(vu, vl), "an entry and its mirror differ"
    return float(np.abs(g.V[up].astype(np.float64) - r64).max(initial=0.0))


def check_rows(g, z64, rows, thr, ties_allowed):
    for i, (above, ties, _) in zip(rows, orc.rows_above(z64, rows, thr)):
        got = g.C[g.R[i]:g.R[i + 1]]
        if ties_allowed:  # density: every pair above r*, and some of those exactly at r*
            assert np.all(np.isin(above, got)) and np.all(np.isin(got, np.union1d(above, ties))), f"row {i}"
        else:
            assert np.array_equal(got, above), f"row {i}"


def test_c2_fp32_exact_with_binary_degree_and_ec(ctx):
    gold = json.loads((GOLDEN / "c2_density.json").read_text())
    N, L = gold["N"], gold["L"]
    say(f"C2 fp32: N={N} L={L} density {gold['density_pct']} %")
    x = bna.synth_factor_model(N, L, gold["K"], gold["phi"], gold["seed"])
    k = bna.density_to_k(N, gold["density_pct"])
    assert k == gold["k"] == 10124775
    with ctx.series(x) as s:
        gw = s.build(density=gold["density_pct"], precision="fp32")
        tw = ctx.timings()
        gb = s.build(density=gold["density_pct"], precision="fp32", weighted=False)
    say(f"  weighted nnz={gw.nnz} r*={gw.thr_used!r} candidates={gw.n_candidates} corr_ms={tw['corr_ms']:.2f}")
    # the oracle's graph, bit for bit
    assert gw.nnz == gb.nnz == 2 * k == gold["nnz"]
    assert gw.thr_used == gold["r_star"], (gw.thr_used, gold["r_star"])
    assert hashlib.sha256(gw.R.astype(np.int64).tobytes()).hexdigest() == gold["sha256_R_int64"]
    assert hashlib.sha256(gw.C.astype(np.int32).tobytes()).hexdigest() == gold["sha256_C_int32"]
    assert np.array_equal(gb.R, gw.R) and np.array_equal(gb.C, gw.C) and gb.V is None
    check_symmetric_sorted(gw)
    z64, _, _, _, valid = orc.znorm(x, 0)
    up, r64 = upper_r64(gw, z64)
    err = max_value_error(gw, up, r64)
    say(f"  max |V - r64| over all {gw.nnz} entries: {err:.3g}")
    assert err <= V_TOL
    # degree: float sums in CSR order, bit for bit; binary: row lengths
    deg = ctx.degree(gw)
    assert np.array_equal(deg, orc.degree(gw.R, gw.V, N))
    assert np.array_equal(ctx.degree(gb), np.diff(gw.R).astype(np.float32))
    # eigenvector centrality on the same CSR: the weighted reference's A*y and
    # the binary reference's unit y
    t0 = time.time()
    ec, lam, it = ctx.eigcent(gw)
    oec, olam, oit = orc.eigcent(gw.R, gw.C, gw.V, N, out_mode=0)
    rel = np.linalg.norm(ec - oec) / np.linalg.norm(oec)
    say(f"  weighted EC: {it} iterations (oracle {oit}), lambda {lam:.9g} (oracle {olam:.9g}), "
        f"rel L2 {rel:.3g}, oracle {time.time() - t0:.1f} s")
    assert rel <= 1e-5 and abs(lam - olam) <= 1e-6 * abs(olam)
    ecb, lamb, itb = ctx.eigcent(gb, output=bna.EC_BINARY_REF, binary=True)
    oecb, olamb, oitb = orc.eigcent(gw.R, gw.C, None, N, out_mode=1)
    relb = np.linalg.norm(ecb - oecb) / np.linalg.norm(oecb)
    say(f"  binary EC: {itb} iterations (oracle {oitb}), rel L2 {relb:.3g}")
    assert relb <= 1e-5 and abs(lamb - olamb) <= 1e-6 * abs(olamb)


def test_c3_all_precisions_against_oracle_rows(ctx):
    N, L, thr = 45000, 1200, 0.4
    say(f"C3: N={N} L={L} r > {thr}")
    x = bna.synth_factor_model(N, L, 20, 0.6, 0xB17A0000 + 3000)
    res = {}
    with ctx.series(x) as s:
        for prec in ("bf16x3", "bf16", "fp32"):
            g = s.build(thr=thr, precision=prec)
            t = ctx.timings()
            say(f"  {prec}: nnz={g.nnz} candidates={g.n_candidates} flips={g.n_flips} band={g.band:.3g} "
                f"corr_ms={t['corr_ms']:.2f}")
            res[prec] = g
        z, _, _ = s.download()
    a = res["bf16x3"]
    for prec in ("bf16", "fp32"):
        assert np.array_equal(res[prec].R, a.R) and np.array_equal(res[prec].C, a.C), prec
    z64, _, _, _, valid = orc.znorm(x, 0)
    assert np.array_equal(z, z64)
    check_symmetric_sorted(a)
    rows = tile_rows(N)
    check_rows(a, z64, rows, thr, ties_allowed=False)
    say(f"  {len(rows)} oracle rows (each tile-row's first and last, and row N - 1) match")
    up, r64 = upper_r64(a, z64)
    for prec, g in res.items():
        err = max_value_error(g, up, r64)
        say(f"  {prec}: max |V - r64| over all {g.nnz} entries = {err:.3g}")
        assert err <= V_TOL


def c4_subjects(n_subj=64, seed=0xB17A0000 + 4000):
    """The batched form's subjects, as the benchmark draws them."""
    rng = np.random.default_rng(seed)
    Ns = rng.integers(8000, 60001, size=n_subj)
    Ls = rng.choice([150, 240, 480, 1200], size=n_subj)
    return [(int(N), int(L), seed + q) for q, (N, L) in enumerate(zip(Ns, Ls))]


def test_c4_batch_equals_solo_and_oracle_rows(ctx):
    subs = c4_subjects()
    say(f"C4: {len(subs)} subjects, N {min(s[0] for s in subs)}..{max(s[0] for s in subs)}")
    xs = [bna.synth_factor_model(N, L, 20, 0.6, seed) for N, L, seed in subs]
    series = [ctx.series(x) for x in xs]
    try:
        opts = [bna.make_opts(density=1.0, precision="bf16x3") for _ in series]
        t0 = time.time()
        res = bna.build_batch(ctx, series, opts)
        t = ctx.timings()
        say(f"  batch: {time.time() - t0:.2f} s with download, corr launches {t['corr_launches']}, "
            f"corr_ms {t['corr_ms']:.1f}")
        for q, ((N, L, seed), x, s, g) in enumerate(zip(subs, xs, series, res)):
            assert not isinstance(g, Exception), (q, g)
            assert g.nnz == 2 * bna.density_to_k(N, 1.0), q
            solo = s.build(density=1.0, precision="bf16x3")
            assert np.array_equal(g.R, solo.R) and np.array_equal(g.C, solo.C) and np.array_equal(g.V, solo.V), q
            assert g.thr_used == solo.thr_used
            z64, _, _, _, _ = orc.znorm(x, 0)
            check_rows(g, z64, [q % N, N - 1 - q], g.thr_used, ties_allowed=True)
    finally:
        for s in series:
            s.close()
    say("  every subject equals its solo build; oracle rows match")


def c5_subjects(n=16, seed=0xB17A0000 + 5000):
    rng = np.random.default_rng(seed)
    subs = []
    for q in range(n):
        N = int(np.exp(rng.uniform(np.log(5000), np.log(150000))))
        if q in (3, 11):  # two big subjects, whatever the draw
            N = int(rng.integers(100000, 150001))
        L = int(rng.integers(150, 1201))
        subs.append((N, L, bool(rng.random() < 0.5), seed + 100 + q))
    return subs


def test_c5_served_stream_equals_solo():
    subs = c5_subjects()
    assert sum(N >= 100000 for N, _, _, _ in subs) >= 2
    assert any(w for *_, w, _ in subs) and not all(w for *_, w, _ in subs)
    say("C5: " + ", ".join(f"{N}x{L}{'w' if w else 'b'}" for N, L, w, _ in subs))
    xs = [bna.synth_factor_model(N, L, 20, 0.6, seed) for N, L, _, seed in subs]
    rng = np.random.default_rng(7)
    with bna.Server(workers=3, max_batch=8) as srv:
        ids = []
        for x, (N, L, w, _) in zip(xs, subs):
            ids.append(srv.submit(x, density=1.0, weighted=w))
            time.sleep(float(rng.exponential(0.02)))
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    say(f"  served: rounds {st['rounds']}, max concurrent rounds {st['max_concurrent_rounds']}, "
        f"max batch {st['max_batch_seen']}, failed {st['failed']}")
    assert st["failed"] == 0 and st["completed"] == len(subs)
    with bna.Context(0) as ctx:
        for q, (x, (N, L, w, _), r) in enumerate(zip(xs, subs, out)):
            assert r.status == 0, (q, r.error)
            g = ctx.build_csr(x, density=1.0, weighted=w)
            h = r.graph
            assert h.nnz == g.nnz == 2 * bna.density_to_k(N, 1.0), q
            assert np.array_equal(h.R, g.R) and np.array_equal(h.C, g.C), q
            assert (h.V is None) == (not w) and (g.V is None) == (not w)
            if w:
                assert np.array_equal(h.V, g.V), q
    say("  latency ms: " + " ".join(f"{r.latency_ms:.0f}" for r in out))
