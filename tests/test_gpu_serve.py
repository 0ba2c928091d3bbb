"""The request loop on the GPU: requests arriving while others run are admitted
by footprint, built in overlapping batched rounds, and every result equals the
same subject's solo build (CSR, degree and eigenvector centrality bit for bit).
"""
import sys
import threading
import time

import numpy as np
import pytest

import cuda_newest20160331_amd as bna

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mix(n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for q in range(n):
        N = int(np.exp(rng.uniform(np.log(500), np.log(4000))))
        L = int(rng.integers(60, 400))
        kw = dict(density=float(rng.choice([0.5, 1.0, 2.0]))) if q % 2 else dict(thr=float(rng.uniform(0.35, 0.5)))
        kw["weighted"] = bool(rng.random() < 0.7)
        x = bna.synth_factor_model(N, L, 8, 0.7, 1000 + q)
        reqs.append((x, kw, q % 3 == 0))
    return reqs


def solo(ctx, x, kw, with_maps):
    g = ctx.build_csr(x, **kw)
    deg = ctx.degree(g) if with_maps else None
    ec = ctx.eigcent(g, output=0 if g.V is not None else 1) if with_maps else None
    return g, deg, ec


def check(res, ref):
    g, deg, ec = ref
    assert res.status == 0, res.error
    h = res.graph
    assert h.nnz == g.nnz and np.array_equal(h.R, g.R) and np.array_equal(h.C, g.C)
    assert (h.V is None) == (g.V is None)
    if g.V is not None:
        assert np.array_equal(h.V, g.V)
    if deg is not None:
        assert np.array_equal(res.degree, deg)
        assert np.array_equal(res.ec, ec[0]) and res.lam == ec[1]


def test_mixed_stream_matches_solo():
    reqs = mix(10, 3)
    with bna.Context(0) as ctx:
        refs = [solo(ctx, x, kw, m) for x, kw, m in reqs]
    with bna.Server(workers=2, max_batch=4) as srv:
        ids = []

        def feed():
            for x, kw, m in reqs:
                ids.append(srv.submit(x, degree=m, ec=m, **kw))
                time.sleep(0.003)

        t = threading.Thread(target=feed)
        t.start()
        t.join()
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    say(f"serve: {st}")
    for res, ref in zip(out, refs):
        check(res, ref)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    say("  latency ms: " + " ".join(f"{r.latency_ms:.1f}" for r in out))


def test_burst_is_packed_into_rounds():
    # one worker: while its first round runs, the rest of the burst queues up
    # and the next rounds take several subjects each
    reqs = mix(10, 5)
    with bna.Context(0) as ctx:
        refs = [solo(ctx, x, kw, m) for x, kw, m in reqs]
    with bna.Server(workers=1, max_batch=4) as srv:
        ids = [srv.submit(x, degree=m, ec=m, **kw) for x, kw, m in reqs]
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    say(f"burst: {st}; batch sizes {[r.batch_size for r in out]}")
    for res, ref in zip(out, refs):
        check(res, ref)
    assert st["max_batch_seen"] > 1 and st["rounds"] < len(reqs)


def test_admission_respects_budget():
    reqs = mix(8, 4)
    fps = [bna.request_footprint(x.shape[1], x.shape[0], **kw) for x, kw, _ in reqs]
    budget = int(max(fps) * 1.5)
    with bna.Context(0) as ctx:
        refs = [solo(ctx, x, kw, False) for x, kw, _ in reqs]
    with bna.Server(hbm_budget=budget, workers=2, max_batch=8) as srv:
        ids = [srv.submit(x, **kw) for x, kw, _ in reqs]
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    say(f"admission: budget {budget} max_reserved {st['max_reserved']}")
    assert st["max_reserved"] <= budget
    for res, ref in zip(out, refs):
        check(res, ref)


def test_cancel_and_poll():
    # one worker, one request per round: a long first request (a density
    # threshold at N = 60000, L = 1200, two tile passes and the refinement) keeps
    # the small ones queued, so the last can be cancelled whatever the device speed
    big = bna.synth_factor_model(60000, 1200, 8, 0.7, 76)
    x = bna.synth_factor_model(6000, 300, 8, 0.7, 77)
    with bna.Server(workers=1, max_batch=1) as srv:
        ids = [srv.submit(big, density=1.0)] + [srv.submit(x, thr=0.5) for _ in range(3)]
        assert srv.cancel(ids[-1]) == 0
        first = srv.poll(ids[0])
        assert first == bna.Server.BUSY or first.status == 0
        res = srv.wait(ids[-1])
        assert res.status == bna.Server.CANCELLED
        for i in ids[1:-1]:
            assert srv.wait(i).status == 0
        if first == bna.Server.BUSY:
            assert srv.wait(ids[0]).status == 0
        assert srv.poll(ids[0]) == bna.Server.NOT_FOUND
        assert srv.stats()["cancelled"] == 1


def test_device_resident_result():
    x = bna.synth_factor_model(3000, 200, 8, 0.7, 91)
    with bna.Context(0) as ctx:
        g = ctx.build_csr(x, density=1.0)
    with bna.Server(workers=2) as srv:
        r = srv.wait(srv.submit(x, density=1.0, device_result=True))
    assert r.status == 0 and r.on_device and r.graph is None and r.nnz == g.nnz
