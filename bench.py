#!/usr/bin/env python3
"""Benchmark of the network-construction hot path on one MI355X.

Workload (the headline configuration): one subject, N = 45,000 voxels x
L = 1,200 time points, seeded factor-model series (K = 20 AR(1) latents,
phi = 0.6), Pearson r > 0.4, weighted CSR.  One step = z-normalise + MFMA
correlation with the fused threshold + FP64 recheck of band pairs + CSR written
to device memory, with the raw series already resident in HBM.

value = voxel pairs per second, N(N-1)/2 / step time, split-BF16 (bf16x3).
The plain FP32 path is timed the same way and reported beside it.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import queue
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import cuda_newest20160331_amd as bna  # noqa: E402

PEAK = {"bf16x3": 2500.0, "bf16": 2500.0, "fp32": 157.3}  # TFLOP/s, MI355X dense


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def time_path(ctx, series, N, L, prec, thr, steps, warmup):
    """Returns (ms per step by host clock, mean corr-kernel ms, last graph stats,
    mean per-stage ms from the library's HIP events)."""
    last = None
    for _ in range(warmup):
        series.normalize(bna.PRECISIONS[prec])
        g = series.build_device(thr=thr, precision=prec)
        last = (g.nnz, g.n_candidates, g.n_flips, g.band, g.max_err)
        series.free_device(g)
    ctx._L.bna_synchronize(ctx._h)
    corr, stages = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        series.normalize(bna.PRECISIONS[prec])
        g = series.build_device(thr=thr, precision=prec)
        t = ctx.timings()  # the build's stages and the normalisation queued before it
        corr.append(t["corr_ms"])
        stages.append(t)
        last = (g.nnz, g.n_candidates, g.n_flips, g.band, g.max_err)
        series.free_device(g)
    ctx._L.bna_synchronize(ctx._h)
    t1 = time.perf_counter()
    keys = ("znorm_ms", "corr_ms", "recheck_ms", "select_ms", "csr_ms", "refine_ms")
    st_ms = {k: round(float(np.mean([d[k] for d in stages])), 4) for k in keys}
    # one more build, untimed, copied to the host: digests of R and C and the
    # values, to compare the precisions' graphs
    series.normalize(bna.PRECISIONS[prec])
    g = series.build_device(thr=thr, precision=prec)
    h = series.to_host(g)
    series.free_device(g)
    sha = hashlib.sha256(h.R.astype(np.int64).tobytes() + h.C.astype(np.int32).tobytes()).hexdigest()
    return (t1 - t0) * 1e3 / steps, float(np.mean(corr)), last, st_ms, sha, h.V


def cpu_info() -> dict:
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count() or 1, "affinity": aff, "model": model}


def cpu_baseline(x: np.ndarray, N: int, L: int, thr: float, target_s: float, c2: bool = True):
    """The reference CPU path (normalise, CorMat_cpu correlation with
    CorMat_cpu_MT's j-striding, FormFullGraph(_s), FormCSRGraph, degree, power
    iteration; orc_ref_pipeline in the oracle) on the host cores: C1 and C2 in
    full, and the headline workload (C3) as a bounded row sample.  Threads: the
    process's CPU share, at most 16 (the box's share for one GPU; the reference
    uses one thread per logical processor)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import orc

    info = cpu_info()
    threads = max(1, min(16, info["affinity"]))
    M1 = N * (N - 1) // 2

    def summary(t, M1_):
        tot = t["t_norm"] + t["t_corr"] + t["t_thr"] + t["t_csr"] + t["t_deg"] + t["t_ec"]
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in t.items()} | {
            "total_s": round(tot, 3), "voxel_pairs_per_s": M1_ / tot}

    out = {}
    # C1: N = 5000, L = 200, r > 0.5, whole path
    x1 = bna.synth_factor_model(5000, 200, 8, 0.7, 0xB17A0000 + 1000)
    out["c1"] = summary(orc.ref_pipeline(x1, thr=0.5, nthreads=threads), 5000 * 4999 // 2)
    log(f"cpu: C1 {out['c1']}")
    if c2:  # C2: N = 45000, L = 240, 1 % density (exact k), whole path
        x2 = bna.synth_factor_model(45000, 240, 20, 0.6, 0xB17A0000 + 2000)
        k2 = orc.density_k(45000, 1.0)
        out["c2"] = summary(orc.ref_pipeline(x2, k=k2, nthreads=threads), 45000 * 44999 // 2)
        del x2
        log(f"cpu: C2 {out['c2']}")
    # C3: rows [0, n) of the headline subject, n sized to about target_s
    probe = orc.ref_pipeline(x, row0=0, row1=16, thr=thr, nthreads=threads)
    rate = probe["pairs"] / max(probe["t_corr"], 1e-6)
    want = rate * target_s
    n, acc = 0, 0
    while acc < want and n < N - 1:
        acc += N - 1 - n
        n += 1
    t = orc.ref_pipeline(x, row0=0, row1=n, thr=thr, nthreads=threads)
    # the normalisation covers the whole subject; the sample carries its share
    sec = t["t_corr"] + t["t_thr"] + t["t_csr"] + t["t_norm"] * t["pairs"] / M1
    out["c3_sample"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in t.items()}
    return {
        "value": t["pairs"] / sec,
        "unit": "voxel-pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows 0..{n - 1} of N={N}, L={L} (all pairs j > i: {t['pairs']} pairs, "
                  f"{t['pairs'] / M1:.2%} of the subject): normalise (whole subject, its share counted), "
                  f"correlate, r>{thr} into the bool matrix, CSR scan; {sec:.1f} s on {threads} threads",
        "cpu": info,
        "c1_full": out["c1"],
        "c2_full": out.get("c2"),
        "c3_sample": out["c3_sample"],
    }


def serve_stream(n_subj: int, rate: float, seed: int, workers: int = 4, chunk_ms: float | None = None,
                 reserve_eighths: int | None = None, distinct: int = 0) -> dict:
    """The serving form: n_subj subjects with Poisson arrivals (mean `rate` per
    second), N log-uniform in [5000, 150000], L uniform in [150, 1200], 1 %
    density, a mix of weighted and binary requests, split-BF16.  Latency is
    submit -> device CSR ready; each subject is then run alone through the same
    server for its solo time.  Results are collected (and their device CSR
    freed) by a second thread in submission order while the stream is still
    being submitted.  distinct > 0: the stream's requests cycle through
    that many generated subjects (a 600 s stream of 1200 requests would otherwise
    need ~140 GB of host series)."""
    rng = np.random.default_rng(seed)
    subs = []
    n_gen = n_subj if distinct <= 0 else min(distinct, n_subj)
    for q in range(n_gen):
        N = int(np.exp(rng.uniform(np.log(5000), np.log(150000))))
        L = int(rng.integers(150, 1201))
        subs.append((N, L, bool(rng.random() < 0.5)))
    gaps = rng.exponential(1.0 / rate, size=n_subj)
    log(f"serve: generating {n_gen} subjects (sum N*L = {sum(N * L for N, L, _ in subs):.3g})")
    xs = [bna.synth_factor_model(N, L, 20, 0.6, 0xB17A0000 + 5000 + q) for q, (N, L, _) in enumerate(subs)]
    gen = (xs, subs)
    if n_gen < n_subj:
        xs = [gen[0][q % n_gen] for q in range(n_subj)]
        subs = [gen[1][q % n_gen] for q in range(n_subj)]
    lat, batch = [], []
    with bna.Server(workers=workers, max_batch=8, reserve_per_worker=8 << 30, chunk_ms=chunk_ms,
                    reserve_eighths=reserve_eighths) as srv:
        # warm the kernels once
        srv.wait(srv.submit(xs[0][:, :2000], density=1.0, device_result=True), graph=False)
        t0 = time.perf_counter()
        pending, res, errs = queue.Queue(), [], []

        def collector():
            while (rid := pending.get()) is not None:
                try:
                    res.append(srv.wait(rid, graph=False))
                except Exception as e:  # re-raised once the stream is submitted
                    errs.append(e)

        th = threading.Thread(target=collector)
        th.start()
        t_next = t0
        for q, (x, (N, L, weighted)) in enumerate(zip(xs, subs)):
            t_next += gaps[q]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            pending.put(srv.submit(x, density=1.0, weighted=weighted, device_result=True))
            if q % 16 == 15:
                log(f"serve: submitted {q + 1}, collected {len(res)}")
        pending.put(None)
        th.join()
        if errs:
            raise errs[0]
        wall = time.perf_counter() - t0
        st = srv.stats()
        failed = [r.error for r in res if r.status != 0]
        lat = np.array([r.latency_ms for r in res])
        qms = np.array([r.queue_ms for r in res])
        rms = np.array([r.run_ms for r in res])
        t_arr = np.cumsum(gaps) * 1e3
        phases = [r.phases_ms for r in res]
        devt = [r.build for r in res]
        batch = [r.batch_size for r in res]
        log(f"serve: stream done in {wall:.1f} s, {st['rounds']} rounds, failed {len(failed)}")
        solo_of = []
        for q, (x, (N, L, weighted)) in enumerate(zip(gen[0], gen[1])):
            r = srv.wait(srv.submit(x, density=1.0, weighted=weighted, device_result=True), graph=False)
            solo_of.append(r.latency_ms)
            if q % 32 == 31:
                log(f"serve: solo {q + 1} of {n_gen}")
        solo = np.array([solo_of[q % n_gen] for q in range(n_subj)])
    ratio = lat / solo
    worst = np.argsort(ratio)[::-1][:6]
    log("serve: worst latency/solo: " + "; ".join(
        f"#{q} N={subs[q][0]} L={subs[q][1]} at {t_arr[q]:.0f} ms: {lat[q]:.0f} (queue {qms[q]:.0f} run {rms[q]:.0f})"
        f"/{solo[q]:.0f} ms" for q in worst))
    for q in worst[:3]:
        log(f"serve:   #{q} phases {dict((k, round(v, 1)) for k, v in phases[q].items())} device "
            f"{dict((k, round(v, 2)) for k, v in devt[q].items() if k.endswith('_ms'))}")
    big = np.argsort([-s_[0] * s_[0] * s_[1] for s_ in subs])[:6]
    log("serve: largest: " + "; ".join(
        f"#{q} N={subs[q][0]} L={subs[q][1]} at {t_arr[q]:.0f} ms: {lat[q]:.0f} (queue {qms[q]:.0f} run {rms[q]:.0f})"
        f"/{solo[q]:.0f} ms" for q in big))
    return {
        "subjects": n_subj,
        "distinct_subjects": n_gen,
        "stream_s": wall,
        "arrival_rate_per_s": rate,
        "N_range": [min(s[0] for s in subs), max(s[0] for s in subs)],
        "L_range": [min(s[1] for s in subs), max(s[1] for s in subs)],
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "p99_latency_over_solo": float(np.percentile(ratio, 99)),
        "max_latency_over_solo": float(ratio.max()),
        "solo_p50_ms": float(np.percentile(solo, 50)),
        "subjects_per_s": n_subj / wall,
        "voxel_pairs_per_s": sum(N * (N - 1) // 2 for N, _, _ in subs) / wall,
        "rounds": st["rounds"],
        "max_batch": max(batch),
        "max_concurrent_rounds": st["max_concurrent_rounds"],
        "failed": len(failed),
        "worst": [[subs[q][0], subs[q][1], round(float(lat[q]), 1), round(float(solo[q]), 1)] for q in worst],
        "config": "density 1 %, split-BF16, Poisson arrivals, N log-uniform [5000, 150000], "
                  "CSR left in device memory (latency: submit -> device CSR ready), "
                  "L uniform [150, 1200], half weighted / half binary, "
                  f"{workers} worker contexts",
    }


# min(d, a + b) updates per second on the vector ALU: MI355X's published FP32
# vector rate, 157.3 TFLOP/s = 7.86e13 lane operations per second (an FMA counts
# two flops), at 1.5 instructions per update (v_add_f32 and half a v_min3_f32)
MINPLUS_PEAK = 157.3e12 / 2 / 1.5


def density_and_maps(ctx, steps: int) -> dict:
    """The 1 %-density form: one subject, N = 45,000 x L = 240 (K = 20, phi = 0.6),
    split-BF16, weighted CSR with exactly k = 10,124,775 edges; then degree and
    eigenvector centrality on the device-resident CSR.  Roofline bytes: degree
    4(N+1) + 4 nnz + 4N; one SpMV iteration 4(N+1) + 8 nnz + 8N (x) + 8N (y)."""
    N, L = 45000, 240
    x = bna.synth_factor_model(N, L, 20, 0.6, 0xB17A0000 + 2000)
    M1 = N * (N - 1) // 2
    out = {}
    with ctx.series(x) as s:
        for _ in range(2):
            s.normalize(bna.BF16X3)
            s.free_device(s.build_device(density=1.0, precision="bf16x3"))
        ctx._L.bna_synchronize(ctx._h)
        t0 = time.perf_counter()
        for _ in range(steps):
            s.normalize(bna.BF16X3)
            g = s.build_device(density=1.0, precision="bf16x3")
            if _ < steps - 1:
                s.free_device(g)
        ctx._L.bna_synchronize(ctx._h)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        st = ctx.timings()
        nnz = g.nnz
        band_pairs = g.n_candidates
        deg_ms = []
        for _ in range(3):  # the first call also pays the kernel's first launch
            ctx.degree_device(g)
            deg_ms.append(ctx.timings()["degree_ms"])
        deg_ms = min(deg_ms)
        ec_ms = None
        for _ in range(2):
            _, lam, it = ctx.eigcent_device(g)
            t = ctx.timings()
            if ec_ms is None or t["ec_ms"] < ec_ms:
                ec_ms, iters = t["ec_ms"], t["ec_iters"]
        cp_ms = None
        for _ in range(2):
            _, cp_mean = ctx.cp_device(g)
            t = ctx.timings()["cp_ms"]
            cp_ms = t if cp_ms is None else min(cp_ms, t)
        lp_ms = None
        for _ in range(2):
            lp_val = ctx.lp(g)
            t = ctx.timings()["lp_ms"]
            lp_ms = t if lp_ms is None else min(lp_ms, t)
        log("maps: weighted Lp (blocked Floyd-Warshall)")
        lpw_val, _, lpw_iso = ctx.lp_weighted(g)
        lpw_ms = ctx.timings()["lpw_ms"]
        rl = np.diff(s.to_host(g).R)
        wedges = int(np.sum(rl.astype(np.int64) ** 2))  # (i, j, k) walks the row kernel tests
        s.free_device(g)
    deg_bytes = 4 * (N + 1) + 4 * nnz + 4 * N
    spmv_bytes = 4 * (N + 1) + 8 * nnz + 16 * N
    spmv_gbs = spmv_bytes * iters / (ec_ms / 1e3) / 1e9
    deg_gbs = deg_bytes / (deg_ms / 1e3) / 1e9
    out["build"] = {"workload": f"one subject, N={N} x L={L}, 1 % density (k = {nnz // 2} edges), split-BF16, "
                                "weighted CSR", "ms_per_step": ms, "voxel_pairs_per_s": M1 / (ms / 1e3),
                    "edges_per_s": (nnz // 2) / (ms / 1e3), "corr_ms": st["corr_ms"], "select_ms": st["select_ms"],
                    "recheck_ms": st["recheck_ms"], "refine_ms": st["refine_ms"], "csr_ms": st["csr_ms"],
                    "tile_launches": st["corr_launches"], "bracket_pairs_rechecked": band_pairs}
    out["degree"] = {"ms": deg_ms, "GB_per_s": deg_gbs, "roofline_frac": deg_gbs / 8000.0, "bytes": deg_bytes,
                     "note": "best of 3 calls"}
    out["eigcent"] = {"ms": ec_ms, "iterations": iters, "lambda": lam, "spmv_GB_per_s": spmv_gbs,
                      "roofline_frac": spmv_gbs / 8000.0, "bytes_per_iteration": spmv_bytes,
                      "note": "whole power iteration on the device (SpMV, FP64 norms, stopping test); "
                              "GB/s = algorithmic bytes x iterations / time; best of 2 runs"}
    out["cp"] = {"ms": cp_ms, "mean_cp": cp_mean, "wedges": wedges, "wedges_per_s": wedges / (cp_ms / 1e3),
                 "note": "weighted clustering coefficient (weight pass + one workgroup per row); "
                         "wedges = sum of squared row lengths, the (i, j, k) walks tested; best of 2"}
    out["lp"] = {"ms": lp_ms, "Lp": lp_val, "source_edges_per_s": N * nnz / (lp_ms / 1e3),
                 "note": "characteristic path length of the binary graph: BFS from all N sources, 256 per pass "
                         "over the CSR (bit planes); source_edges = N x nnz, the edge visits of N separate "
                         "searches; best of 2"}
    Np = (N + 127) // 128 * 128
    upd = Np ** 3 / (lpw_ms / 1e3)
    out["lp_weighted"] = {"ms": lpw_ms, "Lp": lpw_val, "isolated": lpw_iso, "minplus_updates_per_s": upd,
                          "valu_frac": upd / MINPLUS_PEAK,
                          "note": "weighted characteristic path length and nodal efficiency: blocked "
                                  "Floyd-Warshall (128-vertex blocks) on the dense cost matrix 1/w of the "
                                  f"weighted graph ({Np}^2 floats in HBM), Np^3 min-plus updates; peak = "
                                  "vector ALU at one add and half a three-way min per update; one run"}
    return out


def bc_form(ctx) -> dict:
    """Betweenness on a C1-shaped subject's 1 % density graph (N = 5000, L = 200,
    K = 8, phi = 0.7): Brandes from every source on the device CSR."""
    N, L = 5000, 200
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A0000 + 1000)
    with ctx.series(x) as s:
        s.normalize(bna.BF16X3)
        g = s.build_device(density=1.0, precision="bf16x3", weighted=False)
        try:
            ms = None
            for _ in range(2):
                b = ctx.bc(g)
                t = ctx.timings()["bc_ms"]
                ms = t if ms is None else min(ms, t)
            nnz = g.nnz
        finally:
            s.free_device(g)
    return {"workload": f"N={N} x L={L}, 1 % density binary graph (nnz {nnz})", "ms": ms,
            "teps": N * nnz / (ms / 1e3), "max_bc": float(b.max()),
            "note": "Brandes from all N sources (one workgroup each, double path counts); "
                    "teps = N x nnz / time; best of 2"}


def batch_form(n_subj: int, seed: int) -> dict:
    """The batched form: n_subj subjects with N uniform in [8000, 60000] and L in
    {150, 240, 480, 1200}, 1 % density each, split-BF16, built by one
    bna_build_csr_batch (one histogram launch and one emission launch of the
    persistent tile kernel over every subject's tiles)."""
    rng = np.random.default_rng(seed)
    Ns = rng.integers(8000, 60001, size=n_subj)
    Ls = rng.choice([150, 240, 480, 1200], size=n_subj)
    log(f"batch: generating {n_subj} subjects (sum N*L = {float(np.sum(Ns * Ls)):.3g})")
    xs = [bna.synth_factor_model(int(N), int(L), 20, 0.6, seed + q) for q, (N, L) in enumerate(zip(Ns, Ls))]
    pairs = int(sum(int(N) * (int(N) - 1) // 2 for N in Ns))
    with bna.Context(0) as ctx:
        ser = [ctx.series(x) for x in xs]
        del xs
        opts = [bna.make_opts(density=1.0, precision="bf16x3") for _ in ser]

        def run():
            for s_ in ser:
                s_.normalize(bna.BF16X3)
            gs, st = bna.build_batch_device(ctx, ser, opts)
            ctx._L.bna_synchronize(ctx._h)
            return gs, st

        gs, st = run()  # warm: kernels, workspaces
        for g in gs:
            ctx._L.bna_csr_release(ctx._h, bna.C.byref(g))
        t0 = time.perf_counter()
        gs, st = run()
        dt = time.perf_counter() - t0
        tm = ctx.timings()
        edges = sum(g.nnz // 2 for g in gs)
        for g in gs:
            ctx._L.bna_csr_release(ctx._h, bna.C.byref(g))
        for s_ in ser:
            s_.close()
    # algorithmic flops 2L per pair, counted once although a density build makes
    # two tile passes (histogram, then emission)
    flops = float(sum(int(N) * (int(N) - 1) * int(L) for N, L in zip(Ns, Ls)))
    tile_tf = flops / (tm["corr_ms"] / 1e3) / 1e12
    return {"subjects": n_subj, "failed": int(sum(1 for v in st if v)), "ms": dt * 1e3,
            "voxel_pairs_per_s": pairs / dt, "edges": edges, "edges_per_s": edges / dt,
            "corr_ms": tm["corr_ms"], "corr_launches": tm["corr_launches"],
            "tile_algorithmic_TFLOP_per_s": tile_tf, "tile_roofline_frac": tile_tf / PEAK["bf16x3"],
            "N_range": [int(Ns.min()), int(Ns.max())], "L_set": sorted(set(int(v) for v in Ls)),
            "config": "density 1 % each, split-BF16, weighted CSR left in device memory, series resident"}


def group_form(n_subj: int, seed: int, N: int = 45000, L: int = 240) -> dict:
    """Group averaging (the avg = yf path): n_subj subjects of the C2 shape
    added into device-resident Fisher-z sums by the tile kernel's accumulation
    pass, averaged, then a 1 % density graph of the averages."""
    xs = [bna.synth_factor_model(N, L, 20, 0.6, seed + q) for q in range(n_subj)]
    M1 = N * (N - 1) // 2
    with bna.Context(0) as ctx:
        ser = [ctx.series(x) for x in xs]
        del xs
        for s_ in ser:
            s_.normalize(bna.BF16X3)
        with ctx.group(N, fisher=True) as warm:  # kernels and workspaces
            warm.add(ser[0])
            warm.finish()
            warm.build(density=1.0)
        g = ctx.group(N, fisher=True)
        ctx._L.bna_synchronize(ctx._h)
        t0 = time.perf_counter()
        add_ms = []
        for s_ in ser:
            g.add(s_)
            add_ms.append(ctx.timings()["corr_ms"])
        t1 = time.perf_counter()
        g.finish()
        t2 = time.perf_counter()
        G = g.build(density=1.0)
        t3 = time.perf_counter()
        gt = ctx.timings()
        g.close()
        for s_ in ser:
            s_.close()
    add_dt = t1 - t0
    return {"subjects": n_subj, "N": N, "L": L, "add_ms_per_subject": add_dt * 1e3 / n_subj,
            "accumulate_kernel_ms": float(np.mean(add_ms)),
            "voxel_pairs_per_s": M1 * n_subj / add_dt,
            "finish_ms": (t2 - t1) * 1e3, "graph_ms": (t3 - t2) * 1e3,
            "graph_select_ms": gt["select_ms"], "graph_csr_ms": gt["csr_ms"], "edges": G.nnz // 2,
            "r_star": G.thr_used,
            "config": "Fisher-z group mean (avg yf), split-BF16, 1 % density of the averages, "
                      "series resident; graph_ms includes the CSR copy to the host"}


def measured_traffic(N: int, L: int):
    """HBM bytes per launch of the split-BF16 tile kernel at N=45000, L=1200, from
    the committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a
    copy of known size); None for other shapes."""
    f = ROOT / "profiles" / "c3_n45000_l1200_hbm_traffic.txt"
    if (N, L) != (45000, 1200) or not f.exists():
        return None
    for line in f.read_text().splitlines():
        name = line.split(" ")[0].removeprefix("bna::")
        if name in ("corr_tile8_kernel<2,", "corr_tile_kernel<2,") and line.split()[1] == "0>":
            return float(line.split()[-1])
    return None


def calibrate_copy(gib: float = 1.0) -> None:
    """A plain device-to-device copy of known size, to scale the HBM counters
    (FETCH_SIZE / WRITE_SIZE) read under rocprofv3 --pmc: the copy kernel reads and
    writes exactly `gib` GiB."""
    import torch

    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"calibrate: copied {n * 4} bytes 3 times")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=45000)
    ap.add_argument("--l", type=int, default=1200)
    ap.add_argument("--thr", type=float, default=0.4)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-c2", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--serve-subjects", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-workers", type=int, default=3)
    ap.add_argument("--serve-distinct", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-chunk-ms", type=float, default=None)
    ap.add_argument("--serve-reserve-eighths", type=int, default=None)
    ap.add_argument("--no-maps", action="store_true")
    ap.add_argument("--batch-subjects", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--group-subjects", type=int, default=8)
    ap.add_argument("--no-group", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy()
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        return 2

    N, L, thr = a.n, a.l, a.thr
    seed = 0xB17A0000 + 3000
    log(f"bench: generating N={N} L={L} seed={seed:#x}")
    x = bna.synth_factor_model(N, L, 20, 0.6, seed)
    M1 = N * (N - 1) // 2
    flops = float(N) * (N - 1) * L  # 2L per pair

    res = {}
    maps = None
    tuned = {}
    with bna.Context(0) as ctx:
        # tile traversal (band height) per precision and length bucket
        for prec, Lt in (("bf16x3", L), ("fp32", L), ("bf16", L), ("bf16x3", 240)):
            if (prec == "fp32" and a.no_fp32) or (prec == "bf16" and a.no_bf16):
                continue
            gb, gms = ctx.autotune(prec, Lt, N=N)
            tuned[f"{prec}_L{Lt}"] = {"band_rows": gb, f"tile_ms_at_N{N}": round(gms, 4)}
        log(f"bench: autotuned {tuned}")
        with ctx.series(x) as s:
            for prec in (["bf16x3"] + ([] if a.no_fp32 else ["fp32"]) + ([] if a.no_bf16 else ["bf16"])):
                log(f"bench: {prec} warmup={a.warmup} steps={a.steps}")
                ms, corr_ms, st, st_ms, sha, V = time_path(ctx, s, N, L, prec, thr, a.steps, a.warmup)
                res[prec] = {"ms": ms, "corr_ms": corr_ms, "stats": st, "stages_ms": st_ms, "csr_sha": sha, "V": V}
                log(f"bench: {prec} {ms:.3f} ms/step, corr kernel {corr_ms:.3f} ms, nnz={st[0]} "
                    f"cand={st[1]} flips={st[2]} band={st[3]:.3g} max_err={st[4]:.3g} stages={st_ms}")
        if not a.no_maps:
            log("bench: 1 % density form and graph maps")
            maps = density_and_maps(ctx, 5)
            maps["bc"] = bc_form(ctx)
            log(f"bench: maps {maps}")

    b = res["bf16x3"]
    value = M1 / (b["ms"] / 1e3)
    achieved = flops / (b["corr_ms"] / 1e3) / 1e12
    line = {
        "metric": "corr+CSR voxel-pairs/s @N=45k,L=1200 (split-BF16 vs FP32); p99 subject latency",
        "value": value,
        "unit": "voxel-pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": b["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3",
        "data": "synthetic",
        "config": {"workload": f"one subject, N={N} x L={L} fp32 factor-model series (K=20, phi=0.6), "
                               f"r>{thr}, weighted CSR, split-BF16 MFMA + FP64 band recheck",
                   "N": N, "L": L, "thr": thr, "precision": "bf16x3", "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK["bf16x3"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK["bf16x3"], "traffic": measured_traffic(N, L)},
        "autotune": tuned,
        "edges": b["stats"][0] // 2,
        "edges_per_s": (b["stats"][0] // 2) / (b["ms"] / 1e3),
        "band_pairs_rechecked": b["stats"][1],
        "max_abs_err_vs_fp64": b["stats"][4],
        "p99_subject_latency_ms": None,
        "stages_ms": b["stages_ms"],
    }
    # kernel (c), the CSR compaction (fragment rows, scan, placement): its
    # algorithmic bytes are the CSR it writes, 8 nnz + 4 (N + 1); an edge is
    # two entries, 16 bytes, so HBM bounds it at 8 TB/s / 16 B = 5e11 edges/s
    nnz = b["stats"][0]
    csr_s = b["stages_ms"]["csr_ms"] / 1e3
    csr_bytes = 8 * nnz + 4 * (N + 1)
    csr_gbs = csr_bytes / csr_s / 1e9 if csr_s > 0 else None
    line["csr"] = {"ms": b["stages_ms"]["csr_ms"], "bytes": csr_bytes, "GB_per_s": csr_gbs,
                   "roofline_frac": csr_gbs / 8000.0 if csr_gbs else None,
                   "edges_per_s": (nnz // 2) / csr_s if csr_s > 0 else None,
                   "edges_roofline_frac": ((nnz // 2) / csr_s) / (8e12 / 16) if csr_s > 0 else None,
                   "note": "fragment rows + scan + placement (kernel c), HIP events; bytes = 8 nnz + 4 (N + 1)"}
    line["csr_roofline_frac"] = line["csr"]["roofline_frac"]
    line["edges_roofline_frac"] = line["edges_per_s"] / (8e12 / 16)
    if "fp32" in res:
        f = res["fp32"]
        fval = M1 / (f["ms"] / 1e3)
        line["fp32_value"] = fval
        line["speedup_vs_fp32"] = value / fval
        line["fp32_roofline_frac"] = flops / (f["corr_ms"] / 1e3) / 1e12 / PEAK["fp32"]
        line["same_csr_as_fp32"] = f["csr_sha"] == b["csr_sha"]
        line["max_abs_dV_vs_fp32"] = float(np.max(np.abs(f["V"].astype(np.float64) - b["V"]), initial=0.0))
        line["fp32_stages_ms"] = f["stages_ms"]
    if "bf16" in res:
        # the other precision form of the same workload: plain BF16 on MFMA (one
        # pass), a wider error band rechecked in FP64, the same CSR
        p16 = res["bf16"]
        v16 = M1 / (p16["ms"] / 1e3)
        line["bf16"] = {"value": v16, "ms_per_step": p16["ms"], "corr_ms": p16["corr_ms"],
                        "roofline_frac": flops / (p16["corr_ms"] / 1e3) / 1e12 / PEAK["bf16"],
                        "band": p16["stats"][3], "band_pairs_rechecked": p16["stats"][1],
                        "same_csr_as_bf16x3": p16["csr_sha"] == b["csr_sha"],
                        "max_abs_dV_vs_bf16x3": float(np.max(np.abs(p16["V"].astype(np.float64) - b["V"]),
                                                             initial=0.0)),
                        "stages_ms": p16["stages_ms"]}
        if "fp32" in res:
            line["bf16"]["speedup_vs_fp32"] = v16 / line["fp32_value"]
    if maps is not None:
        line["density_form"] = maps["build"]
        line["degree"] = maps["degree"]
        line["eigcent"] = maps["eigcent"]
        line["cp"] = maps["cp"]
        line["lp"] = maps["lp"]
        line["bc"] = maps["bc"]
        line["lp_weighted"] = maps["lp_weighted"]
    if not a.no_batch:
        log("bench: batched form")
        line["batch"] = batch_form(a.batch_subjects, 0xB17A0000 + 4000)
        log(f"bench: batch {line['batch']}")
    if not a.no_group:
        log("bench: group average")
        line["group"] = group_form(a.group_subjects, 0xB17A0000 + 6000)
        log(f"bench: group {line['group']}")
    if not a.no_serve:
        log("bench: serving stream")
        sv = serve_stream(a.serve_subjects, a.serve_rate, 0xB17A0000 + 5000, a.serve_workers,
                          a.serve_chunk_ms, a.serve_reserve_eighths, a.serve_distinct)
        line["serving"] = sv
        line["p99_subject_latency_ms"] = sv["p99_ms"]
    if not a.no_cpu:
        log("bench: CPU baseline")
        line["cpu_baseline"] = cpu_baseline(x, N, L, thr, a.cpu_seconds, c2=not a.no_cpu_c2)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
