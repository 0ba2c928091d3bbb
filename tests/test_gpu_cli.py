"""The command-line tools end to end on the GPU: NIfTI directory in, .csr / .nm /
_evalue.txt files out, checked against the oracle on the ingested series.

File names follow CorMat.cpp:294-350 (`<stem>_<N>_spa%.3f%%_cor%.3f[_weighted].csr`
under weighted/ or unweighted/), Degree.cpp:93 (`_deg.nm`) and
weighted-net.cu:300-301 (`_ec.nm`, `_evalue.txt`).
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc
from nifti import write_nii

pytestmark = pytest.mark.gpu
BIN = Path(bna.__file__).resolve().parent / "bin"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(tool, *args):
    say(f"  run {tool} {' '.join(map(str, args))}")
    r = subprocess.run([str(BIN / tool), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


@pytest.fixture(scope="module")
def subjects(tmp_path_factory):
    d = tmp_path_factory.mktemp("bold")
    X, Y, Z = 16, 16, 10
    rng = np.random.default_rng(11)
    mask = (rng.random((X, Y, Z)) < 0.85).astype(np.uint8)
    write_nii(d / "mask.nii", mask)
    out = {}
    for name, L, dt, seed in (("subjA", 150, np.float32, 5), ("subjB", 96, np.float64, 6)):
        ser = bna.synth_factor_model(X * Y * Z, L, 8, 0.7, seed)  # (L, voxels)
        vol = ser.T.reshape(X, Y, Z, L, order="F").astype(dt)
        write_nii(d / f"{name}.nii", vol)
        out[name] = L
    return d, mask, out


def ingest(d, name):
    sel, N = bna.nifti_mask(str(d / "mask.nii"), 0.5)
    return bna.nifti_bold(str(d / f"{name}.nii"), sel), N


def test_cormat_r_threshold_weighted_files(subjects):
    d, mask, Ls = subjects
    say("cli: r thresholds, weighted")
    run("bna_cormat", d, d / "mask.nii", 0.5, "n", "n", "r", 0.5, 0.45)
    for name in Ls:
        x, N = ingest(d, name)
        z64, _, _, _, valid = orc.znorm(x, 0)
        M1 = N * (N - 1) / 2
        for thr in (0.45, 0.5):
            og = orc.build_threshold(z64, valid, np.float32(thr))
            spa = np.float32(100.0 * og.nnz / M1 / 2)
            f = d / "weighted" / f"{name}_{N}_spa{spa:.3f}%_cor{np.float32(thr):.3f}_weighted.csr"
            assert f.exists(), sorted(p.name for p in (d / "weighted").iterdir())
            g = bna.read_csr(f)
            assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C)
            assert np.abs(g.V.astype(np.float64) - og.V64).max(initial=0) <= 1e-6


def test_cormat_density_binary_and_maskless_form(subjects):
    d, mask, Ls = subjects
    say("cli: density, binary, mask path left out")
    run("bna_cormat", "--binary", d, 0.5, "n", "n", "s", 2.0)
    for name in Ls:
        x, N = ingest(d, name)
        z64, _, _, _, valid = orc.znorm(x, 0)
        k = orc.density_k(N, np.float32(2.0))
        og = orc.build_density(z64, valid, k)
        f = d / "unweighted" / f"{name}_{N}_spa{np.float32(2.0):.3f}%_cor{np.float32(og.r_star):.3f}.csr"
        assert f.exists(), sorted(p.name for p in (d / "unweighted").iterdir())
        g = bna.read_csr(f)
        assert g.V is None
        assert np.array_equal(g.R, og.R) and np.array_equal(g.C, og.C)


def test_degree_and_ec_tools(subjects):
    d, mask, Ls = subjects
    say("cli: degree + EC over the weighted and unweighted outputs")
    for sub in ("weighted", "unweighted"):
        if not (d / sub).exists() or not list((d / sub).glob("*.csr")):
            pytest.skip("construction outputs missing")
        run("bna_degree", d / sub)
        run("bna_ec", d / sub)
        iso = []
        for f in sorted((d / sub).glob("*.csr")):
            g = bna.read_csr(f)
            R64 = g.R.astype(np.int64)
            deg = bna.read_nm(str(f)[:-4] + "_deg.nm")
            assert np.array_equal(deg, orc.degree(R64, g.V, g.N))
            iso.append(int(np.sum(np.diff(R64) == 0)))
            ec = bna.read_nm(str(f)[:-4] + "_ec.nm")
            oec, olam, _ = orc.eigcent(R64, g.C, g.V, g.N, out_mode=0 if g.V is not None else 1)
            assert np.linalg.norm(ec - oec) <= 1e-5 * np.linalg.norm(oec)
            lam = float(Path(str(f)[:-4] + "_evalue.txt").read_text())
            assert abs(lam - olam) <= 1e-5 * abs(olam)  # printed with 6 significant digits
        if sub == "unweighted":
            lines = (d / sub / "isolated_v_mark.txt").read_text().split()
            assert list(map(int, lines)) == iso


def read_cormat(path, N):
    raw = np.fromfile(path, dtype=np.uint8)
    head = int(raw[:4].view(np.uint32)[0])
    M1 = N * (N - 1) // 2
    assert head == M1 & 0xFFFFFFFF
    vals = raw[4:].view(np.float32)
    assert vals.size == M1
    return vals


def edge_set(R, Cc):
    rows = np.repeat(np.arange(len(R) - 1), np.diff(R))
    k = rows < Cc
    return set(zip(rows[k].tolist(), Cc[k].tolist()))


def test_cormat_group_average_fisher(subjects):
    d, mask, Ls = subjects
    say("cli: group average (Fisher), .cormat")
    run("bna_cormat", d, d / "mask.nii", 0.5, "yf", "y", "r", 0.3)
    packs, N = [], None
    for name in sorted(Ls):
        x, N = ingest(d, name)
        z64, _, _, _, valid = orc.znorm(x, 0)
        packs.append(orc.all_pairs(z64, valid))
    ref = orc.group_average(packs, N, fisher=True)
    base = f"{d.name}_{len(Ls)}_{N}"
    cm = read_cormat(d / f"{base}.cormat", N)
    assert np.abs(cm.astype(np.float64) - ref).max() <= 2e-5
    files = sorted(d.glob(f"{base}_spa*_cor0.300.csr"))
    assert len(files) == 1, sorted(p.name for p in d.iterdir())
    g = bna.read_csr(files[0])
    og = orc.threshold_packed(ref, N, np.float32(0.3))
    for (i, j) in edge_set(g.R, g.C) ^ edge_set(og.R, og.C):
        p = i * N - (i + 1) * i // 2 + j - i - 1
        assert abs(float(ref[p]) - 0.3) <= 2e-5
    spa = np.float32(100.0 * g.nnz / (N * (N - 1) / 2) / 2)
    assert files[0].name == f"{base}_spa{spa:.3f}%_cor0.300.csr"


def test_cormat_both_modes_with_subject_cormat(subjects):
    d, mask, Ls = subjects
    say("cli: avg bn (each subject and the group mean), density, .cormat")
    run("bna_cormat", d, 0.5, "bn", "y", "s", 2.0)
    packs, N = [], None
    for name in sorted(Ls):
        x, N = ingest(d, name)
        z64, _, _, _, valid = orc.znorm(x, 0)
        p = orc.all_pairs(z64, valid)
        packs.append(p)
        cm = read_cormat(d / "weighted" / f"{name}_{N}.cormat", N)
        exp = np.where(np.isnan(p), 0.0, p)
        assert np.abs(cm - exp).max() <= 4e-6
        assert any((d / "weighted").glob(f"{name}_{N}_spa2.000%_cor*_weighted.csr"))
    ref = orc.group_average(packs, N, fisher=False)
    base = f"{d.name}_{len(Ls)}_{N}"
    assert np.abs(read_cormat(d / f"{base}.cormat", N) - ref).max() <= 4e-6
    files = sorted(d.glob(f"{base}_spa2.000%_cor*.csr"))
    assert len(files) == 1
    g = bna.read_csr(files[0])
    k = orc.density_k(N, np.float32(2.0))
    assert g.nnz == 2 * k and g.V is not None


def test_cormat_group_binary(subjects):
    d, mask, Ls = subjects
    say("cli: group average, binary graph")
    run("bna_cormat", "--binary", d, 0.5, "yn", "n", "s", 3.0)
    x, N = ingest(d, sorted(Ls)[0])
    base = f"{d.name}_{len(Ls)}_{N}"
    files = sorted(d.glob(f"{base}_spa3.000%_cor*.csr"))
    assert len(files) == 1, sorted(p.name for p in d.iterdir())
    g = bna.read_csr(files[0])
    assert g.V is None and g.nnz == 2 * orc.density_k(N, np.float32(3.0))


def test_cp_tool(subjects):
    d, mask, Ls = subjects
    say("cli: clustering coefficient over the weighted and unweighted outputs")
    for sub in ("weighted", "unweighted"):
        if not (d / sub).exists() or not list((d / sub).glob("*.csr")):
            pytest.skip("construction outputs missing")
        run("bna_cp", d / sub, 0)
        for f in sorted((d / sub).glob("*.csr")):
            g = bna.read_csr(f)
            ocp, omean = orc.cp(g.R.astype(np.int64), g.C, g.V, g.N)
            cp = bna.read_nm(str(f)[:-4] + ".cp")
            if g.V is None:
                assert np.array_equal(cp, ocp)
            else:
                assert np.allclose(cp, ocp, rtol=1e-6, atol=0)
            mean = float(Path(str(f)[:-4] + "_cp.txt").read_text())
            assert abs(mean - omean) <= 1e-5 * abs(omean)  # six significant digits


def test_modularity_tool(subjects):
    d, mask, Ls = subjects
    say("cli: modularity over the unweighted outputs")
    sub = d / "unweighted"
    if not sub.exists() or not list(sub.glob("*.csr")):
        pytest.skip("construction outputs missing")
    run("bna_modularity", sub, 0)
    for f in sorted(sub.glob("*.csr")):
        g = bna.read_csr(f)
        olab, oQ, omods, rounds, its = orc.modularity(g.R.astype(np.int64), g.C, g.N)
        lab = bna.read_nm(str(f)[:-4] + "_modu.nm")
        assert np.array_equal(lab, olab.astype(np.float32))
        txt = Path(str(f)[:-4] + "_modu.txt").read_text()
        assert txt.startswith(f"Number of Modules: {omods},")
        assert abs(float(txt.split("Q=")[1]) - oQ) <= 1e-5 * abs(oQ)


def test_lp_and_bc_tools(subjects):
    d, mask, Ls = subjects
    say("cli: path length and betweenness over the unweighted outputs")
    sub = d / "unweighted"
    if not sub.exists() or not list(sub.glob("*.csr")):
        pytest.skip("construction outputs missing")
    run("bna_lp", sub, 0)
    run("bna_bc", sub, 0)
    for f in sorted(sub.glob("*.csr")):
        g = bna.read_csr(f)
        olp, _ = orc.lp(g.R.astype(np.int64), g.C, g.N)
        lp = float(Path(str(f)[:-4] + "_Lp.txt").read_text())
        assert abs(lp - olp) <= 1e-5 * abs(olp)  # six significant digits
        ob = orc.bc(g.R.astype(np.int64), g.C, g.N)
        b = bna.read_nm(str(f)[:-4] + "_bc.nm")
        assert np.all(np.abs(b - ob) <= 2e-6 * np.abs(ob) + 1e-6)


def test_lpw_tool(subjects, tmp_path):
    d, mask, Ls = subjects
    say("cli: weighted path length and nodal efficiency")
    files = sorted((d / "weighted").glob("*.csr")) if (d / "weighted").exists() else []
    if not files:
        pytest.skip("construction outputs missing")
    # one weighted graph in a directory of its own (the tool takes every .csr)
    f = tmp_path / files[0].name
    f.write_bytes(files[0].read_bytes())
    run("bna_lpw", tmp_path, 0, "gn")
    g = bna.read_csr(f)
    olp, oeff, _ = orc.lpw(g.R.astype(np.int64), g.C, g.V, g.N)
    lp = float(Path(str(f)[:-4] + "_Lp.txt").read_text())
    assert abs(lp - olp) <= 1e-5 * abs(olp)  # six significant digits
    assert np.array_equal(bna.read_nm(str(f)[:-4] + "_eff.nm"), oeff)


def test_random_network_tools(subjects, tmp_path):
    # With --seed the tools' Maslov networks are bna_maslov(g, seed + l), so the
    # lines they add can be recomputed here: Cp, binary Lp and weighted Lp of the
    # same random networks, and lambda = Lp / mean of the random networks' Lp
    # (Cp_main.cpp:130-151, CUBFS_LP_main.cpp:110-123, CUBFW_Lp_main.cpp:219-253)
    d, mask, Ls = subjects
    say("cli: random-network lines of the Cp and Lp tools")
    wf = sorted((d / "weighted").glob("*.csr")) if (d / "weighted").exists() else []
    uf = sorted((d / "unweighted").glob("*.csr")) if (d / "unweighted").exists() else []
    if not wf or not uf:
        pytest.skip("construction outputs missing")
    w = tmp_path / "w"
    u = tmp_path / "u"
    w.mkdir()
    u.mkdir()
    fw = w / wf[0].name
    fw.write_bytes(wf[0].read_bytes())
    fu = u / uf[0].name
    fu.write_bytes(uf[0].read_bytes())
    gw, gu = bna.read_csr(fw), bna.read_csr(fu)

    run("bna_cp", "--seed=5", w, 2)
    lines = [float(v) for v in Path(str(fw)[:-4] + "_cp.txt").read_text().split()]
    want = [orc.cp(g.R.astype(np.int64), g.C, g.V, g.N)[1] for g in (gw, bna.maslov(gw, 5), bna.maslov(gw, 6))]
    assert len(lines) == 3
    assert all(abs(a - b) <= 1e-5 * abs(b) for a, b in zip(lines, want)), (lines, want)

    run("bna_lp", "--seed=3", u, 2)
    lines = [float(v) for v in Path(str(fu)[:-4] + "_Lp.txt").read_text().split()]
    want = [orc.lp(g.R.astype(np.int64), g.C, g.N)[0] for g in (gu, bna.maslov(gu, 3), bna.maslov(gu, 4))]
    assert len(lines) == 3
    assert all(abs(a - b) <= 1e-5 * abs(b) for a, b in zip(lines, want)), (lines, want)

    run("bna_lpw", "--seed=9", w, 1, "gl")
    lines = [float(v) for v in Path(str(fw)[:-4] + "_Lp.txt").read_text().split()]
    want = [orc.lpw(g.R.astype(np.int64), g.C, g.V, g.N)[0] for g in (gw, bna.maslov(gw, 9))]
    assert len(lines) == 2
    assert all(abs(a - b) <= 1e-5 * abs(b) for a, b in zip(lines, want)), (lines, want)
    lam = float(Path(str(fw)[:-4] + "_lambda.txt").read_text())
    assert abs(lam - want[0] / want[1]) <= 1e-5 * abs(want[0] / want[1])
