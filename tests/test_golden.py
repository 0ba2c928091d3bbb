"""Golden fixtures (tests/golden, written by make_golden.py): the oracle must
reproduce them exactly, and the library on the GPU must reproduce them to the
parity bar (R, C, r*, degree exact; V within 1e-6; EC within 1e-5 relative L2).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(GOLD / name, allow_pickle=False)


def test_oracle_reproduces_c1_small():
    f = load("c1_small.npz")
    x = f["x"]
    N = x.shape[1]
    z64, _, _, _, valid = orc.znorm(x, 0)
    g = orc.build_threshold(z64, valid, float(f["thr"]))
    assert np.array_equal(g.R, f["R"]) and np.array_equal(g.C, f["C"]) and np.array_equal(g.V64, f["V64"])
    d = orc.build_density(z64, valid, int(f["k"]))
    assert np.array_equal(d.R, f["dR"]) and np.array_equal(d.C, f["dC"]) and d.r_star == float(f["r_star"])
    assert orc.density_k(N, float(f["dens"])) == int(f["k"])
    assert np.array_equal(orc.degree(g.R, g.V, N), f["deg"])
    ec, lam, it = orc.eigcent(g.R, g.C, g.V, N, out_mode=0)
    assert np.array_equal(ec, f["ecw"]) and lam == float(f["lamw"]) and it == int(f["itw"])
    ec, lam, it = orc.eigcent(g.R, g.C, None, N, out_mode=1)
    assert np.array_equal(ec, f["ecb"]) and lam == float(f["lamb"]) and it == int(f["itb"])


def test_oracle_reproduces_ties():
    f = load("ties.npz")
    z64, _, _, _, valid = orc.znorm(f["x"], 0)
    d = orc.build_density(z64, valid, int(f["k"]))
    assert np.array_equal(d.R, f["R"]) and np.array_equal(d.C, f["C"])
    assert d.r_star == float(f["r_star"]) and d.n_ties == int(f["n_ties"]) and d.n_ties > 1


def test_oracle_ec_kat():
    kat = json.loads((GOLD / "ec_kat.json").read_text())
    R = np.array(kat["R"])
    C = np.array(kat["C"], np.int32)
    y, lam, it = orc.eigcent(R, C, None, 4, out_mode=1)
    np.testing.assert_allclose(y, kat["v"], atol=2e-6)
    assert lam == pytest.approx(kat["lambda"], abs=1e-5) and it == kat["iterations"]


@pytest.mark.gpu
def test_library_reproduces_golden():
    f = load("c1_small.npz")
    x = f["x"]
    N = x.shape[1]
    with bna.Context(0) as ctx:
        for prec in ("bf16x3", "bf16", "fp32"):
            g = ctx.build_csr(x, thr=float(f["thr"]), precision=prec)
            assert np.array_equal(g.R, f["R"]) and np.array_equal(g.C, f["C"]), prec
            assert np.abs(g.V.astype(np.float64) - f["V64"]).max(initial=0) <= 1e-6
            d = ctx.build_csr(x, density=float(f["dens"]), precision=prec)
            assert np.array_equal(d.R, f["dR"]) and np.array_equal(d.C, f["dC"]), prec
            assert d.thr_used == float(f["r_star"])
        g = ctx.build_csr(x, thr=float(f["thr"]))
        assert np.array_equal(ctx.degree(g), f["deg"])
        ec, lam, it = ctx.eigcent(g, output=bna.EC_WEIGHTED_REF)
        assert np.linalg.norm(ec - f["ecw"]) <= 1e-5 * np.linalg.norm(f["ecw"])
        ec, lam, it = ctx.eigcent(g, output=bna.EC_BINARY_REF, binary=True)
        assert np.linalg.norm(ec - f["ecb"]) <= 1e-5 * np.linalg.norm(f["ecb"])
        t = load("ties.npz")
        d = ctx.build_csr(t["x"], density=float(t["dens"]))
        assert np.array_equal(d.R, t["R"]) and np.array_equal(d.C, t["C"]) and d.thr_used == float(t["r_star"])
        kat = json.loads((GOLD / "ec_kat.json").read_text())
        kg = bna.Graph.__new__(bna.Graph)
        kg.N, kg.nnz = 4, len(kat["C"])
        kg.R, kg.C, kg.V = np.array(kat["R"], np.int64), np.array(kat["C"], np.int32), None
        y, lam, it = ctx.eigcent(kg, output=bna.EC_BINARY_REF)
        np.testing.assert_allclose(y, kat["v"], atol=2e-6)
        assert lam == pytest.approx(kat["lambda"], abs=1e-5) and it == kat["iterations"]
