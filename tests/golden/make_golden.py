"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture holds a seeded input and the oracle's outputs for it:
  c1_small.npz   N = 240, L = 200 factor-model series (K = 8, phi = 0.7): the
                 threshold graph r > 0.5 (R, C, V64), the 1 % density graph
                 (R, C, r*), degree, weighted and binary eigenvector centrality
  ties.npz       duplicated voxels (exact r ties) under density 1.3 %: R, C, r*
                 and the number of ties at r* (the row-major tie rule decides)
  ec_kat.json    the reference's in-tree EC test graph and its answer
                 (B/src/cuda_EC/cuda_EC/ecGPU.cu:217-233)
The reference ships no fixture of its own for this path; these pin the oracle
against regressions and are the common target of the oracle and GPU tests.
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import cuda_newest20160331_amd as bna  # noqa: E402
import orc  # noqa: E402


def c1_small():
    N, L = 240, 200
    x = bna.synth_factor_model(N, L, 8, 0.7, 0xB17A0000 + 1001)
    x[:, 7] = 2.5  # one constant voxel
    z64, _, _, _, valid = orc.znorm(x, 0)
    g = orc.build_threshold(z64, valid, 0.5)
    k = orc.density_k(N, 1.0)
    d = orc.build_density(z64, valid, k)
    deg = orc.degree(g.R, g.V, N)
    ecw, lamw, itw = orc.eigcent(g.R, g.C, g.V, N, out_mode=0)
    ecb, lamb, itb = orc.eigcent(g.R, g.C, None, N, out_mode=1)
    np.savez_compressed(HERE / "c1_small.npz", x=x, thr=0.5, R=g.R, C=g.C, V64=g.V64, dens=1.0, k=k,
                        dR=d.R, dC=d.C, dV64=d.V64, r_star=d.r_star, deg=deg, ecw=ecw, lamw=lamw, itw=itw,
                        ecb=ecb, lamb=lamb, itb=itb)


def ties():
    rng = np.random.default_rng(5)
    base = bna.synth_factor_model(60, 80, 4, 0.6, 9)
    x = np.ascontiguousarray(base[:, rng.integers(0, 60, size=300)])
    z64, _, _, _, valid = orc.znorm(x, 0)
    N = x.shape[1]
    k = orc.density_k(N, 1.3)
    d = orc.build_density(z64, valid, k)
    np.savez_compressed(HERE / "ties.npz", x=x, dens=1.3, k=k, R=d.R, C=d.C, r_star=d.r_star, n_ties=d.n_ties)


def ec_kat():
    R = [0, 2, 3, 6, 8]
    C = [2, 3, 2, 0, 1, 3, 0, 2]
    kat = {"source": "B/src/cuda_EC/cuda_EC/ecGPU.cu:217-233", "R": R, "C": C,
           "lambda": 2.170086486626, "v": [0.5227207, 0.2818452, 0.6116285, 0.5227207],
           "iterations": 32}
    (HERE / "ec_kat.json").write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == "__main__":
    c1_small()
    ties()
    ec_kat()
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))
