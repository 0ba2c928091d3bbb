"""Writes c2_density.json: the FP64 oracle's 1 % density graph of the C2 shape.

    python tests/golden/make_c2_golden.py [threads]

The input is the seeded factor-model series the GPU test generates (N = 45000,
L = 240, K = 20, phi = 0.6, seed 0xB17A0000 + 2000).  The oracle computes all
1,012,477,500 pair correlations in its fixed FP64 order, selects the
k = 10,124,775 largest with the reference's tie rule and forms the CSR; the
fixture keeps r*, the tie count, nnz and SHA-256 digests of R (int64) and C
(int32), so the GPU test can check its CSR bit for bit without the oracle's
several minutes of CPU work.
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import cuda_newest20160331_amd as bna  # noqa: E402
import orc  # noqa: E402

N, L, K, PHI, SEED, DENSITY = 45000, 240, 20, 0.6, 0xB17A0000 + 2000, 1.0


def main() -> None:
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    x = bna.synth_factor_model(N, L, K, PHI, SEED)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(N, DENSITY)
    t0 = time.time()
    g = orc.build_density(z64, valid, k, nthreads=threads)
    dt = time.time() - t0
    out = {
        "N": N, "L": L, "K": K, "phi": PHI, "seed": SEED, "density_pct": DENSITY, "k": k,
        "nnz": int(g.nnz), "r_star": float(g.r_star), "r_star_hex": float(g.r_star).hex(),
        "n_ties": int(g.n_ties),
        "sha256_R_int64": hashlib.sha256(np.ascontiguousarray(g.R, np.int64).tobytes()).hexdigest(),
        "sha256_C_int32": hashlib.sha256(np.ascontiguousarray(g.C, np.int32).tobytes()).hexdigest(),
        "V64_min": float(g.V64.min()), "V64_max": float(g.V64.max()),
        "oracle_seconds": round(dt, 1), "oracle_threads": threads,
    }
    (HERE / "c2_density.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
