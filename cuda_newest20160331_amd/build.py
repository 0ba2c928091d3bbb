"""Builds libbna.so (HIP kernels for gfx950 + the C ABI) and the command-line
tools from cuda_newest20160331_amd/csrc, in-tree, with hipcc.

    python -m cuda_newest20160331_amd.build          # library + tools
    python -m cuda_newest20160331_amd.build --lib    # library only
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# No floating-point contraction anywhere: the z-normalisation, the FP64 recheck
# and the degree sums must round exactly as written.  Packed FP32 instructions
# are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["k_znorm.hip", "k_corr.hip", "k_csr.hip", "k_select.hip", "k_graph.hip", "k_group.hip", "k_cp.hip", "k_modularity.hip", "k_paths.hip", "k_fw.hip", "api.hip", "host_util.cpp", "serve.cpp"]
TOOLS = {"bna_cormat": "cli_cormat.cpp", "bna_degree": "cli_degree.cpp", "bna_ec": "cli_ec.cpp",
         "bna_convert_nii": "cli_convert.cpp", "bna_cp": "cli_cp.cpp", "bna_modularity": "cli_modularity.cpp",
         "bna_lp": "cli_lp.cpp", "bna_bc": "cli_bc.cpp",
         "bna_lpw": "cli_lpw.cpp"}


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(
            ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln
        )
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{msg}")


def _compile(src: str) -> Path:
    s = CSRC / src
    o = OBJ / (src + ".o")
    if s.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(s), "-o", str(o)]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", str(s), "-o", str(o)]
    _run(cmd)
    return o


def build_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    out = LIBDIR / "libbna.so"
    tmp = LIBDIR / "libbna.so.tmp"
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_tools() -> list[Path]:
    BINDIR.mkdir(exist_ok=True)
    outs = []
    for name, src in TOOLS.items():
        s = CSRC / src
        if not s.exists():
            continue
        o = BINDIR / name
        _run([
            "g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(s), "-o", str(o),
            f"-L{LIBDIR}", "-lbna", f"-Wl,-rpath,$ORIGIN/../lib",
        ])
        outs.append(o)
    return outs


def build_oracle() -> Path:
    """The CPU oracle is test infrastructure; built beside the library so tests
    and the CPU baseline can load it."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "libbna_oracle.so"


def main(argv: list[str]) -> int:
    lib = build_lib()
    print(f"built {lib}")
    if "--lib" not in argv:
        for t in build_tools():
            print(f"built {t}")
        print(f"built {build_oracle()}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
