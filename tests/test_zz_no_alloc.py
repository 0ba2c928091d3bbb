"""bna_eigcent_noalloc, the entry point that allocates nothing, against
bna_eigcent on the same device CSRs, through the C ABI.

Two calls in a row on a context that never built anything (so it owns no
workspace), the second on a smaller graph.  The no-alloc path runs the plain
L2-gather SpMV while bna_eigcent copies these graphs (16 or more entries per
row) into its blocked layout, so the two sum each row in a different order:
the iteration counts must be equal and the eigenvector and lambda must agree
within 1e-12 relative (FP64 rounding; the bar for eigenvector centrality is
1e-5 relative L2).  The context's workspace must still be empty afterwards.
Sizes beyond the static scratch and host CSRs are refused on the host before
any launch.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import cuda_newest20160331_amd as bna

pytestmark = pytest.mark.gpu

MAX_N = 1048576  # BNA_NOALLOC_EC_MAX_N


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def device_doubles(h, n):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), 8 * n) == 0
    return p


def download(h, p, n):
    out = np.empty(n, np.float64)
    assert h.hipMemcpy(out.ctypes.data, p, 8 * n, 2) == 0  # hipMemcpyDeviceToHost
    return out


def test_noalloc_eigcent_equals_eigcent_twice_second_smaller():
    h = hip()
    builder = bna.Context(0)
    user = bna.Context(0)  # builds nothing: no workspace of its own
    cases = [(5000, 150, dict(thr=0.3), bna.EC_WEIGHTED_REF),
             (2100, 96, dict(thr=0.25, weighted=False), bna.EC_BINARY_REF)]
    keep = []
    try:
        assert user.workspace_bytes() == 0
        for seed, (N, L, kw, mode) in enumerate(cases):
            say(f"no-alloc eigcent: N={N} L={L} {kw} output={mode}")
            x = bna.synth_factor_model(N, L, 12, 0.6, 7100 + seed, bna.TIME_MAJOR)
            s = bna.Series(builder, x)
            g = s.build_device(**kw)
            keep.append((s, g))
            assert g.on_device == 1 and g.nnz > 0
            ec_ref, lam_ref, it_ref = builder.eigcent_device(g, output=mode)
            ec_dev = device_doubles(h, N)
            try:
                lam, it = user.eigcent_noalloc(g, ec_dev.value, output=mode)
                ec = download(h, ec_dev, N)
            finally:
                h.hipFree(ec_dev)
            say(f"  nnz={g.nnz} iterations {it} (eigcent {it_ref}) lambda {lam!r} ({lam_ref!r})")
            rel = np.linalg.norm(ec - ec_ref) / np.linalg.norm(ec_ref)
            say(f"  rel L2 vs eigcent {rel:.2e}")
            assert it == it_ref and it > 0
            assert abs(lam - lam_ref) <= 1e-12 * abs(lam_ref)
            assert rel <= 1e-12
            assert user.workspace_bytes() == 0
        # refused on the host, before any launch
        s, g = keep[0]
        host = bna.CSR()
        host.N, host.nnz, host.R, host.C, host.V, host.on_device = g.N, g.nnz, g.R, g.C, g.V, 0
        with pytest.raises(bna.BnaError, match="device CSR"):
            user.eigcent_noalloc(host, 1 << 20)
        big = bna.CSR()
        big.N, big.nnz, big.R, big.C, big.V, big.on_device = MAX_N + 1, g.nnz, g.R, g.C, g.V, 1
        with pytest.raises(bna.BnaError, match="exceeds the scratch maximum"):
            user.eigcent_noalloc(big, 1 << 20)
        with pytest.raises(bna.BnaError, match="NULL"):
            user.eigcent_noalloc(g, 0)
    finally:
        for s, g in keep:
            s.free_device(g)
            s.close()
        user.close()
        builder.close()
