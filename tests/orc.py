"""ctypes binding of the CPU oracle (oracle/bna_oracle.c) for the tests.

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "_build" / "libbna_oracle.so"
NTHREADS = min(16, os.cpu_count() or 1)

_o = None


def olib() -> C.CDLL:
    global _o
    if _o is not None:
        return _o
    if not ORACLE_SO.exists():
        raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
    L = C.CDLL(str(ORACLE_SO))
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "orc_cpu_has_fma": (i32, []),
        "orc_lp": (dbl, [vp, vp, i32, vp, i32]),
        "orc_bc": (None, [vp, vp, i32, vp]),
        "orc_lpw": (dbl, [vp, vp, vp, i32, i32, vp, vp, vp]),
        "orc_sssp64": (None, [vp, vp, vp, i32, vp]),
        "orc_znorm": (None, [vp, i32, i32, i32, vp, vp, vp, vp, vp]),
        "orc_dot64": (dbl, [vp, vp, i32]),
        "orc_corr_pairs": (None, [vp, i32, vp, vp, i64, vp]),
        "orc_ref_corr_rows": (None, [vp, i32, i32, i32, i32, vp]),
        "orc_build_threshold": (vp, [vp, vp, i32, i32, dbl, i32]),
        "orc_density_k": (i64, [i32, C.c_float]),
        "orc_all_pairs": (None, [vp, vp, i32, i32, i32, vp]),
        "orc_select_packed": (vp, [vp, i32, i64, i32]),
        "orc_build_density": (vp, [vp, vp, i32, i32, i64, i32, i32]),
        "orc_graph_nnz": (i64, [vp]),
        "orc_graph_rstar": (dbl, [vp]),
        "orc_graph_ties": (i64, [vp]),
        "orc_graph_copy": (None, [vp, vp, vp, vp, vp]),
        "orc_graph_free": (None, [vp]),
        "orc_degree": (None, [vp, vp, i32, vp]),
        "orc_cp": (dbl, [vp, vp, vp, i32, vp]),
        "orc_eigcent": (dbl, [vp, vp, vp, i32, dbl, i32, i32, vp, C.POINTER(C.c_int)]),
        "orc_ref_cpu_rows": (i64, [vp, i32, i32, i32, i32, C.c_float, i32, C.POINTER(C.c_double)]),
        "orc_fisher": (C.c_float, [C.c_float]),
        "orc_inv_fisher": (C.c_float, [C.c_float]),
        "orc_group_add": (None, [vp, vp, i64, i32]),
        "orc_group_finish": (None, [vp, i64, i32, i32]),
        "orc_threshold_packed": (vp, [vp, i32, dbl]),
        "orc_modularity": (i32, [vp, vp, i32, i32, dbl, vp, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                 C.POINTER(C.c_longlong)]),
        "orc_modularity_first": (dbl, [vp, vp, i32, i32, dbl, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _o = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def znorm(x: np.ndarray, layout: int = 0):
    """x float32, (L, N) time-major (layout 0) or (N, L) voxel-major (layout 1).
    Returns z64 (N, L), z32 (N, L), mu, sd, valid."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    L, N = x.shape if layout == 0 else x.shape[::-1]
    z64 = np.empty((N, L), np.float64)
    z32 = np.empty((N, L), np.float32)
    mu = np.empty(N, np.float64)
    sd = np.empty(N, np.float64)
    valid = np.empty(N, np.uint8)
    olib().orc_znorm(_p(x), N, L, layout, _p(mu), _p(sd), _p(z64), _p(z32), _p(valid))
    return z64, z32, mu, sd, valid


def dot64(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return olib().orc_dot64(_p(a), _p(b), a.shape[0])


class OGraph:
    def __init__(self, h):
        L = olib()
        if not h:
            raise ValueError("oracle returned no graph")
        n = L.orc_graph_nnz(h)
        # N from the handle is not exported; callers pass it via attribute below
        self._h = h
        self.nnz = n
        self.r_star = L.orc_graph_rstar(h)
        self.n_ties = L.orc_graph_ties(h)

    def load(self, N: int):
        L = olib()
        self.N = N
        self.R = np.empty(N + 1, np.int64)
        self.C = np.empty(self.nnz, np.int32)
        self.V = np.empty(self.nnz, np.float32)
        self.V64 = np.empty(self.nnz, np.float64)
        L.orc_graph_copy(self._h, _p(self.R), _p(self.C), _p(self.V), _p(self.V64))
        L.orc_graph_free(self._h)
        self._h = None
        return self


def build_threshold(z64, valid, thr: float, nthreads: int = NTHREADS) -> OGraph:
    N, L = z64.shape
    h = olib().orc_build_threshold(_p(z64), _p(valid), N, L, thr, nthreads)
    return OGraph(h).load(N)


def density_k(N: int, spa: float) -> int:
    return int(olib().orc_density_k(N, spa))


def all_pairs(z64, valid, nthreads: int = NTHREADS) -> np.ndarray:
    N, L = z64.shape
    out = np.empty(N * (N - 1) // 2, np.float64)
    olib().orc_all_pairs(_p(z64), _p(valid), N, L, nthreads, _p(out))
    return out


def select_packed(packed: np.ndarray, N: int, k: int, literal: bool) -> OGraph:
    packed = np.ascontiguousarray(packed, np.float64)
    h = olib().orc_select_packed(_p(packed), N, k, int(literal))
    return OGraph(h).load(N)


def build_density(z64, valid, k: int, literal: bool = False, nthreads: int = NTHREADS) -> OGraph:
    N, L = z64.shape
    h = olib().orc_build_density(_p(z64), _p(valid), N, L, k, nthreads, int(literal))
    return OGraph(h).load(N)


def degree(R, V, N):
    R = np.ascontiguousarray(R, np.int64)
    V = None if V is None else np.ascontiguousarray(V, np.float32)
    deg = np.empty(N, np.float32)
    olib().orc_degree(_p(R), _p(V), N, _p(deg))
    return deg


def cp(R, Cc, V, N):
    """Clustering coefficient per vertex and the float mean of |Cp| (orc_cp)."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    V = None if V is None else np.ascontiguousarray(V, np.float32)
    out = np.empty(N, np.float32)
    mean = olib().orc_cp(_p(R), _p(Cc), _p(V), N, _p(out))
    return out, mean


def lp(R, Cc, N, hist_len=64):
    """(Lp as CUBFS_Lp sums it, ordered-pair count per distance) (orc_lp)."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    hist = np.zeros(hist_len, np.int64)
    v = olib().orc_lp(_p(R), _p(Cc), N, _p(hist), hist_len)
    return v, hist


def bc(R, Cc, N):
    """Brandes betweenness per vertex in double (orc_bc)."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    out = np.empty(N, np.float64)
    olib().orc_bc(_p(R), _p(Cc), N, _p(out))
    return out


def lpw(R, Cc, V, N, B=128, with_dist=False):
    """Weighted Lp as CUBFW_Lp sums it over blocked Floyd-Warshall distances
    (block size B): (Lp, nodal efficiency, isolated count[, distances])."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    V = None if V is None else np.ascontiguousarray(V, np.float32)
    eff = np.empty(N, np.float32)
    iso = C.c_int()
    dist = np.empty((N, N), np.float32) if with_dist else None
    v = olib().orc_lpw(_p(R), _p(Cc), _p(V), N, B, _p(eff), C.byref(iso), _p(dist))
    return (v, eff, iso.value, dist) if with_dist else (v, eff, iso.value)


def sssp64(R, Cc, V, N):
    """All-pairs Dijkstra over the float costs 1.0f / w with double path sums."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    V = None if V is None else np.ascontiguousarray(V, np.float32)
    out = np.empty((N, N), np.float64)
    olib().orc_sssp64(_p(R), _p(Cc), _p(V), N, _p(out))
    return out


def eigcent(R, Cc, V, N, tol=1e-6, max_iter=10000, out_mode=0):
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    V = None if V is None else np.ascontiguousarray(V, np.float32)
    ec = np.empty(N, np.float64)
    it = C.c_int()
    lam = olib().orc_eigcent(_p(R), _p(Cc), _p(V), N, tol, max_iter, out_mode, _p(ec), C.byref(it))
    return ec, lam, it.value


def modularity(R, Cc, N, max_iter=10000, eps=1e-6):
    """Newman leading-eigenvector partition of a binary graph: (labels from 1,
    Q, modules, rounds, power iterations)."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    lab = np.empty(N, np.int32)
    q = C.c_double()
    rounds = C.c_int()
    its = C.c_longlong()
    mods = olib().orc_modularity(_p(R), _p(Cc), N, max_iter, C.c_double(eps), _p(lab), C.byref(q),
                                 C.byref(rounds), C.byref(its))
    return lab, q.value, mods, rounds.value, its.value


def modularity_first(R, Cc, N, max_iter=10000, eps=1e-6):
    """Round 0 alone: (leading vector, eigenvalue after the shift)."""
    R = np.ascontiguousarray(R, np.int64)
    Cc = np.ascontiguousarray(Cc, np.int32)
    v = np.empty(N, np.float64)
    lam = olib().orc_modularity_first(_p(R), _p(Cc), N, max_iter, C.c_double(eps), _p(v))
    return v, lam


def ref_cpu_rows(z32, i0, i1, thr, nthreads=NTHREADS):
    N, L = z32.shape
    cs = C.c_double()
    e = olib().orc_ref_cpu_rows(_p(np.ascontiguousarray(z32, np.float32)), N, L, i0, i1, thr,
                                nthreads, C.byref(cs))
    return int(e), cs.value


def ref_corr_rows(z32, i0, i1):
    N, L = z32.shape
    n = sum(N - 1 - i for i in range(i0, i1))
    out = np.empty(n, np.float32)
    olib().orc_ref_corr_rows(_p(np.ascontiguousarray(z32, np.float32)), N, L, i0, i1, _p(out))
    return out


def fisher(r: float) -> float:
    return float(olib().orc_fisher(r))


def inv_fisher(z: float) -> float:
    return float(olib().orc_inv_fisher(z))


def group_average(packed_list, N: int, fisher: bool) -> np.ndarray:
    """post_block over each subject's packed r (float64, NaN for invalid pairs),
    then the average (float32 packed)."""
    M1 = N * (N - 1) // 2
    acc = np.zeros(M1, np.float32)
    for p in packed_list:
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == (M1,)
        olib().orc_group_add(acc.ctypes.data, p.ctypes.data, M1, 1 if fisher else 0)
    olib().orc_group_finish(acc.ctypes.data, M1, len(packed_list), 1 if fisher else 0)
    return acc


def threshold_packed(packed: np.ndarray, N: int, thr: float) -> OGraph:
    packed = np.ascontiguousarray(packed, dtype=np.float64)
    return OGraph(olib().orc_threshold_packed(packed.ctypes.data, N, thr)).load(N)


def rows_above(z64, rows, thr, band=1e-9):
    """For each row i of `rows`: (the columns j != i with r64(i, j) > thr, the
    columns with r64(i, j) == thr).  r comes from a float64 product of z64 (its
    order differs from the oracle's by ~1e-13 at L = 1200); every pair within
    `band` of thr is recomputed in the oracle's order, so the decisions are the
    oracle's."""
    z64 = np.ascontiguousarray(z64, np.float64)
    N, L = z64.shape
    rows = np.asarray(rows, np.int64)
    r = z64[rows] @ z64.T
    out = []
    for a, i in enumerate(rows):
        ri = r[a]
        near = np.nonzero(np.abs(ri - thr) <= band)[0].astype(np.int32)
        if near.size:
            ii = np.full(near.size, i, np.int32)
            ex = np.empty(near.size)
            olib().orc_corr_pairs(_p(z64), L, _p(ii), _p(near), near.size, _p(ex))
            ri[near] = ex
        ri[i] = -np.inf
        out.append((np.nonzero(ri > thr)[0], np.nonzero(ri == thr)[0], ri))
    return out


class RefTimes(C.Structure):
    _fields_ = [("t_norm", C.c_double), ("t_corr", C.c_double), ("t_thr", C.c_double),
                ("t_csr", C.c_double), ("t_deg", C.c_double), ("t_ec", C.c_double),
                ("pairs", C.c_int64), ("edges", C.c_int64), ("nnz", C.c_int64),
                ("ec_iters", C.c_int), ("r_star", C.c_double), ("lam", C.c_double)]


def ref_pipeline(x, layout=0, row0=0, row1=None, thr=0.0, k=0, nthreads=NTHREADS) -> dict:
    """The reference CPU path end to end (orc_ref_pipeline): stage seconds,
    pairs, edges, nnz, r*, and for a whole subject the centrality's iterations
    and eigenvalue.  k > 0: density selection of k edges; else r > thr."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    L, N = x.shape if layout == 0 else x.shape[::-1]
    row1 = N if row1 is None else row1
    t = RefTimes()
    L_ = olib()
    L_.orc_ref_pipeline.restype = C.c_int64
    L_.orc_ref_pipeline.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_int64, C.c_int, C.POINTER(RefTimes)]
    e = L_.orc_ref_pipeline(_p(x), N, L, layout, row0, row1, thr, k, nthreads, C.byref(t))
    if e < 0:
        raise ValueError("orc_ref_pipeline: bad arguments")
    return {f: getattr(t, f) for f, _ in RefTimes._fields_}
