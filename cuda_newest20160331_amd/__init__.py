"""cuda_newest20160331_amd -- MI355X network construction for BNAPlatform.

The product is libbna.so (HIP kernels for gfx950 behind the C ABI declared in
include/bna.h) and the command-line tools bna_cormat / bna_degree / bna_ec.
This module is a thin ctypes binding to that library, used by the tests and
the benchmark; every call goes through the C ABI.  There is no fallback: if the
library is missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libbna.so"

TIME_MAJOR, VOXEL_MAJOR = 0, 1
FP32, BF16, BF16X3 = 0, 1, 2
THR_R, THR_DENSITY = 0, 1
EC_WEIGHTED_REF, EC_BINARY_REF = 0, 1
# bna_status
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_DEVICE_NOMEM, ERR_IO, ERR_FORMAT, ERR_OVERFLOW = 0, -1, -2, -3, -4, -5, -6, -7
PRECISIONS = {"fp32": FP32, "bf16": BF16, "bf16x3": BF16X3}

# every function declared in include/bna.h
EXPORTS = [
    "bna_opts_default", "bna_csr_free", "bna_create", "bna_destroy", "bna_synchronize",
    "bna_status_string", "bna_last_error", "bna_get_timings", "bna_workspace_bytes",
    "bna_build_csr", "bna_series_create", "bna_series_destroy", "bna_series_normalize",
    "bna_build_csr_series", "bna_csr_to_host", "bna_series_download", "bna_degree", "bna_cp",
    "bna_eigcent", "bna_density_to_k", "bna_synth_factor_model", "bna_write_csr",
    "bna_read_csr", "bna_write_nm", "bna_nifti_mask", "bna_nifti_bold", "bna_free",
    "bna_build_csr_batch", "bna_trim", "bna_server_config_default", "bna_server_create",
    "bna_server_destroy", "bna_submit", "bna_poll", "bna_wait", "bna_cancel", "bna_server_get_stats",
    "bna_result_free", "bna_request_footprint", "bna_memory_info", "bna_reserve",
    "bna_csr_release", "bna_set_chunking", "bna_set_grid_limit", "bna_grid_size", "bna_autotune",
    "bna_autotune_load", "bna_autotune_get", "bna_autotune_set", "bna_group_create",
    "bna_group_destroy", "bna_group_add", "bna_group_count", "bna_group_finish", "bna_group_download",
    "bna_group_set_exact", "bna_group_band_pairs", "bna_lp", "bna_bc", "bna_lp_weighted",
    "bna_group_build_csr", "bna_write_cormat", "bna_nm_to_nifti", "bna_corr_packed", "bna_select_density",
    "bna_build_csr_into", "bna_set_graphs", "bna_batch_graph_stats", "bna_eigcent_noalloc",
    "bna_modularity", "bna_check_csr", "bna_maslov",
]


class Opts(C.Structure):
    _fields_ = [
        ("thr_type", C.c_int), ("thr", C.c_double), ("weighted", C.c_int),
        ("precision", C.c_int), ("recheck", C.c_int), ("ref_compat_eps", C.c_int),
        ("tile", C.c_int), ("refine", C.c_int),
    ]


class CSR(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("nnz", C.c_int64), ("R", C.POINTER(C.c_int32)),
        ("C", C.POINTER(C.c_int32)), ("V", C.POINTER(C.c_float)), ("on_device", C.c_int),
        ("thr_used", C.c_double), ("density_pct", C.c_double), ("n_candidates", C.c_int64),
        ("n_flips", C.c_int64), ("band", C.c_double), ("max_err", C.c_double),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("znorm_ms", C.c_double), ("corr_ms", C.c_double), ("recheck_ms", C.c_double),
        ("csr_ms", C.c_double), ("refine_ms", C.c_double), ("select_ms", C.c_double),
        ("total_ms", C.c_double),
        ("corr_launches", C.c_int64),
        ("degree_ms", C.c_double), ("ec_ms", C.c_double), ("ec_iters", C.c_int64), ("cp_ms", C.c_double),
        ("modularity_ms", C.c_double), ("modularity_iters", C.c_int64),
        ("lp_ms", C.c_double), ("bc_ms", C.c_double), ("lpw_ms", C.c_double),
    ]


_lib = None


def lib() -> C.CDLL:
    """Loads libbna.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m cuda_newest20160331_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64, dbl, sz = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
    fp = C.POINTER(C.c_float)
    sig = {
        "bna_opts_default": (None, [C.POINTER(Opts)]),
        "bna_csr_free": (None, [C.POINTER(CSR)]),
        "bna_create": (i32, [C.POINTER(vp), i32, sz]),
        "bna_destroy": (i32, [vp]),
        "bna_synchronize": (i32, [vp]),
        "bna_status_string": (C.c_char_p, [i32]),
        "bna_last_error": (C.c_char_p, [vp]),
        "bna_get_timings": (i32, [vp, C.POINTER(Timings)]),
        "bna_workspace_bytes": (i64, [vp]),
        "bna_build_csr": (i32, [vp, fp, i32, i32, i32, C.POINTER(Opts), C.POINTER(CSR)]),
        "bna_series_create": (i32, [vp, vp, i32, i32, i32, i32, C.POINTER(vp)]),
        "bna_series_destroy": (i32, [vp, vp]),
        "bna_series_normalize": (i32, [vp, vp, i32]),
        "bna_build_csr_series": (i32, [vp, vp, C.POINTER(Opts), C.POINTER(CSR)]),
        "bna_csr_to_host": (i32, [vp, C.POINTER(CSR), C.POINTER(CSR)]),
        "bna_series_download": (i32, [vp, vp, vp, vp, vp]),
        "bna_degree": (i32, [vp, C.POINTER(CSR), fp]),
        "bna_cp": (i32, [vp, C.POINTER(CSR), fp, C.POINTER(C.c_double)]),
        "bna_eigcent": (i32, [vp, C.POINTER(CSR), dbl, i32, i32, C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "bna_eigcent_noalloc": (i32, [vp, C.POINTER(CSR), dbl, i32, i32, vp,
                                      C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "bna_modularity": (i32, [vp, C.POINTER(CSR), C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]),
        "bna_check_csr": (i32, [C.POINTER(CSR), i32]),
        "bna_maslov": (i32, [C.POINTER(CSR), i32, C.c_uint64, C.POINTER(CSR)]),
        "bna_density_to_k": (i64, [i32, C.c_float]),
        "bna_synth_factor_model": (i32, [fp, i32, i32, i32, dbl, C.c_uint64, i32]),
        "bna_write_csr": (i32, [C.c_char_p, C.POINTER(CSR)]),
        "bna_read_csr": (i32, [C.c_char_p, C.POINTER(CSR)]),
        "bna_write_nm": (i32, [C.c_char_p, fp, i32]),
        "bna_nifti_mask": (i32, [C.c_char_p, C.c_float, C.POINTER(C.POINTER(C.c_uint8)),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
        "bna_nifti_bold": (i32, [C.c_char_p, C.POINTER(C.c_uint8), i64, i32,
                                 C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int)]),
        "bna_free": (None, [vp]),
        "bna_build_csr_batch": (i32, [vp, C.POINTER(vp), C.POINTER(Opts), i32, C.POINTER(CSR),
                                      C.POINTER(C.c_int)]),
        "bna_trim": (i32, [vp]),
        "bna_csr_release": (i32, [vp, C.POINTER(CSR)]),
        "bna_autotune": (i32, [vp, i32, i32, i32, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "bna_autotune_load": (i32, [vp, C.c_char_p]),
        "bna_autotune_get": (i32, [vp, i32, i32, C.POINTER(C.c_int)]),
        "bna_autotune_set": (i32, [vp, i32, i32, i32]),
        "bna_group_create": (i32, [vp, i32, i32, C.POINTER(vp)]),
        "bna_group_destroy": (i32, [vp, vp]),
        "bna_group_add": (i32, [vp, vp, vp, i32]),
        "bna_group_count": (i32, [vp]),
        "bna_group_set_exact": (i32, [vp, vp, i32]),
        "bna_lp": (i32, [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), i32]),
        "bna_bc": (i32, [vp, vp, fp]),
        "bna_lp_weighted": (i32, [vp, vp, C.POINTER(C.c_double), fp, C.POINTER(C.c_int)]),
        "bna_group_band_pairs": (i64, [vp]),
        "bna_group_finish": (i32, [vp, vp]),
        "bna_group_download": (i32, [vp, vp, fp]),
        "bna_group_build_csr": (i32, [vp, vp, C.POINTER(Opts), C.POINTER(CSR)]),
        "bna_write_cormat": (i32, [C.c_char_p, fp, i32]),
        "bna_nm_to_nifti": (i32, [C.c_char_p, dbl, fp, i32, C.c_char_p]),
        "bna_corr_packed": (i32, [vp, vp, i32, fp]),
        "bna_build_csr_into": (i32, [vp, vp, C.POINTER(Opts), vp, vp, vp, i64, C.POINTER(C.c_int64)]),
        "bna_select_density": (i32, [vp, vp, i32, dbl, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "bna_set_graphs": (i32, [vp, i32]),
        "bna_batch_graph_stats": (i32, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class BnaError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def nifti_mask(path: str, mask_thr: float) -> tuple[np.ndarray, int]:
    """Voxel selection of a NIfTI mask (mask >= mask_thr): (selected uint8[total], N)."""
    p = C.POINTER(C.c_uint8)()
    total, n = C.c_int64(), C.c_int()
    rc = lib().bna_nifti_mask(str(path).encode(), mask_thr, C.byref(p), C.byref(total), C.byref(n))
    if rc:
        raise BnaError(rc, f"bna_nifti_mask({path})")
    try:
        sel = np.ctypeslib.as_array(p, shape=(total.value,)).copy()
    finally:
        lib().bna_free(C.cast(p, C.c_void_p))
    return sel, n.value


def nifti_bold(path: str, selected: np.ndarray) -> np.ndarray:
    """Selected voxels of a NIfTI series, time-major (L, N) float32."""
    selected = np.ascontiguousarray(selected, np.uint8)
    N = int(selected.sum())
    p = C.POINTER(C.c_float)()
    L = C.c_int()
    rc = lib().bna_nifti_bold(str(path).encode(), selected.ctypes.data_as(C.POINTER(C.c_uint8)),
                              selected.size, N, C.byref(p), C.byref(L))
    if rc:
        raise BnaError(rc, f"bna_nifti_bold({path})")
    try:
        x = np.ctypeslib.as_array(p, shape=(L.value, N)).copy()
    finally:
        lib().bna_free(C.cast(p, C.c_void_p))
    return x


def read_csr(path: str) -> "Graph":
    """A .csr file as a host Graph (V is None for a binary file)."""
    h = CSR()
    rc = lib().bna_read_csr(str(path).encode(), C.byref(h))
    if rc:
        raise BnaError(rc, f"bna_read_csr({path})")
    try:
        return Graph(h)
    finally:
        lib().bna_csr_free(C.byref(h))


def check_csr(g: "Graph", ascending: bool = True) -> int:
    """bna_check_csr on a host Graph: 0 (BNA_OK) or the error status."""
    h = _host_csr(g)
    return lib().bna_check_csr(C.byref(h[0]), int(ascending))


def maslov(g: "Graph", seed: int, permute_weights: bool = False) -> "Graph":
    """A degree-preserving random network of a host Graph (bna_maslov): the
    Maslov rewiring the Cp and Lp tools compare against."""
    h = _host_csr(g)
    out = CSR()
    rc = lib().bna_maslov(C.byref(h[0]), int(permute_weights), C.c_uint64(seed & (2**64 - 1)), C.byref(out))
    if rc:
        raise BnaError(rc, "bna_maslov")
    try:
        return Graph(out)
    finally:
        lib().bna_csr_free(C.byref(out))


def read_nm(path: str) -> np.ndarray:
    """A .nm node map: int32 N then float32[N]."""
    raw = np.fromfile(path, dtype=np.int32)
    n = int(raw[0])
    if raw.size != n + 1:
        raise ValueError(f"{path}: {raw.size - 1} values for N={n}")
    return raw[1:].view(np.float32).copy()


def synth_factor_model(N: int, L: int, K: int, phi: float, seed: int,
                       layout: int = TIME_MAJOR) -> np.ndarray:
    """Seeded factor-model series (float32); shape (L, N) time-major or (N, L)."""
    x = np.empty((L, N) if layout == TIME_MAJOR else (N, L), dtype=np.float32)
    rc = lib().bna_synth_factor_model(_f32p(x), N, L, K, phi, seed, layout)
    if rc:
        raise BnaError(rc, "bna_synth_factor_model failed")
    return x


def density_to_k(N: int, spa_percent: float) -> int:
    return int(lib().bna_density_to_k(N, spa_percent))


def make_opts(thr: float = 0.0, *, density: float | None = None, weighted: bool = True,
              precision: int | str = BF16X3, recheck: bool = True,
              ref_compat_eps: bool = False, refine: bool = True) -> Opts:
    o = Opts()
    lib().bna_opts_default(C.byref(o))
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    if density is not None:
        o.thr_type, o.thr = THR_DENSITY, float(density)
    else:
        o.thr_type, o.thr = THR_R, float(thr)
    o.weighted = int(bool(weighted))
    o.precision = int(precision)
    o.recheck = int(bool(recheck))
    o.ref_compat_eps = int(bool(ref_compat_eps))
    o.refine = int(bool(refine))
    return o


class Graph:
    """A host CSR copied out of a bna_csr (int64 R for convenience)."""

    def __init__(self, g: CSR):
        n, nnz = g.N, g.nnz
        self.N = n
        self.nnz = nnz
        self.R = np.ctypeslib.as_array(g.R, shape=(n + 1,)).astype(np.int64).copy()
        self.C = (np.ctypeslib.as_array(g.C, shape=(nnz,)).copy() if nnz
                  else np.zeros(0, np.int32))
        self.V = None
        if g.V:
            self.V = (np.ctypeslib.as_array(g.V, shape=(nnz,)).copy() if nnz
                      else np.zeros(0, np.float32))
        self.thr_used = g.thr_used
        self.density_pct = g.density_pct
        self.n_candidates = g.n_candidates
        self.n_flips = g.n_flips
        self.band = g.band
        self.max_err = g.max_err

    @classmethod
    def from_arrays(cls, R, Cc, V=None) -> "Graph":
        """A host graph from CSR arrays (R of N + 1 offsets, C, optional V)."""
        g = cls.__new__(cls)
        g.R = np.ascontiguousarray(R, np.int64)
        g.N = int(g.R.size - 1)
        g.C = np.ascontiguousarray(Cc, np.int32)
        g.nnz = int(g.C.size)
        g.V = None if V is None else np.ascontiguousarray(V, np.float32)
        g.thr_used = g.density_pct = g.band = g.max_err = 0.0
        g.n_candidates = g.n_flips = 0
        return g


class Context:
    """One bna_ctx on a HIP device."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._L = lib()
        self._h = C.c_void_p()
        rc = self._L.bna_create(C.byref(self._h), device, hbm_budget)
        if rc:
            raise BnaError(rc, self._L.bna_last_error(None).decode())

    def close(self):
        if self._h:
            self._L.bna_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            raise BnaError(rc, f"{what}: {self._L.bna_last_error(self._h).decode()}")

    def build_csr(self, x: np.ndarray, layout: int = TIME_MAJOR, **kw) -> Graph:
        x = np.ascontiguousarray(x, dtype=np.float32)
        if layout == TIME_MAJOR:
            L, N = x.shape
        else:
            N, L = x.shape
        o = make_opts(**kw)
        g = CSR()
        self._check(self._L.bna_build_csr(self._h, _f32p(x), N, L, layout, C.byref(o), C.byref(g)),
                    "bna_build_csr")
        try:
            return Graph(g)
        finally:
            self._L.bna_csr_free(C.byref(g))

    def workspace_bytes(self) -> int:
        """Device bytes the context's workspace holds (bna_workspace_bytes)."""
        return int(self._L.bna_workspace_bytes(self._h))

    def timings(self) -> dict:
        t = Timings()
        self._check(self._L.bna_get_timings(self._h, C.byref(t)), "bna_get_timings")
        return {k: getattr(t, k) for k, _ in Timings._fields_}

    # device-resident series
    def set_graphs(self, on: bool):
        self._check(self._L.bna_set_graphs(self._h, int(on)), "bna_set_graphs")

    def batch_graph_stats(self) -> tuple[int, int]:
        """(captures, replays) of the batched build's graph."""
        c, r = C.c_int64(), C.c_int64()
        self._check(self._L.bna_batch_graph_stats(self._h, C.byref(c), C.byref(r)), "bna_batch_graph_stats")
        return c.value, r.value

    def series(self, x: np.ndarray, layout: int = TIME_MAJOR) -> "Series":
        return Series(self, x, layout)

    def group(self, N: int, fisher: bool = False, exact: bool = True) -> "Group":
        return Group(self, N, fisher, exact)

    def degree(self, g: Graph) -> np.ndarray:
        h = _host_csr(g)
        deg = np.empty(g.N, np.float32)
        self._check(self._L.bna_degree(self._h, C.byref(h[0]), _f32p(deg)), "bna_degree")
        return deg

    def cp(self, g: Graph):
        """Clustering coefficient per vertex and the mean |Cp| (bna_cp)."""
        h = _host_csr(g)
        out = np.empty(g.N, np.float32)
        mean = C.c_double()
        self._check(self._L.bna_cp(self._h, C.byref(h[0]), _f32p(out), C.byref(mean)), "bna_cp")
        return out, mean.value

    def lp(self, g, hist_len: int = 0):
        """Characteristic path length of the binary graph (bna_lp); with hist_len,
        also the ordered-pair count per distance (index 0: pairs with no path).
        g: a Graph (host) or a CSR (device)."""
        h = (g, None) if isinstance(g, CSR) else _host_csr(g)
        lp = C.c_double()
        hist = np.zeros(max(hist_len, 0), np.int64)
        self._check(self._L.bna_lp(self._h, C.byref(h[0]), C.byref(lp),
                                   hist.ctypes.data_as(C.POINTER(C.c_int64)) if hist_len else None, hist_len),
                    "bna_lp")
        return (lp.value, hist) if hist_len else lp.value

    def lp_weighted(self, g):
        """Weighted characteristic path length (bna_lp_weighted): returns (lp, eff,
        isolated) -- lp over costs 1 / w by blocked Floyd-Warshall, eff the nodal
        efficiency per vertex, isolated the number of vertices whose efficiency
        is below 1 / N (set to 0).  g: Graph (host) or CSR (device)."""
        h = (g, None) if isinstance(g, CSR) else _host_csr(g)
        lp = C.c_double()
        eff = np.empty(h[0].N, np.float32)
        iso = C.c_int()
        self._check(self._L.bna_lp_weighted(self._h, C.byref(h[0]), C.byref(lp), _f32p(eff), C.byref(iso)),
                    "bna_lp_weighted")
        return lp.value, eff, iso.value

    def bc(self, g) -> np.ndarray:
        """Betweenness centrality per vertex (bna_bc); g: Graph or device CSR."""
        h = (g, None) if isinstance(g, CSR) else _host_csr(g)
        out = np.empty(h[0].N, np.float32)
        self._check(self._L.bna_bc(self._h, C.byref(h[0]), _f32p(out)), "bna_bc")
        return out

    def cp_device(self, g: CSR):
        """Clustering coefficient of a device-resident CSR."""
        out = np.empty(g.N, np.float32)
        mean = C.c_double()
        self._check(self._L.bna_cp(self._h, C.byref(g), _f32p(out), C.byref(mean)), "bna_cp")
        return out, mean.value

    def autotune(self, precision: str, L: int, N: int = 16384, cache: str | None = None):
        """Tunes the tile traversal band height for (precision, L's bucket);
        returns (band height, best tile-kernel ms)."""
        g = C.c_int()
        ms = C.c_double()
        self._check(self._L.bna_autotune(self._h, PRECISIONS[precision], L, N,
                                         cache.encode() if cache else None, C.byref(g), C.byref(ms)),
                    "bna_autotune")
        return g.value, ms.value

    def autotune_load(self, path: str):
        self._check(self._L.bna_autotune_load(self._h, path.encode()), "bna_autotune_load")

    def band_rows(self, precision: str, L: int) -> int:
        g = C.c_int()
        self._check(self._L.bna_autotune_get(self._h, PRECISIONS[precision], L, C.byref(g)), "bna_autotune_get")
        return g.value

    def set_band_rows(self, precision: str, L: int, rows: int):
        self._check(self._L.bna_autotune_set(self._h, PRECISIONS[precision], L, rows), "bna_autotune_set")

    def degree_device(self, g: CSR) -> np.ndarray:
        """Degree of a device-resident CSR (bna_csr from Series.build_device)."""
        deg = np.empty(g.N, np.float32)
        self._check(self._L.bna_degree(self._h, C.byref(g), _f32p(deg)), "bna_degree")
        return deg

    def eigcent_device(self, g: CSR, tol: float = 1e-6, max_iter: int = 10000,
                       output: int = EC_WEIGHTED_REF):
        """Eigenvector centrality of a device-resident CSR."""
        ec = np.empty(g.N, np.float64)
        lam = C.c_double()
        it = C.c_int()
        self._check(self._L.bna_eigcent(self._h, C.byref(g), tol, max_iter, output,
                                        ec.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lam),
                                        C.byref(it)), "bna_eigcent")
        return ec, lam.value, it.value

    def modularity(self, g: "Graph | CSR"):
        """bna_modularity: (labels numbered from 1, Q, module count) of a host
        Graph or a device-resident bna_csr (taken as binary)."""
        if isinstance(g, CSR):
            ref, N = g, g.N
        else:
            h = _host_csr(g, drop_values=True)
            ref, N = h[0], g.N
        lab = np.empty(N, np.int32)
        q = C.c_double()
        mods = C.c_int()
        self._check(self._L.bna_modularity(self._h, C.byref(ref), lab.ctypes.data_as(C.POINTER(C.c_int32)),
                                           C.byref(q), C.byref(mods)), "bna_modularity")
        return lab, q.value, mods.value

    def eigcent_noalloc(self, g: CSR, ec_dev: int, tol: float = 1e-6, max_iter: int = 10000,
                        output: int = EC_WEIGHTED_REF):
        """bna_eigcent_noalloc: device CSR in, N doubles written at device
        address ec_dev; returns (lambda, iterations)."""
        lam = C.c_double()
        it = C.c_int()
        self._check(self._L.bna_eigcent_noalloc(self._h, C.byref(g), tol, max_iter, output,
                                                C.c_void_p(ec_dev), C.byref(lam), C.byref(it)),
                    "bna_eigcent_noalloc")
        return lam.value, it.value

    def eigcent(self, g: Graph, tol: float = 1e-6, max_iter: int = 10000,
                output: int = EC_WEIGHTED_REF, binary: bool = False):
        h = _host_csr(g, drop_values=binary)
        ec = np.empty(g.N, np.float64)
        lam = C.c_double()
        it = C.c_int()
        self._check(self._L.bna_eigcent(self._h, C.byref(h[0]), tol, max_iter, output,
                                        ec.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lam),
                                        C.byref(it)), "bna_eigcent")
        return ec, lam.value, it.value


def _host_csr(g: Graph, drop_values: bool = False):
    """A bna_csr view of a Graph's arrays (kept alive by the returned tuple)."""
    R = np.ascontiguousarray(g.R, dtype=np.int32)
    Cc = np.ascontiguousarray(g.C, dtype=np.int32)
    V = None if (drop_values or g.V is None) else np.ascontiguousarray(g.V, dtype=np.float32)
    h = CSR()
    h.N = g.N
    h.nnz = g.nnz
    h.R = R.ctypes.data_as(C.POINTER(C.c_int32))
    h.C = Cc.ctypes.data_as(C.POINTER(C.c_int32))
    h.V = V.ctypes.data_as(C.POINTER(C.c_float)) if V is not None else C.POINTER(C.c_float)()
    h.on_device = 0
    return (h, R, Cc, V)


class Series:
    """A device-resident series (bna_series)."""

    def __init__(self, ctx: Context, x: np.ndarray, layout: int = TIME_MAJOR):
        self.ctx = ctx
        x = np.ascontiguousarray(x, dtype=np.float32)
        if layout == TIME_MAJOR:
            self.L, self.N = x.shape
        else:
            self.N, self.L = x.shape
        self._h = C.c_void_p()
        ctx._check(ctx._L.bna_series_create(ctx._h, x.ctypes.data, self.N, self.L, layout, 0,
                                            C.byref(self._h)), "bna_series_create")

    def close(self):
        if self._h:
            self.ctx._L.bna_series_destroy(self.ctx._h, self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def normalize(self, precision: int = BF16X3):
        self.ctx._check(self.ctx._L.bna_series_normalize(self.ctx._h, self._h, precision),
                        "bna_series_normalize")

    def download(self):
        z = np.empty((self.N, self.L), np.float64)
        mu = np.empty(self.N, np.float64)
        sd = np.empty(self.N, np.float64)
        self.ctx._check(self.ctx._L.bna_series_download(self.ctx._h, self._h, z.ctypes.data,
                                                        mu.ctypes.data, sd.ctypes.data),
                        "bna_series_download")
        return z, mu, sd

    def build_device(self, **kw) -> CSR:
        """Builds and keeps the CSR on the device; returns the raw bna_csr
        (release with free_device)."""
        o = make_opts(**kw)
        g = CSR()
        self.ctx._check(self.ctx._L.bna_build_csr_series(self.ctx._h, self._h, C.byref(o),
                                                         C.byref(g)), "bna_build_csr_series")
        return g

    def to_host(self, g: CSR) -> Graph:
        h = CSR()
        self.ctx._check(self.ctx._L.bna_csr_to_host(self.ctx._h, C.byref(g), C.byref(h)),
                        "bna_csr_to_host")
        try:
            return Graph(h)
        finally:
            self.ctx._L.bna_csr_free(C.byref(h))

    def free_device(self, g: CSR):
        self.ctx._L.bna_csr_release(self.ctx._h, C.byref(g))

    def build_into(self, R, Cc, V, **kw) -> int:
        """Builds into caller arrays (numpy int32 R[N+1], int32 C[cap], float32 V[cap]
        or None); R=None is a size query.  Returns nnz."""
        o = make_opts(**kw)
        nnz = C.c_int64()
        cap = 0 if Cc is None else int(Cc.size)
        self.ctx._check(self.ctx._L.bna_build_csr_into(
            self.ctx._h, self._h, C.byref(o), None if R is None else R.ctypes.data,
            None if Cc is None else Cc.ctypes.data, None if V is None else V.ctypes.data, cap, C.byref(nnz)),
            "bna_build_csr_into")
        return nnz.value

    def corr_packed(self, precision: str = "bf16x3") -> np.ndarray:
        """The packed r matrix (M1 floats; 0 outside [-1, 1] and for invalid voxels)."""
        out = np.empty(self.N * (self.N - 1) // 2, np.float32)
        self.ctx._check(self.ctx._L.bna_corr_packed(self.ctx._h, self._h, PRECISIONS[precision], _f32p(out)),
                        "bna_corr_packed")
        return out

    def select_density(self, density: float, precision: str = "bf16x3") -> tuple[float, int]:
        """(r*, k): the k-th largest r for a density in percent."""
        r = C.c_double()
        k = C.c_int64()
        self.ctx._check(self.ctx._L.bna_select_density(self.ctx._h, self._h, PRECISIONS[precision], density,
                                                       C.byref(r), C.byref(k)), "bna_select_density")
        return r.value, k.value

    def build(self, **kw) -> Graph:
        g = self.build_device(**kw)
        try:
            return self.to_host(g)
        finally:
            self.free_device(g)


class Group:
    """Group averaging of correlation matrices (bna_group): add subjects, finish,
    then threshold the averages into a graph."""

    def __init__(self, ctx: Context, N: int, fisher: bool = False, exact: bool = True):
        self.ctx = ctx
        self.N = N
        self.M1 = N * (N - 1) // 2
        self._h = C.c_void_p()
        ctx._check(ctx._L.bna_group_create(ctx._h, N, 1 if fisher else 0, C.byref(self._h)),
                   "bna_group_create")
        if not exact:
            ctx._check(ctx._L.bna_group_set_exact(ctx._h, self._h, 0), "bna_group_set_exact")

    @property
    def band_pairs(self) -> int:
        """Pairs the last graph build recomputed at the threshold."""
        return int(self.ctx._L.bna_group_band_pairs(self._h))

    def close(self):
        if self._h:
            self.ctx._L.bna_group_destroy(self.ctx._h, self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, s: "Series", precision: str = "bf16x3"):
        self.ctx._check(self.ctx._L.bna_group_add(self.ctx._h, self._h, s._h, PRECISIONS[precision]),
                        "bna_group_add")

    @property
    def count(self) -> int:
        return self.ctx._L.bna_group_count(self._h)

    def finish(self):
        self.ctx._check(self.ctx._L.bna_group_finish(self.ctx._h, self._h), "bna_group_finish")

    def packed(self) -> np.ndarray:
        out = np.empty(self.M1, np.float32)
        self.ctx._check(self.ctx._L.bna_group_download(self.ctx._h, self._h, _f32p(out)),
                        "bna_group_download")
        return out

    def build(self, thr: float = 0.0, *, density: float | None = None, weighted: bool = True) -> Graph:
        o = make_opts(thr, density=density, weighted=weighted)
        g = CSR()
        self.ctx._check(self.ctx._L.bna_group_build_csr(self.ctx._h, self._h, C.byref(o), C.byref(g)),
                        "bna_group_build_csr")
        h = CSR()
        try:
            self.ctx._check(self.ctx._L.bna_csr_to_host(self.ctx._h, C.byref(g), C.byref(h)),
                            "bna_csr_to_host")
        finally:
            self.ctx._L.bna_csr_release(self.ctx._h, C.byref(g))
        try:
            return Graph(h)
        finally:
            self.ctx._L.bna_csr_free(C.byref(h))


def build_batch_device(ctx: Context, series: list, opts: list) -> tuple[list, list]:
    """bna_build_csr_batch over Series objects: (device CSRs, per-subject status)."""
    n = len(series)
    hs = (C.c_void_p * n)(*[s._h.value for s in series])
    os_ = (Opts * n)(*opts)
    out = (CSR * n)()
    st = (C.c_int * n)()
    rc = ctx._L.bna_build_csr_batch(ctx._h, hs, os_, n, out, st)
    if rc and all(v == 0 for v in st):
        ctx._check(rc, "bna_build_csr_batch")
    return [out[q] for q in range(n)], list(st)


def build_batch(ctx: Context, series: list, opts: list) -> list:
    """Host Graphs (or a BnaError per failed subject) for a batch."""
    gs, st = build_batch_device(ctx, series, opts)
    res = []
    for g, rc in zip(gs, st):
        if rc:
            res.append(BnaError(rc, "subject failed"))
            continue
        h = CSR()
        ctx._check(ctx._L.bna_csr_to_host(ctx._h, C.byref(g), C.byref(h)), "bna_csr_to_host")
        try:
            res.append(Graph(h))
        finally:
            ctx._L.bna_csr_free(C.byref(h))
            ctx._L.bna_csr_release(ctx._h, C.byref(g))
    return res


# ------------------------------------------------------------------ serving
class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t), ("workers", C.c_int),
                ("max_batch", C.c_int), ("reserve_per_worker", C.c_size_t),
                ("big_round_ms", C.c_double), ("chunk_ms", C.c_double), ("reserve_eighths", C.c_int)]


class Request(C.Structure):
    _fields_ = [("x", C.c_void_p), ("N", C.c_int), ("L", C.c_int), ("layout", C.c_int),
                ("opts", Opts), ("want_degree", C.c_int), ("want_ec", C.c_int),
                ("device_result", C.c_int)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("error", C.c_char * 256), ("csr", CSR),
                ("degree", C.POINTER(C.c_float)), ("ec", C.POINTER(C.c_double)),
                ("lam", C.c_double), ("ec_iters", C.c_int), ("batch_size", C.c_int),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("upload_ms", C.c_double), ("build_ms", C.c_double), ("maps_ms", C.c_double),
                ("download_ms", C.c_double), ("build", Timings)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "submitted", "completed", "failed", "cancelled", "rounds", "queued",
        "max_concurrent_rounds", "max_batch_seen", "max_reserved", "budget")]


class Served:
    """A finished request: the graph, optional degree / EC, and its timings."""

    def __init__(self, r: Result, N: int, graph: bool = True):
        self.status = r.status
        self.error = r.error.decode(errors="replace")
        self.nnz = r.csr.nnz if r.status == 0 else 0
        self.on_device = bool(r.csr.on_device)
        self.graph = Graph(r.csr) if (r.status == 0 and graph and not r.csr.on_device) else None
        self.degree = np.ctypeslib.as_array(r.degree, shape=(N,)).copy() if r.degree else None
        self.ec = np.ctypeslib.as_array(r.ec, shape=(N,)).copy() if r.ec else None
        self.lam = r.lam
        self.ec_iters = r.ec_iters
        self.batch_size = r.batch_size
        self.queue_ms, self.run_ms, self.latency_ms = r.queue_ms, r.run_ms, r.latency_ms
        self.phases_ms = {"upload": r.upload_ms, "build": r.build_ms, "maps": r.maps_ms,
                          "download": r.download_ms}
        self.build = {k: getattr(r.build, k) for k, _ in Timings._fields_}


_SERVER_SIGS = None


def _server_lib():
    global _SERVER_SIGS
    L = lib()
    if _SERVER_SIGS is None:
        vp = C.c_void_p
        sigs = {
            "bna_server_config_default": (None, [C.POINTER(ServerConfig)]),
            "bna_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(vp)]),
            "bna_server_destroy": (C.c_int, [vp]),
            "bna_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(C.c_int64)]),
            "bna_poll": (C.c_int, [vp, C.c_int64, C.POINTER(Result)]),
            "bna_wait": (C.c_int, [vp, C.c_int64, C.POINTER(Result)]),
            "bna_cancel": (C.c_int, [vp, C.c_int64]),
            "bna_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
            "bna_result_free": (None, [C.POINTER(Result)]),
            "bna_request_footprint": (C.c_int64, [C.c_int, C.c_int, C.POINTER(Opts)]),
            "bna_memory_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _SERVER_SIGS = sigs
    return L


class Server:
    """bna_server: submit series, collect graphs; rounds are batched on the GPU."""

    BUSY, CANCELLED, NOT_FOUND = -9, -10, -8

    def __init__(self, device: int = 0, hbm_budget: int = 0, workers: int = 2, max_batch: int = 16,
                 reserve_per_worker: int = 0, big_round_ms: float | None = None,
                 chunk_ms: float | None = None, reserve_eighths: int | None = None):
        self._L = _server_lib()
        cfg = ServerConfig()
        self._L.bna_server_config_default(C.byref(cfg))
        cfg.device, cfg.hbm_budget, cfg.workers, cfg.max_batch = device, hbm_budget, workers, max_batch
        cfg.reserve_per_worker = reserve_per_worker
        if big_round_ms is not None:
            cfg.big_round_ms = big_round_ms
        if chunk_ms is not None:
            cfg.chunk_ms = chunk_ms
        if reserve_eighths is not None:
            cfg.reserve_eighths = reserve_eighths
        self._h = C.c_void_p()
        rc = self._L.bna_server_create(C.byref(cfg), C.byref(self._h))
        if rc:
            raise BnaError(rc, f"bna_server_create: {self._L.bna_last_error(None).decode()}")
        self._N = {}

    def close(self):
        if self._h:
            self._L.bna_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, x: np.ndarray, layout: int = TIME_MAJOR, degree: bool = False, ec: bool = False,
               device_result: bool = False, **kw) -> int:
        x = np.ascontiguousarray(x, dtype=np.float32)
        L, N = x.shape if layout == TIME_MAJOR else x.shape[::-1]
        r = Request(x.ctypes.data, N, L, layout, make_opts(**kw), int(degree), int(ec), int(device_result))
        rid = C.c_int64()
        rc = self._L.bna_submit(self._h, C.byref(r), C.byref(rid))
        if rc:
            raise BnaError(rc, "bna_submit")
        self._N[rid.value] = N
        return rid.value

    def _collect(self, fn, rid: int, graph: bool):
        r = Result()
        rc = fn(self._h, rid, C.byref(r))
        if rc == self.BUSY or rc == self.NOT_FOUND:
            return rc
        try:
            return Served(r, self._N.pop(rid), graph)
        finally:
            self._L.bna_result_free(C.byref(r))

    def poll(self, rid: int, graph: bool = True):
        """A Served, or Server.BUSY / Server.NOT_FOUND."""
        return self._collect(self._L.bna_poll, rid, graph)

    def wait(self, rid: int, graph: bool = True):
        return self._collect(self._L.bna_wait, rid, graph)

    def cancel(self, rid: int) -> int:
        return self._L.bna_cancel(self._h, rid)

    def stats(self) -> dict:
        s = ServerStats()
        self._L.bna_server_get_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


def request_footprint(N: int, L: int, **kw) -> int:
    o = make_opts(**kw)
    return int(_server_lib().bna_request_footprint(N, L, C.byref(o)))
