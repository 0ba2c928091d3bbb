"""Host-side checks of libbna that need no GPU: the library loads and exports
every entry point declared in include/bna.h, and the host utilities (density
rule, generator, file formats, argument errors) behave as the reference does."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
import orc

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "bna.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(bna_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    L = bna.lib()
    decl = declared_functions()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(L, name), f"{name} declared in bna.h but not exported"
    assert sorted(bna.EXPORTS) == decl


def test_density_to_k_matches_oracle_and_reference_count():
    assert bna.density_to_k(45000, 1.0) == 10124775
    for N, s in [(5000, 1.0), (8000, 1.0), (60000, 1.0), (150000, 1.0), (45000, 0.1), (777, 33.3)]:
        assert bna.density_to_k(N, s) == orc.density_k(N, s)


def test_synth_is_seeded_and_layout_consistent():
    a = bna.synth_factor_model(500, 64, 8, 0.7, 42, bna.TIME_MAJOR)
    b = bna.synth_factor_model(500, 64, 8, 0.7, 42, bna.VOXEL_MAJOR)
    c = bna.synth_factor_model(500, 64, 8, 0.7, 43, bna.TIME_MAJOR)
    assert a.shape == (64, 500) and b.shape == (500, 64)
    assert np.array_equal(a, b.T)
    assert not np.array_equal(a, c)
    assert np.array_equal(a, bna.synth_factor_model(500, 64, 8, 0.7, 42, bna.TIME_MAJOR))


def test_synth_factor_structure():
    # voxel variance = a1^2 + a2^2 + 1 with a ~ N(0, 0.25): E = 1.5; AR(1)
    # latents make lag-1 autocorrelation of a voxel about phi * (a1^2+a2^2)/var.
    x = bna.synth_factor_model(4000, 2000, 20, 0.6, 5, bna.VOXEL_MAJOR).astype(np.float64)
    var = x.var(axis=1)
    assert abs(var.mean() - 1.5) < 0.05
    assert var.min() > 0.8
    r = np.corrcoef(x[:300])
    off = r[np.triu_indices(300, 1)]
    # ~19.5% of pairs share a factor (1 - C(18,2)/C(20,2)); null pairs have
    # r ~ N(0, 1/sqrt(2000)), so |r| > 0.15 (7 sigma) only comes from sharing
    frac = np.mean(np.abs(off) > 0.15)
    assert 0.02 < frac < 0.195


def _host_graph(R, Cc, V):
    R = np.asarray(R, np.int32)
    Cc = np.asarray(Cc, np.int32)
    g = bna.CSR()
    g.N = len(R) - 1
    g.nnz = len(Cc)
    g.R = R.ctypes.data_as(C.POINTER(C.c_int32))
    g.C = Cc.ctypes.data_as(C.POINTER(C.c_int32))
    if V is not None:
        V = np.asarray(V, np.float32)
        g.V = V.ctypes.data_as(C.POINTER(C.c_float))
    g.on_device = 0
    return g, (R, Cc, V)


@pytest.mark.parametrize("weighted", [True, False])
def test_csr_file_round_trip_and_layout(tmp_path, weighted):
    R = [0, 2, 3, 4]
    Cc = [1, 2, 0, 0]
    V = [0.5, 0.25, 0.5, 0.25] if weighted else None
    g, keep = _host_graph(R, Cc, V)
    p = tmp_path / "a.csr"
    assert bna.lib().bna_write_csr(str(p).encode(), C.byref(g)) == 0
    raw = np.fromfile(p, dtype=np.int32)
    # int32 N+1, R, int32 nnz, C [, int32 nnz, float V]  (CorMat.cpp:353-360)
    exp = [4] + R + [4] + Cc
    if weighted:
        exp += [4] + np.asarray(V, np.float32).view(np.int32).tolist()
    assert raw.tolist() == exp
    h = bna.CSR()
    assert bna.lib().bna_read_csr(str(p).encode(), C.byref(h)) == 0
    G = bna.Graph(h)
    bna.lib().bna_csr_free(C.byref(h))
    assert G.R.tolist() == R and G.C.tolist() == Cc
    assert (G.V is None) == (not weighted)
    if weighted:
        assert G.V.tolist() == V


def test_read_csr_rejects_damaged_files(tmp_path):
    L = bna.lib()
    h = bna.CSR()
    p = tmp_path / "bad.csr"
    # V section length differs from C's (weighted-net.cu:283-284 calls it damaged)
    np.array([3, 0, 1, 2, 2, 1, 0, 1, 0], np.int32).tofile(p)
    assert L.bna_read_csr(str(p).encode(), C.byref(h)) == bna.BnaError(-6, "").status
    # R not ending at nnz
    np.array([3, 0, 1, 3, 2, 1, 0], np.int32).tofile(p)
    assert L.bna_read_csr(str(p).encode(), C.byref(h)) == -6
    # column out of range
    np.array([3, 0, 1, 2, 2, 1, 5], np.int32).tofile(p)
    assert L.bna_read_csr(str(p).encode(), C.byref(h)) == -6
    # truncated
    np.array([3, 0, 1], np.int32).tofile(p)
    assert L.bna_read_csr(str(p).encode(), C.byref(h)) == -6
    assert L.bna_read_csr(b"/nonexistent/x.csr", C.byref(h)) == -5


def test_nm_file(tmp_path):
    v = np.array([1.5, -2.0, 3.25], np.float32)
    p = tmp_path / "x_deg.nm"
    assert bna.lib().bna_write_nm(str(p).encode(), v.ctypes.data_as(C.POINTER(C.c_float)), 3) == 0
    raw = np.fromfile(p, dtype=np.int32)
    assert raw[0] == 3 and np.array_equal(raw[1:].view(np.float32), v)


def test_argument_errors_without_device():
    L = bna.lib()
    h = C.c_void_p()
    assert L.bna_create(None, 0, 0) == -1
    assert L.bna_create(C.byref(h), -1, 0) in (-1, -2)
    assert L.bna_synchronize(None) == -1
    assert L.bna_status_string(-7) == b"result exceeds the int32 CSR format"
    x = np.zeros(4, np.float32)
    assert L.bna_synth_factor_model(x.ctypes.data_as(C.POINTER(C.c_float)), 2, 2, 1, 0.5, 1, 0) == -1
    assert L.bna_synth_factor_model(x.ctypes.data_as(C.POINTER(C.c_float)), 2, 2, 2, 1.0, 1, 0) == -1


def test_check_csr_rejects_unsorted_and_duplicate_columns(tmp_path):
    # a .csr whose row 0 lists its columns out of order, as a file written by
    # another tool might; bna_cp needs ascending columns and refuses it
    R = [0, 3, 4, 5, 6]
    ok = [1, 2, 3, 0, 0, 0]
    for cols, want in ((ok, 0), ([2, 1, 3, 0, 0, 0], bna.ERR_FORMAT), ([1, 1, 3, 0, 0, 0], bna.ERR_FORMAT)):
        g, keep = _host_graph(R, cols, None)
        p = tmp_path / "u.csr"
        assert bna.lib().bna_write_csr(str(p).encode(), C.byref(g)) == 0
        back = bna.read_csr(p)
        assert bna.check_csr(back) == want
        assert bna.check_csr(back, ascending=False) == 0
    bad_range, keep = _host_graph([0, 1, 1], [5], None)
    assert bna.lib().bna_check_csr(C.byref(bad_range), 0) == bna.ERR_FORMAT
