"""NIfTI + mask ingest (CorMat.cpp:141-238) and the command-line tools' argument
handling, without a GPU.

Closed form for the ingest: voxel k of a volume stored x-fastest is selected iff
mask[k] >= thr, selected voxels keep ascending k, and BOLD[l*N + i] is time point
l of the i-th selected voxel (doubles narrowed to float).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cuda_newest20160331_amd as bna
from nifti import write_nii

BIN = Path(bna.__file__).resolve().parent / "bin"


def _expected(data, mask, thr):
    X, Y, Z, T = data.shape
    flat = data.reshape(X * Y * Z, T, order="F")  # row k = voxel x + X*(y + Y*z)
    sel = mask.reshape(-1, order="F").astype(np.float32) >= np.float32(thr)
    return sel, np.ascontiguousarray(flat[sel].T.astype(np.float32))


@pytest.mark.parametrize("mask_dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("data_dtype", [np.float32, np.float64])
def test_nifti_ingest_matches_closed_form(tmp_path, mask_dtype, data_dtype):
    rng = np.random.default_rng(4)
    X, Y, Z, T = 7, 5, 4, 23
    mask = (rng.random((X, Y, Z)) * 2).astype(mask_dtype) if mask_dtype == np.uint8 else \
        rng.random((X, Y, Z)).astype(np.float32)
    data = rng.standard_normal((X, Y, Z, T)).astype(data_dtype)
    write_nii(tmp_path / "mask.nii", mask)
    write_nii(tmp_path / "s1.nii", data)
    thr = 0.5
    sel, N = bna.nifti_mask(str(tmp_path / "mask.nii"), thr)
    esel, ex = _expected(data, mask, thr)
    assert N == int(esel.sum()) and np.array_equal(sel.astype(bool), esel)
    x = bna.nifti_bold(str(tmp_path / "s1.nii"), sel)
    assert x.shape == (T, N)
    assert np.array_equal(x, ex)


def test_nifti_ingest_rejects_bad_files(tmp_path):
    rng = np.random.default_rng(1)
    mask = np.ones((4, 4, 2), np.uint8)
    write_nii(tmp_path / "mask.nii", mask)
    sel, N = bna.nifti_mask(str(tmp_path / "mask.nii"), 0.5)
    assert N == 32
    # volume size differs from the mask's
    write_nii(tmp_path / "other.nii", rng.standard_normal((4, 4, 3, 5)).astype(np.float32))
    with pytest.raises(bna.BnaError):
        bna.nifti_bold(str(tmp_path / "other.nii"), sel)
    # truncated data
    write_nii(tmp_path / "short.nii", rng.standard_normal((4, 4, 2, 5)).astype(np.float32))
    b = (tmp_path / "short.nii").read_bytes()
    (tmp_path / "short.nii").write_bytes(b[:-8])
    with pytest.raises(bna.BnaError):
        bna.nifti_bold(str(tmp_path / "short.nii"), sel)
    # series of 8-bit values (neither float nor double)
    write_nii(tmp_path / "u8.nii", np.ones((4, 4, 2, 5), np.uint8))
    with pytest.raises(bna.BnaError):
        bna.nifti_bold(str(tmp_path / "u8.nii"), sel)
    # mask of double values (only uint8 and float masks are read)
    write_nii(tmp_path / "m64.nii", np.ones((4, 4, 2), np.float64))
    with pytest.raises(bna.BnaError):
        bna.nifti_mask(str(tmp_path / "m64.nii"), 0.5)
    with pytest.raises(bna.BnaError):
        bna.nifti_mask(str(tmp_path / "missing.nii"), 0.5)


def _run(tool, *args):
    return subprocess.run([str(BIN / tool), *map(str, args)], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not (BIN / "bna_cormat").exists(), reason="tools not built")
def test_cli_argument_handling(tmp_path):
    r = _run("bna_cormat", tmp_path)
    assert r.returncode == 1 and "usage" in r.stderr
    write_nii(tmp_path / "mask.nii", np.ones((2, 2, 2), np.uint8))
    # malformed avg / save_cormat / threshold letters are refused before any device work
    r = _run("bna_cormat", tmp_path, tmp_path / "mask.nii", 0.5, "q", "n", "r", 0.3)
    assert r.returncode == 1 and "avg type" in r.stderr
    r = _run("bna_cormat", tmp_path, tmp_path / "mask.nii", 0.5, "y", "n", "r", 0.3)
    assert r.returncode == 1 and "second letter" in r.stderr
    r = _run("bna_cormat", tmp_path, 0.5, "n", "x", "r", 0.3)
    assert r.returncode == 1 and "to_save_cor_matrix" in r.stderr
    r = _run("bna_cormat", tmp_path, 0.5, "n", "n", "q", 0.3)
    assert r.returncode == 1 and "threshold type" in r.stderr
    r = _run("bna_cormat", "--precision=fp64", tmp_path, 0.5, "n", "n", "r", 0.3)
    assert r.returncode == 1 and "precision" in r.stderr
    r = _run("bna_degree", tmp_path / "nope")
    assert r.returncode == 1
    r = _run("bna_ec")
    assert r.returncode == 1 and "usage" in r.stderr


def test_request_footprint_and_server_without_device():
    a = bna.request_footprint(5000, 200, density=1.0)
    b = bna.request_footprint(10000, 200, density=1.0)
    c = bna.request_footprint(10000, 1200, density=1.0)
    assert 0 < a < b < c
    # series + operand panel + CSR at 1 % density are all inside the estimate
    N, L = 45000, 240
    est = bna.request_footprint(N, L, density=1.0)
    assert est > 4 * N * L + 8 * 2 * 0.01 * N * (N - 1) / 2
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(bna.BnaError):
            bna.Server()


@pytest.mark.parametrize("mask_dtype", [np.uint8, np.float32])
def test_convert_nii_scatters_maps_onto_the_mask(tmp_path, mask_dtype):
    # ConvertNII (W/ConvertNII/ConvertNII/ConvertNII.cpp:65-167): the inverse of
    # the ingest's gather.  A volume built by the tool and read back through
    # bna_nifti_bold with the same mask gives the map again; voxels outside the
    # mask are 0; the header is the mask's with float data.
    rng = np.random.default_rng(9)
    X, Y, Z = 7, 6, 5
    if mask_dtype == np.uint8:
        mask = (rng.random((X, Y, Z)) < 0.6).astype(np.uint8)
        thr = 0.5
    else:
        mask = rng.random((X, Y, Z)).astype(np.float32)
        thr = 0.4
    write_nii(tmp_path / "mask.nii", mask)
    sel, N = bna.nifti_mask(str(tmp_path / "mask.nii"), thr)
    d = tmp_path / "maps"
    d.mkdir()
    v = rng.normal(size=N).astype(np.float32)
    assert bna.lib().bna_write_nm(str(d / "s_deg.nm").encode(), bna._f32p(v), N) == 0
    w = rng.normal(size=N + 3).astype(np.float32)
    assert bna.lib().bna_write_nm(str(d / "t_ec.nm").encode(), bna._f32p(w), N + 3) == 0
    r = _run("bna_convert_nii", d, tmp_path / "mask.nii", thr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "doesn't match" in r.stdout
    assert not (d / "t_ec.nii").exists()
    out = (d / "s_deg.nii").read_bytes()
    mh = (tmp_path / "mask.nii").read_bytes()[:352]
    h = np.frombuffer(out[:352], np.int16)
    m16 = np.frombuffer(mh, np.int16)
    assert h[35] == 16 and h[36] == 32
    keep = np.ones(176, bool)
    keep[[35, 36]] = False
    assert np.array_equal(h[keep], m16[keep])
    vol = np.frombuffer(out[352:], np.float32)
    flat_mask = mask.reshape(-1, order="F").astype(np.float64) >= thr
    assert vol.size == X * Y * Z
    assert np.array_equal(vol[flat_mask], v)
    assert np.all(vol[~flat_mask] == 0)
    back = bna.nifti_bold(str(d / "s_deg.nii"), sel)
    assert np.array_equal(back.reshape(-1), v)
