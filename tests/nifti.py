"""Minimal NIfTI-1 writer for test inputs (single-file .nii, little-endian).

Fields written: sizeof_hdr, dim[0..4], datatype, bitpix, pixdim, vox_offset 352
and the "n+1" magic.  The reader under test takes dims from short offsets
21..24, datatype from 35 and bitpix from 36 of the header viewed as int16[176].
"""
import numpy as np

DT = {np.dtype(np.uint8): (2, 8), np.dtype(np.float32): (16, 32), np.dtype(np.float64): (64, 64)}


def write_nii(path, data: np.ndarray):
    """data has shape (X, Y, Z) or (X, Y, Z, T), stored with x fastest."""
    data = np.asarray(data)
    code, bits = DT[data.dtype]
    hdr = bytearray(352)
    hdr[0:4] = np.int32(348).tobytes()
    dims = np.ones(8, np.int16)
    dims[0] = data.ndim
    dims[1:1 + data.ndim] = data.shape
    hdr[40:56] = dims.tobytes()
    hdr[70:72] = np.int16(code).tobytes()
    hdr[72:74] = np.int16(bits).tobytes()
    hdr[76:108] = np.ones(8, np.float32).tobytes()
    hdr[108:112] = np.float32(352.0).tobytes()
    hdr[344:348] = b"n+1\x00"
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        # volume-major with x fastest: element (x, y, z, t) at x + X*(y + Y*(z + Z*t))
        f.write(np.ascontiguousarray(data.transpose()).tobytes())
