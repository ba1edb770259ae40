"""NIfTI-1 single-file volumes (.nii, .nii.gz) as numpy arrays.

Arrays are in this package's numpy order: a NIfTI volume [nx, ny, nz] (x
fastest on disk) is the C-order array of shape (nz, ny, nx).  Written from the
published NIfTI-1 header layout (348-byte header, 4 extension bytes, data at
vox_offset), like the command line tool's reader (apps/nifti_io.cpp); the
reference reads and writes through nifti/nifti1_io.c.
"""
import gzip
import struct

import numpy as np

_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16,
           768: np.uint32}


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def read(path):
    """-> (float32 array (nz, ny, nx), voxel size (dx, dy, dz)).  Float32 data
    are returned as stored; other scalar types are converted, with the
    header's scl_slope / scl_inter applied when the slope is non-zero."""
    with _open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 348:
        raise ValueError("%s: shorter than a NIfTI-1 header" % path)
    for end in "<>":
        if struct.unpack(end + "i", buf[:4])[0] == 348:
            break
    else:
        raise ValueError("%s: not a NIfTI-1 file" % path)
    if buf[344:348] != b"n+1\0":
        raise ValueError("%s: not a single-file NIfTI-1 volume" % path)
    dim = struct.unpack(end + "8h", buf[40:56])
    if not 1 <= dim[0] <= 7 or any(d > 1 for d in dim[4:dim[0] + 1]):
        raise ValueError("%s: only volumes of up to three dimensions are read" % path)
    n = [max(int(d), 1) if i < dim[0] else 1 for i, d in enumerate(dim[1:4])]
    dt = struct.unpack(end + "h", buf[70:72])[0]
    if dt not in _DTYPES:
        raise ValueError("%s: unsupported datatype %d" % (path, dt))
    pix = struct.unpack(end + "8f", buf[76:108])
    off = int(struct.unpack(end + "f", buf[108:112])[0])
    slope, inter = struct.unpack(end + "2f", buf[112:120])
    count = n[0] * n[1] * n[2]
    a = np.frombuffer(buf, dtype=np.dtype(_DTYPES[dt]).newbyteorder(end), count=count, offset=off)
    if dt != 16 and slope != 0:
        a = a.astype(np.float64) * slope + inter
    return np.ascontiguousarray(a, dtype=np.float32).reshape(n[2], n[1], n[0]), tuple(pix[1:4])


def write(path, volume, voxel=(1.0, 1.0, 1.0)):
    """Writes a (nz, ny, nx) volume as float32 (gzip-compressed for .gz)."""
    v = np.ascontiguousarray(volume, dtype="<f4")
    if v.ndim != 3:
        raise ValueError("volume must be three-dimensional")
    nz, ny, nx = v.shape
    h = bytearray(352)
    struct.pack_into("<i", h, 0, 348)
    h[38:39] = b"r"
    struct.pack_into("<8h", h, 40, 3, nx, ny, nz, 1, 1, 1, 1)
    struct.pack_into("<2h", h, 70, 16, 32)
    struct.pack_into("<8f", h, 76, 1.0, voxel[0], voxel[1], voxel[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<3f", h, 108, 352.0, 1.0, 0.0)
    h[148:148 + 15] = b"niftyrec-mi355x"
    h[344:348] = b"n+1\0"
    with _open(path, "wb") as f:
        f.write(bytes(h))
        f.write(v.tobytes())
